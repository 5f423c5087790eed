// Pressure projection of the LES step (press.f95:47-162, bondFG.f95:18-42,
// boundp.f95:7-78): bondfg + rhs with its volume-weighted mean removed,
// red-black SOR, and removal of the pressure mean followed by the
// boundp1/boundp2 ghost fills.
//
// Arithmetic: the stencils are FP32 in the reference's operation order; the
// three reductions (mean of rhs, sum of reltmp^2, mean of p) accumulate the
// FP32 terms in FP64 with a fixed block partition and summation order, so
// results are run-to-run reproducible.  No kernel waits on another block:
// each reduction writes per-block partials and a later launch adds them.
//
// SOR: one launch per iteration updates the red cells (i+j+k odd, nrd=0 at
// press.f95:92) and then the black cells of the same iteration.  A block
// owns a TX x TY column of cells and a range of k levels and marches k
// upward: at level k it computes the new red values of plane k over its
// tile plus a one-cell halo (from the previous iterate, held in an LDS ring
// with a two-cell halo), then the new black values of plane k-1 from those
// red values.  The previous iterate is read from one buffer and the new one
// written to the other (the two alternate), so blocks never read what
// another block writes.  Ghost reads reproduce boundp1/boundp2 exactly:
// across i=0/i=im+1 and k=0/k=km+1 a cell reads its own pre-update value
// (the ghost is a copy of it: boundp.f95:16-21,44-58 run after the
// colour that last changed it), and across j=0/j=jm+1 it reads the
// periodic image (boundp.f95:63-69): the red sweep sees the previous
// iterate, the black sweep sees the new red value, or -- when jm is odd
// and the image has the same colour -- the previous iterate.
//
// The stop test (press.f95:119-121) runs one iteration behind: block 0 of
// iteration l+1 adds the partials of iteration l and records the result;
// iteration l+2 then finds the flag and exits.  When iteration l converged,
// iteration l+1 has written only the other buffer, so iteration l's result
// is intact and is the one kept.
#include "les_kernels.h"

namespace les {

#define BX 64
#define BY 4
#define NT (BX * BY)

static inline dim3 blk() { return dim3(BX, BY, 1); }

// at most this many blocks write partials in one reduction launch
constexpr int kMaxRedBlocks = 2048;

// k-blocks of a reduction launch: at most `cap` blocks in total unless one
// k-block alone has more (the partial buffer holds max(cap, gx*gy)).
static inline int kblocks(int gx, int gy, int nk, int cap) {
  int gz = cap / (gx * gy);
  if (gz > nk) gz = nk;
  if (gz < 1) gz = 1;
  return gz;
}

// ---------------------------------------------------------------- rhs
// bondfg (bondFG.f95:18-42) folded into the divergence: f(0)=f(1),
// g(i,0)=g(i,jm), h(k=0)=h(km)=0; the ghosts are also stored so the arrays
// end as the reference leaves them.  rhs = div(F) + div(u)/dt
// (press.f95:49-60); block partials of dx1*dy1*dzn*rhs (press.f95:63-72).
__global__ __launch_bounds__(NT) void k_rhs(Grid g, Metrics m, const float* __restrict__ u,
                                            const float* __restrict__ v, const float* __restrict__ w,
                                            float* __restrict__ f, float* __restrict__ gg,
                                            float* __restrict__ h, float* __restrict__ rhs,
                                            double* __restrict__ part, Ctrl* ctrl, float dt) {
  __shared__ double lds[NT / kWave];
  const int tid = threadIdx.y * BX + threadIdx.x;
  const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  if (bid == 0 && tid == 0) {
    ctrl->sor_done = 0;
    ctrl->sor_iters = 0;
    ctrl->nonfinite = 0;
  }
  const int i = 1 + blockIdx.x * BX + threadIdx.x, j = 1 + blockIdx.y * BY + threadIdx.y;
  double acc = 0.0;
  if (i <= g.im && j <= g.jm) {
    const float dxi = M1(m.dx1, i), dyj = M1(m.dy1, j);
    for (int k = 1 + blockIdx.z; k <= g.km; k += gridDim.z) {
      const int c = ix(g, i, j, k);
      const float dzk = M1(m.dzn, k);
      const float fc = f[c], gc = gg[c];
      const float fm = (i == 1) ? fc : f[c - 1];
      const float gm = (j == 1) ? gg[ix(g, i, g.jm, k)] : gg[c - g.sj];
      const float hc = (k == g.km) ? 0.0f : h[c];
      const float hm = (k == 1) ? 0.0f : h[c - g.sk];
      if (i == 1) f[c - 1] = fc;
      if (j == 1) gg[c - g.sj] = gm;
      if (k == 1) h[c - g.sk] = 0.0f;
      if (k == g.km) h[c] = 0.0f;
      const float r = (-u[c - 1] + u[c]) / dxi + (-v[c - g.sj] + v[c]) / dyj + (-w[c - g.sk] + w[c]) / dzk;
      const float rr = (fc - fm) / dxi + (gc - gm) / dyj + (hc - hm) / dzk + r / dt;
      rhs[c] = rr;
      acc += (double)(dxi * dyj * dzk * rr);
    }
  }
  const double s = block_sum<NT>(acc, tid, lds);
  if (tid == 0) part[bid] = s;
}

// One block: total of n partials in a fixed order; what=0 -> rhsav,
// what=1 -> pav (each divided by the volume).
__global__ __launch_bounds__(256) void k_finalize(const double* __restrict__ part, int n, Ctrl* ctrl,
                                                  int what, double denom) {
  __shared__ double lds[4];
  double s = 0.0;
  for (int b = threadIdx.x; b < n; b += 256) s += part[b];
  s = block_sum<256>(s, threadIdx.x, lds);
  if (threadIdx.x == 0) {
    if (what == 0)
      ctrl->rhsav = s / denom;
    else
      ctrl->pav = s / denom;
  }
}

// rhs = rhs - rhsav (press.f95:81-87), the subtraction done in FP64.
__global__ __launch_bounds__(NT) void k_rhs_sub(Grid g, float* __restrict__ rhs, const Ctrl* ctrl) {
  const int i = 1 + blockIdx.x * BX + threadIdx.x, j = 1 + blockIdx.y * BY + threadIdx.y;
  if (i > g.im || j > g.jm) return;
  const double av = ctrl->rhsav;
  for (int k = 1 + blockIdx.z; k <= g.km; k += gridDim.z) {
    const int c = ix(g, i, j, k);
    rhs[c] = (float)((double)rhs[c] - av);
  }
}

// ---------------------------------------------------------------- SOR
constexpr int STX = 64;            // tile width (i)
constexpr int STY = 8;             // tile height (j)
constexpr int SNT = 256;           // threads per block
constexpr int AXW = STX + 4;       // previous-iterate plane with a 2-cell halo
constexpr int AYW = STY + 4;
constexpr int MXW = STX + 2;       // merged plane (new red, old black) with a 1-cell halo
constexpr int MYW = STY + 2;

struct SorBufs {
  const float* p;     // iterate read by iteration 1 (the field p)
  float* s0;          // alternating iterates: iteration l writes s[l&1]
  float* s1;
  double* part0;      // per-block sum(reltmp^2): iteration l writes part[l&1]
  double* part1;
};

__device__ __forceinline__ int wrapj(int j, int jm) {
  int r = (j - 1) % jm;
  if (r < 0) r += jm;
  return r + 1;
}

template <bool SEP>
__device__ __forceinline__ float sor_new(const Grid& g, const Metrics& m, const float* __restrict__ cn1,
                                         const float* __restrict__ rhs, int i, int j, int k, float own,
                                         float pim, float pip, float pjm, float pjp, float pkm, float pkp,
                                         float omega, float* reltmp_out) {
  const int c = ix(g, i, j, k);
  float cn;
  if (SEP)
    cn = 1.0f / ((M1(m.cnax, i) + M1(m.cnby, j)) + M1(m.cncz, k));  // init.f95:100-107
  else
    cn = cn1[c];
  const float reltmp = omega * (cn * (M1(m.cn2l, i) * pip + M1(m.cn2s, i) * pim + M1(m.cn3l, j) * pjp +
                                      M1(m.cn3s, j) * pjm + M1(m.cn4l, k) * pkp + M1(m.cn4s, k) * pkm -
                                      rhs[c]) -
                                own);
  *reltmp_out = reltmp;
  return own + reltmp;
}

template <bool SEP>
__global__ __launch_bounds__(SNT) void k_sor_iter(Grid g, Metrics m, SorBufs sb,
                                                  const float* __restrict__ cn1,
                                                  const float* __restrict__ rhs, Ctrl* ctrl, int iter,
                                                  int kchunk, int nparts, float omega, float pjuge) {
  __shared__ float As[4][AYW][AXW];
  __shared__ float Ms[3][MYW][MXW];
  __shared__ double lds[SNT / kWave];
  __shared__ int skip;
  const int tid = threadIdx.x;
  // Block 0 of this launch may set the flag while other blocks start; read
  // it once per block so the whole block takes one decision.  A block that
  // sees the flag set by this launch's block 0 leaves its part of an
  // iterate that is discarded anyway.
  if (tid == 0) skip = (ctrl->sor_done != 0);
  __syncthreads();
  if (skip) return;
  const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  if (iter >= 2 && bid == 0) {
    // stop test of the previous iteration (press.f95:119-121)
    const double* pp = ((iter - 1) & 1) ? sb.part1 : sb.part0;
    double s = 0.0;
    for (int b = tid; b < nparts; b += SNT) s += pp[b];
    s = block_sum<SNT>(s, tid, lds);
    if (tid == 0) {
      if (iter - 1 <= kSorHist) ctrl->sor_hist[iter - 2] = s;
      ctrl->sor_iters = iter - 1;
      if (!(s == s) || s > 1.0e300) ctrl->nonfinite = 1;
      if (s < (double)pjuge || ctrl->nonfinite) ctrl->sor_done = iter - 1;
    }
  }
  const float* __restrict__ A = (iter == 1) ? sb.p : (((iter - 1) & 1) ? sb.s1 : sb.s0);
  float* __restrict__ B = (iter & 1) ? sb.s1 : sb.s0;
  double* part = (iter & 1) ? sb.part1 : sb.part0;
  const int im = g.im, jm = g.jm, km = g.km;
  const int i0 = 1 + blockIdx.x * STX, j0 = 1 + blockIdx.y * STY;
  const int k0 = 1 + blockIdx.z * kchunk;
  const int k1 = min(km, k0 + kchunk - 1);

  auto loadA = [&](int k) {
    if (k < 1 || k > km) return;
    float(*P)[AXW] = As[k & 3];
    for (int e = tid; e < AXW * AYW; e += SNT) {
      const int r = e / AXW, q = e - r * AXW;
      const int i = i0 - 2 + q;
      const int j = wrapj(j0 - 2 + r, jm);
      P[r][q] = (i >= 1 && i <= im) ? A[ix(g, i, j, k)] : 0.0f;
    }
  };

  double acc = 0.0;
  loadA(k0 - 2);
  loadA(k0 - 1);
  for (int k = k0 - 1; k <= k1 + 1; k++) {
    loadA(k + 1);
    __syncthreads();
    // ---- red cells of plane k over the tile and a one-cell halo
    if (k >= 1 && k <= km) {
      const float(*P)[AXW] = As[k & 3];
      const float(*Pd)[AXW] = As[(k - 1) & 3];
      const float(*Pu)[AXW] = As[(k + 1) & 3];
      float(*M)[MXW] = Ms[k % 3];
      for (int e = tid; e < MXW * MYW; e += SNT) {
        const int r = e / MXW, q = e - r * MXW;
        const int i = i0 - 1 + q;
        const int jj = j0 - 1 + r;
        const int j = wrapj(jj, jm);
        const float own = P[r + 1][q + 1];
        float val = own;
        if (i >= 1 && i <= im && ((i + j + k) & 1)) {
          const float pim = (i == 1) ? own : P[r + 1][q];
          const float pip = (i == im) ? own : P[r + 1][q + 2];
          const float pkm = (k == 1) ? own : Pd[r + 1][q + 1];
          const float pkp = (k == km) ? own : Pu[r + 1][q + 1];
          float rel;
          val = sor_new<SEP>(g, m, cn1, rhs, i, j, k, own, pim, pip, P[r][q + 1], P[r + 2][q + 1], pkm, pkp,
                             omega, &rel);
          if (r >= 1 && r <= STY && q >= 1 && q <= STX && jj <= jm && k >= k0 && k <= k1)
            acc += (double)rel * (double)rel;
        }
        M[r][q] = val;
      }
    }
    __syncthreads();
    // ---- black cells of plane kb = k-1 of the tile; store the plane
    const int kb = k - 1;
    if (kb >= k0 && kb <= k1) {
      const float(*P)[AXW] = As[kb & 3];
      const float(*M)[MXW] = Ms[kb % 3];
      const float(*Md)[MXW] = Ms[(kb + 2) % 3];  // plane kb-1
      const float(*Mu)[MXW] = Ms[(kb + 1) % 3];  // plane kb+1
      for (int e = tid; e < STX * STY; e += SNT) {
        const int r = e / STX, q = e - r * STX;
        const int i = i0 + q, j = j0 + r;
        if (i > im || j > jm) continue;
        float val;
        if ((i + j + kb) & 1) {
          val = M[r + 1][q + 1];  // red, already updated
        } else {
          const float own = P[r + 2][q + 2];
          const float pim = (i == 1) ? own : M[r + 1][q];
          const float pip = (i == im) ? own : M[r + 1][q + 2];
          const float pkm = (kb == 1) ? own : Md[r + 1][q + 1];
          const float pkp = (kb == km) ? own : Mu[r + 1][q + 1];
          float rel;
          val = sor_new<SEP>(g, m, cn1, rhs, i, j, kb, own, pim, pip, M[r][q + 1], M[r + 2][q + 1], pkm, pkp,
                             omega, &rel);
          acc += (double)rel * (double)rel;
        }
        B[ix(g, i, j, kb)] = val;
      }
    }
  }
  const double s = block_sum<SNT>(acc, tid, lds);
  if (tid == 0) part[bid] = s;
}

// After the iteration loop: when no iteration met the stop test, the last
// one (nmaxp) is summed here.  Records the iteration count and sum.
__global__ __launch_bounds__(256) void k_sor_final(SorBufs sb, Ctrl* ctrl, int nmaxp, int nparts,
                                                   float pjuge) {
  __shared__ double lds[4];
  if (ctrl->sor_done != 0) return;  // uniform across the block
  const double* pp = (nmaxp & 1) ? sb.part1 : sb.part0;
  double s = 0.0;
  for (int b = threadIdx.x; b < nparts; b += 256) s += pp[b];
  s = block_sum<256>(s, threadIdx.x, lds);
  if (threadIdx.x == 0) {
    if (nmaxp <= kSorHist) ctrl->sor_hist[nmaxp - 1] = s;
    ctrl->sor_iters = nmaxp;
    if (!(s == s) || s > 1.0e300) ctrl->nonfinite = 1;
  }
}

__device__ __forceinline__ const float* sor_result(const SorBufs& sb, const Ctrl* ctrl) {
  const int L = ctrl->sor_iters;
  return (L == 0) ? sb.p : ((L & 1) ? sb.s1 : sb.s0);
}

// pav partials: p*dx1*dy1*dzn over the SOR result (press.f95:124-133).
__global__ __launch_bounds__(NT) void k_pav(Grid g, Metrics m, SorBufs sb, double* __restrict__ part,
                                            const Ctrl* ctrl) {
  __shared__ double lds[NT / kWave];
  const int tid = threadIdx.y * BX + threadIdx.x;
  const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  const float* __restrict__ p = sor_result(sb, ctrl);
  const int i = 1 + blockIdx.x * BX + threadIdx.x, j = 1 + blockIdx.y * BY + threadIdx.y;
  double acc = 0.0;
  if (i <= g.im && j <= g.jm) {
    const float dxi = M1(m.dx1, i), dyj = M1(m.dy1, j);
    for (int k = 1 + blockIdx.z; k <= g.km; k += gridDim.z)
      acc += (double)(p[ix(g, i, j, k)] * dxi * dyj * M1(m.dzn, k));
  }
  const double s = block_sum<NT>(acc, tid, lds);
  if (tid == 0) part[bid] = s;
}

// p = p - pav on the interior (press.f95:145-150, FP64 subtraction) and the
// ghost cells of boundp1 then boundp2 (press.f95:155-159): every ghost
// receives the adjusted value of the interior cell the copies trace it to
// (i: 0->1, im+1->im; j: 0->jm, jm+1->1; k: 0->1, km+1->km).  Reads the SOR
// result, writes p; covers i,j,k = 0..im+1, 0..jm+1, 0..km+1.
__global__ __launch_bounds__(NT) void k_padj(Grid g, SorBufs sb, float* __restrict__ p, const Ctrl* ctrl) {
  const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
  if (i > g.im + 1 || j > g.jm + 1) return;
  const float* __restrict__ src = sor_result(sb, ctrl);
  const double av = ctrl->pav;
  const int ii = (i == 0) ? 1 : ((i == g.im + 1) ? g.im : i);
  const int jj = (j == 0) ? g.jm : ((j == g.jm + 1) ? 1 : j);
  for (int k = blockIdx.z; k <= g.km + 1; k += gridDim.z) {
    const int kk = (k == 0) ? 1 : ((k == g.km + 1) ? g.km : k);
    p[ix(g, i, j, k)] = (float)((double)src[ix(g, ii, jj, kk)] - av);
  }
}

void launch_press(const DevDomain& d, float dt) {
  const Grid& g = d.g;
  const int gx = (g.im + BX - 1) / BX, gy = (g.jm + BY - 1) / BY;
  const int gz = kblocks(gx, gy, g.km, kMaxRedBlocks);
  const int nred = gx * gy * gz;
  TIMED(d, KC_RHS, hipLaunchKernelGGL(k_rhs, dim3(gx, gy, gz), blk(), 0, d.stream, g, d.m, d.F.u, d.F.v,
                                      d.F.w, d.F.f, d.F.g, d.F.h, d.F.rhs, d.partials, d.ctrl, dt));
  TIMED(d, KC_RHS, hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, d.stream, d.partials, nred, d.ctrl, 0,
                                      d.area));
  TIMED(d, KC_RHS, hipLaunchKernelGGL(k_rhs_sub, dim3(gx, gy, g.km), blk(), 0, d.stream, g, d.F.rhs,
                                      (const Ctrl*)d.ctrl));
  SorBufs sb;
  sb.p = d.F.p;
  sb.s0 = d.F.ps[0];
  sb.s1 = d.F.ps[1];
  sb.part0 = d.partials + d.npartials;
  sb.part1 = d.partials + d.npartials + d.sor_parts_cap;
  const int tx = (g.im + STX - 1) / STX, ty = (g.jm + STY - 1) / STY;
  const int nkc = (g.km + d.sor_kchunk - 1) / d.sor_kchunk;
  const int nparts = tx * ty * nkc;
  for (int l = 1; l <= d.nmaxp; l++) {
    if (d.cn1_separable)
      TIMED(d, KC_SOR, hipLaunchKernelGGL(k_sor_iter<true>, dim3(tx, ty, nkc), dim3(SNT), 0, d.stream, g, d.m, sb,
                                          (const float*)d.F.cn1, (const float*)d.F.rhs, d.ctrl, l, d.sor_kchunk,
                                          nparts, d.omega, d.pjuge));
    else
      TIMED(d, KC_SOR, hipLaunchKernelGGL(k_sor_iter<false>, dim3(tx, ty, nkc), dim3(SNT), 0, d.stream, g, d.m, sb,
                                          (const float*)d.F.cn1, (const float*)d.F.rhs, d.ctrl, l, d.sor_kchunk,
                                          nparts, d.omega, d.pjuge));
  }
  TIMED(d, KC_PAV, hipLaunchKernelGGL(k_sor_final, dim3(1), dim3(256), 0, d.stream, sb, d.ctrl, d.nmaxp, nparts,
                                      d.pjuge));
  TIMED(d, KC_PAV, hipLaunchKernelGGL(k_pav, dim3(gx, gy, gz), blk(), 0, d.stream, g, d.m, sb, d.partials,
                                      (const Ctrl*)d.ctrl));
  TIMED(d, KC_PAV, hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, d.stream, d.partials, nred, d.ctrl, 1,
                                      d.area));
  TIMED(d, KC_PADJ, hipLaunchKernelGGL(k_padj, dim3((g.im + 2 + BX - 1) / BX, (g.jm + 2 + BY - 1) / BY, g.km + 2),
                                       blk(), 0, d.stream, g, sb, d.F.p, (const Ctrl*)d.ctrl));
}

int sor_parts_needed(const Grid& g, int kchunk) {
  const int tx = (g.im + STX - 1) / STX, ty = (g.jm + STY - 1) / STY;
  return tx * ty * ((g.km + kchunk - 1) / kchunk);
}

int reduction_parts_needed() { return kMaxRedBlocks; }

}  // namespace les
