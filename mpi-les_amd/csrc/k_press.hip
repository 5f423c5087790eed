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
// SOR storage: the red cells (i+j+k odd, updated first: nrd=0 at
// press.f95:92) and the black cells live in two compact arrays, pair t of a
// (j,k) row holding i = 2t+1 and i = 2t+2, so a colour sweep streams
// contiguous memory: the same-row neighbours i-1, i+1 of a cell at t are at
// t-1+d and t+d of the other colour (d = 0 when the cell is the odd-i one),
// its j and k neighbours at t.  Each colour has two buffers: iteration l
// reads the buffers of parity l-1 and writes those of parity l, so a sweep
// never reads what it writes.  Ghost reads reproduce boundp1/boundp2
// exactly: across i=0/im+1 and k=0/km+1 a cell reads its own pre-update
// value (the ghost copy, boundp.f95:16-21,44-58, was taken after the last
// change of that cell), across j=0/jm+1 the periodic image (boundp.f95:
// 63-69) -- of the other colour when jm is even (the previous iterate for
// the red sweep, the new red value for the black sweep), of the same colour
// when jm is odd (always the previous iterate).
//
// The stop test (press.f95:119-121) runs one iteration behind: block 0 of
// iteration l+1's red sweep adds the partials of iteration l and records
// the result; later sweeps find the flag and exit.  Iteration l's values
// stay intact because iteration l+1 writes the other buffers.
#include <algorithm>
#include <type_traits>

#include "les_kernels.h"

namespace les {

#define BX 64
#define BY 4
#define NT (BX * BY)

static inline dim3 blk() { return dim3(BX, BY, 1); }

// at most this many blocks write partials in one reduction launch
constexpr int kMaxRedBlocks = 2048;

// k-blocks of a launch: at most `cap` blocks in total unless one k-block
// alone has more (partial buffers hold max(cap, gx*gy)).
static inline int kblocks(int gx, int gy, int nk, int cap) {
  int gz = cap / (gx * gy);
  if (gz > nk) gz = nk;
  if (gz < 1) gz = 1;
  return gz;
}

__device__ __forceinline__ int cix(const SplitGrid& s, int t, int j, int k) {
  return t + (j - 1) * s.cp + (k - 1) * s.cplane;
}

// ---------------------------------------------------------------- rhs
// bondfg (bondFG.f95:18-42) folded into the divergence: f(0)=f(1),
// g(i,0)=g(i,jm), h(k=0)=h(km)=0; the ghosts are also stored so the arrays
// end as the reference leaves them.  rhs = div(F) + div(u)/dt
// (press.f95:49-60); block partials of dx1*dy1*dzn*rhs (press.f95:63-72).
__device__ __forceinline__ void rhs_body(const VBlk& vb, const Grid& g, const Metrics& m,
                                         const float* __restrict__ u, const float* __restrict__ v,
                                         const float* __restrict__ w, float* __restrict__ f, float* __restrict__ gg,
                                         float* __restrict__ h, float* __restrict__ rhs, double* __restrict__ part,
                                         Ctrl* ctrl, float dt) {
  __shared__ double lds[NT / kWave];
  const int tid = threadIdx.y * BX + threadIdx.x;
  const int bid = vlinear(vb);
  if (bid == 0 && tid == 0) {
    ctrl->sor_done = 0;
    ctrl->sor_iters = 0;
    ctrl->nonfinite = 0;
    ctrl->p_ghost_diff = 0;
    ctrl->p_final = -1;
    ctrl->fix_iter = 0;
  }
  const int i = 1 + vb.x * BX + threadIdx.x, j = 1 + vb.y * BY + wave_row();
  double acc = 0.0;
  if (i <= g.im && j <= g.jm) {
    const float dxi = M1(m.dx1, i), dyj = M1(m.dy1, j);
    int k0, k1;
    kspan(vb, 1, g.km, k0, k1);
    for (int k = k0; k <= k1; k++) {
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
__global__ __launch_bounds__(NT) void k_rhs(Grid g, Metrics m, const float* __restrict__ u,
                                            const float* __restrict__ v, const float* __restrict__ w,
                                            float* __restrict__ f, float* __restrict__ gg,
                                            float* __restrict__ h, float* __restrict__ rhs,
                                            double* __restrict__ part, Ctrl* ctrl, float dt) {
  rhs_body(hw_blk<true>(), g, m, u, v, w, f, gg, h, rhs, part, ctrl, dt);
}

// One block: total of n partials in a fixed order; what=0 -> rhsav,
// what=1 -> pav (each divided by the volume).
__device__ __forceinline__ void finalize_body(const double* __restrict__ part, int n, Ctrl* ctrl, int what,
                                              double denom) {
  __shared__ double lds[4];
  double s = 0.0;
  for (int b = threadIdx.x; b < n; b += 256) s += part[b];
  s = block_sum<256>(s, threadIdx.x, lds);
  if (threadIdx.x == 0) {
    if (what == 0)
      ctrl->rhsav = s / denom;
    else if (what == 1)
      ctrl->pav = s / denom;
    else
      ctrl->rhsav64 = s / denom;
  }
}
__global__ __launch_bounds__(256) void k_finalize(const double* __restrict__ part, int n, Ctrl* ctrl,
                                                  int what, double denom) {
  finalize_body(part, n, ctrl, what, denom);
}

// rhs = rhs - rhsav (press.f95:81-87, the subtraction in FP64), and the
// colour-split copies the SOR reads: rhs, the starting iterate p, and cn1
// when it is not recomputed.  Also compares, bit for bit, every ghost of p
// next to the interior with the value boundp gives it (i and k: the
// neighbouring interior cell, boundp.f95:7-78; j: the periodic image) and
// raises ctrl->p_ghost_diff on any difference.  Without one, SOR iteration 1
// reading the entry ghosts equals an iteration that mirrors the current
// values, so it can run as any later iteration.
__device__ __forceinline__ void rhs_split_body(const VBlk& vb, const Grid& g, const SplitGrid& s,
                                               float* __restrict__ rhs, const float* __restrict__ p,
                                               const float* __restrict__ cn1, const SplitBufs& sb,
                                               Ctrl* ctrl) {
  const int i = 1 + vb.x * BX + threadIdx.x, j = 1 + vb.y * BY + wave_row();
  if (i > g.im || j > g.jm) return;
  const double av = ctrl->rhsav;
  const int t = (i - 1) >> 1;
  int k0, k1;
  kspan(vb, 1, g.km, k0, k1);
  bool diff = false;
  const auto same = [](float a, float b) { return __float_as_uint(a) == __float_as_uint(b); };
  for (int k = k0; k <= k1; k++) {
    const int c = ix(g, i, j, k);
    const int col = (i + j + k) & 1;  // 1 = red
    const float r = (float)((double)rhs[c] - av);
    rhs[c] = r;
    const int q = cix(s, t, j, k);
    sb.rhs[col][q] = r;
    const float pc = p[c];
    sb.p[0][col][q] = pc;
    if (sb.cn1[0]) sb.cn1[col][q] = cn1[c];
    if (i == 1) diff |= !same(p[c - 1], pc);
    if (i == g.im) diff |= !same(p[c + 1], pc);
    if (j == 1) diff |= !same(p[ix(g, i, 0, k)], p[ix(g, i, g.jm, k)]);
    if (j == g.jm) diff |= !same(p[ix(g, i, g.jm + 1, k)], p[ix(g, i, 1, k)]);
    if (k == 1) diff |= !same(p[c - g.sk], pc);
    if (k == g.km) diff |= !same(p[c + g.sk], pc);
  }
  if (diff) ctrl->p_ghost_diff = 1;
}
__global__ __launch_bounds__(NT) void k_rhs_split(Grid g, SplitGrid s, float* __restrict__ rhs,
                                                  const float* __restrict__ p, const float* __restrict__ cn1,
                                                  SplitBufs sb, Ctrl* ctrl) {
  rhs_split_body(hw_blk<false>(), g, s, rhs, p, cn1, sb, ctrl);
}

// The head of press with two SOR iterations per launch, in one pass: the
// rhs of k_rhs (the same expression, grid, block partials and ghost stores
// of f, g, h), stored into the colour-split rhsr before its mean is removed
// (k_rhs_sub then writes the split rhs), the split copy of p (iterate set
// 0), and the ghost comparison of k_rhs_split.  Each block's comparison result goes to part[nred + block]
// for k_finalize_head, so that no block clears a flag another one may have
// raised.  The reference-layout rhs is not written here (a download gathers
// it from the split rhs).
__global__ __launch_bounds__(NT) void k_rhs_head(Grid g, SplitGrid s, Metrics m, const float* __restrict__ u,
                                                 const float* __restrict__ v, const float* __restrict__ w,
                                                 float* __restrict__ f, float* __restrict__ gg,
                                                 float* __restrict__ h, const float* __restrict__ p, SplitBufs sb,
                                                 double* __restrict__ part, int nred, Ctrl* ctrl, float dt) {
  __shared__ double lds[NT / kWave];
  const VBlk vb = hw_blk<true>();
  const int tid = threadIdx.y * BX + threadIdx.x;
  const int bid = vlinear(vb);
  if (bid == 0 && tid == 0) {
    ctrl->sor_done = 0;
    ctrl->sor_iters = 0;
    ctrl->nonfinite = 0;
    ctrl->p_final = -1;
    ctrl->fix_iter = 0;
  }
  const int i = 1 + vb.x * BX + threadIdx.x, j = 1 + vb.y * BY + wave_row();
  double acc = 0.0;
  bool diff = false;
  if (i <= g.im && j <= g.jm) {
    const float dxi = M1(m.dx1, i), dyj = M1(m.dy1, j);
    const int t = (i - 1) >> 1;
    const auto same = [](float a, float b) { return __float_as_uint(a) == __float_as_uint(b); };
    const bool lane0 = threadIdx.x == 0;
    int k0, k1;
    kspan(vb, 1, g.km, k0, k1);
    // w and h of level k-1 come from the thread's previous level, u and f
    // of cell i-1 from the neighbouring lane: the same values as the loads
    // of k_rhs.  The operands of level k+1 are loaded before level k is
    // computed; nothing the loop writes (ghosts of f, g and h) is read by a
    // later load.
    struct Lv {
      float f, g, u, w, h, v, vm, gm, p, fl, ul;
    };
    const auto load = [&](int k) {
      const int c = ix(g, i, j, k);
      Lv a;
      a.f = f[c];
      a.g = gg[c];
      a.u = u[c];
      a.w = w[c];
      a.h = h[c];
      a.v = v[c];
      a.vm = v[c - g.sj];
      a.gm = (j == 1) ? gg[ix(g, i, g.jm, k)] : gg[c - g.sj];
      a.p = p[c];
      a.fl = lane0 ? f[c - 1] : 0.0f;  // the block's first lane: cell i-1 belongs to the previous block
      a.ul = lane0 ? u[c - 1] : 0.0f;
      return a;
    };
    float wkm = 0.0f, hkm = 0.0f;
    Lv cur{};
    if (k0 <= k1) {
      wkm = w[ix(g, i, j, k0 - 1)];
      hkm = h[ix(g, i, j, k0 - 1)];
      cur = load(k0);
    }
    for (int k = k0; k <= k1; k++) {
      const Lv nxt = (k < k1) ? load(k + 1) : cur;
      const int c = ix(g, i, j, k);
      const float dzk = M1(m.dzn, k);
      const float fc = cur.f, gc = cur.g, uc = cur.u, wc = cur.w, hk = cur.h;
      float fl = __shfl_up(fc, 1, kWave), ul = __shfl_up(uc, 1, kWave);
      if (lane0) {
        fl = cur.fl;
        ul = cur.ul;
      }
      const float fm = (i == 1) ? fc : fl;
      const float gm = cur.gm;
      const float hc = (k == g.km) ? 0.0f : hk;
      const float hm = (k == 1) ? 0.0f : hkm;
      if (i == 1) f[c - 1] = fc;
      if (j == 1) gg[c - g.sj] = gm;
      if (k == 1) h[c - g.sk] = 0.0f;
      if (k == g.km) h[c] = 0.0f;
      const float r = (-ul + uc) / dxi + (-cur.vm + cur.v) / dyj + (-wkm + wc) / dzk;
      const float rr = (fc - fm) / dxi + (gc - gm) / dyj + (hc - hm) / dzk + r / dt;
      wkm = wc;
      hkm = hk;
      acc += (double)(dxi * dyj * dzk * rr);
      const int col = (i + j + k) & 1;  // 1 = red
      const int q = cix(s, t, j, k);
      sb.rhsr[col][q] = rr;
      const float pc = cur.p;
      sb.p[0][col][q] = pc;
      cur = nxt;
      if (i == 1) diff |= !same(p[c - 1], pc);
      if (i == g.im) diff |= !same(p[c + 1], pc);
      if (j == 1) diff |= !same(p[ix(g, i, 0, k)], p[ix(g, i, g.jm, k)]);
      if (j == g.jm) diff |= !same(p[ix(g, i, g.jm + 1, k)], p[ix(g, i, 1, k)]);
      if (k == 1) diff |= !same(p[c - g.sk], pc);
      if (k == g.km) diff |= !same(p[c + g.sk], pc);
    }
  }
  const double sum = block_sum<NT>(acc, tid, lds);
  const int anyd = __syncthreads_or(diff ? 1 : 0);
  if (tid == 0) {
    part[bid] = sum;
    part[nred + bid] = anyd ? 1.0 : 0.0;
  }
}

// rhsav from the partials of k_rhs_head (as k_finalize: the same order and
// sum), and ctrl->p_ghost_diff from its per-block flags.
__global__ __launch_bounds__(256) void k_finalize_head(const double* __restrict__ part, int n, Ctrl* ctrl,
                                                       double denom) {
  __shared__ double lds[4];
  double s = 0.0, fl = 0.0;
  for (int b = threadIdx.x; b < n; b += 256) {
    s += part[b];
    fl += part[n + b];
  }
  s = block_sum<256>(s, threadIdx.x, lds);
  fl = block_sum<256>(fl, threadIdx.x, lds);
  if (threadIdx.x == 0) {
    ctrl->rhsav = s / denom;
    ctrl->p_ghost_diff = fl != 0.0 ? 1 : 0;
  }
}

// rhs = rhsr - rhsav over both colour-split arrays (press.f95:81-87, the
// subtraction in FP64 as in k_rhs_split), which lie one after the other:
// n4 groups of four floats (the row padding included, which nothing reads).
constexpr int kRhsSubBlocks = 2048;
__global__ __launch_bounds__(256) void k_rhs_sub(const float4* __restrict__ in, float4* __restrict__ out, long n4,
                                                 const Ctrl* ctrl) {
  const double av = ctrl->rhsav;
  for (long q = blockIdx.x * 256L + threadIdx.x; q < n4; q += (long)gridDim.x * 256L) {
    const float4 x = in[q];
    float4 y;
    y.x = (float)((double)x.x - av);
    y.y = (float)((double)x.y - av);
    y.z = (float)((double)x.z - av);
    y.w = (float)((double)x.w - av);
    out[q] = y;
  }
}

// rhs(i,j,k) = the split rhs, on the interior (press.f95:81-87 leave the
// ghosts of rhs untouched).
__global__ __launch_bounds__(NT) void k_rhs_gather(Grid g, SplitGrid s, SplitBufs sb, float* __restrict__ rhs) {
  const VBlk vb = hw_blk<false>();
  const int i = 1 + vb.x * BX + threadIdx.x, j = 1 + vb.y * BY + wave_row();
  if (i > g.im || j > g.jm) return;
  const int t = (i - 1) >> 1;
  int k0, k1;
  kspan(vb, 1, g.km, k0, k1);
  for (int k = k0; k <= k1; k++) rhs[ix(g, i, j, k)] = sb.rhs[(i + j + k) & 1][cix(s, t, j, k)];
}

// ---------------------------------------------------------------- SOR
// One colour sweep of iteration `iter` (press.f95:92-104).  COL 1 = red
// (the first sweep), 0 = black.  Red sweeps of iteration >= 2 run the
// lagged stop test in block 0.
// Workgroups are dispatched to the 8 XCDs round-robin (workgroup b runs on
// XCD b % 8).  Renumber them so each XCD receives a contiguous range of
// tiles: neighbouring j rows and k planes then share that XCD's L2.
__device__ __forceinline__ int xcd_contiguous(int b, int nb) {
  constexpr int kXcd = 8;
  const int x = b % kXcd, l = b / kXcd;
  const int base = nb / kXcd, rem = nb % kXcd;
  return x * base + (x < rem ? x : rem) + l;
}

template <int COL, bool SEP>
__global__ __launch_bounds__(NT) void k_sweep(Grid g, SplitGrid s, Metrics m, SplitBufs sb, Ctrl* ctrl,
                                              int iter, int nparts, float omega, float pjuge, int gx,
                                              int gy, int kchunk, int remap,
                                              const float* __restrict__ pent) {
  __shared__ double lds[NT / kWave];
  __shared__ int skip;
  const int tid = threadIdx.y * BX + threadIdx.x;
  // Block 0 of a red sweep may set the flag while other blocks start; each
  // block reads it once.  Values a skipped block leaves unwritten belong to
  // an iterate that is discarded.
  if (tid == 0) skip = (ctrl->sor_done != 0);
  __syncthreads();
  if (skip) return;
  const int bid = blockIdx.x;
  if (COL == 1 && iter >= 2 && bid == 0) {
    const int pp = (iter - 1) & 1;
    double sum = 0.0;
    for (int b = tid; b < nparts; b += NT) sum += sb.part[pp][1][b] + sb.part[pp][0][b];
    sum = block_sum<NT>(sum, tid, lds);
    if (tid == 0) {
      if (iter - 1 <= kSorHist) ctrl->sor_hist[iter - 2] = sum;
      ctrl->sor_iters = iter - 1;
      if (!(sum == sum) || sum > 1.0e300) ctrl->nonfinite = 1;
      if (sum < (double)pjuge || ctrl->nonfinite) ctrl->sor_done = iter - 1;  // press.f95:119-121
    }
  }
  const int prv = (iter - 1) & 1, cur = iter & 1;
  const float* __restrict__ own_in = sb.p[prv][COL];
  float* __restrict__ out = sb.p[cur][COL];
  // other colour: the previous iterate for the red sweep, the new red values for the black sweep
  const float* __restrict__ oth = (COL == 1) ? sb.p[prv][0] : sb.p[cur][1];
  const float* __restrict__ rhsc = sb.rhs[COL];
  const float* __restrict__ cnc = sb.cn1[COL];
  const int im = g.im, jm = g.jm, km = g.km;
  const int lb = remap ? xcd_contiguous(bid, nparts) : bid;
  const int bx = lb % gx, by = (lb / gx) % gy, bz = lb / (gx * gy);
  const int t = bx * BX + threadIdx.x, j = 1 + by * BY + threadIdx.y;
  const int k0 = 1 + bz * kchunk, k1 = min(km, k0 + kchunk - 1);
  double acc = 0.0;
  if (j <= jm && t < s.nt && k0 <= km) {
    const bool jodd = (jm & 1) != 0;
    const int jlo = (j == 1) ? jm : j - 1, jhi = (j == jm) ? 1 : j + 1;
    // j neighbours across the periodic boundary have this colour when jm is odd
    const float* __restrict__ src_jlo = (j == 1 && jodd) ? own_in : oth;
    const float* __restrict__ src_jhi = (j == jm && jodd) ? own_in : oth;
    const int djlo = (jlo - j) * s.cp, djhi = (jhi - j) * s.cp;
    const float c3l = M1(m.cn3l, j), c3s = M1(m.cn3s, j);
    const float byj = SEP ? M1(m.cnby, j) : 0.0f;
    // Per level k the cell is (t, j, k) with i = 2t+1+d.  Its other-colour
    // neighbours: a = same t (i-1 when d=1, i+1 when d=0), b = t+1 or t-1
    // (the other i side), and, at levels k-1 and k+1, the a of those levels.
    // Loads for level k+1 are issued before level k is computed.
    struct Lv {
      float own, a, b, jl, jh, r, cn;
    };
    auto load = [&](int k, Lv& L) {
      const int d = (COL + 1 + j + k) & 1;
      const int q = cix(s, t, j, k);
      int tb = t + (d ? 1 : -1);
      tb = (tb < 0 || tb >= s.nt) ? t : tb;  // outside the row only where the boundary rule applies
      L.own = own_in[q];
      L.a = oth[q];
      L.b = oth[q - t + tb];
      L.jl = src_jlo[q + djlo];
      L.jh = src_jhi[q + djhi];
      L.r = rhsc[q];
      L.cn = SEP ? 0.0f : cnc[q];
    };
    Lv cur, nxt;
    float a_prev = (k0 > 1) ? oth[cix(s, t, j, k0 - 1)] : 0.0f;
    load(k0, cur);
    for (int k = k0; k <= k1; k++) {
      float a_next = 0.0f;
      if (k < k1)
        load(k + 1, nxt);
      else if (k < km)
        a_next = oth[cix(s, t, j, k + 1)];
      if (k < k1) a_next = nxt.a;
      const int d = (COL + 1 + j + k) & 1;
      const int i = 2 * t + 1 + d;
      if (i <= im) {
        const float own = cur.own;
        const float lo = d ? cur.a : cur.b, hi = d ? cur.b : cur.a;  // values at i-1, i+1
        float pim = (i == 1) ? own : lo;
        float pip = (i == im) ? own : hi;
        float pkm = (k == 1) ? own : a_prev;
        float pkp = (k == km) ? own : a_next;
        float pjm = cur.jl, pjp = cur.jh;
        if (pent) {  // iteration 1: the ghost values p had on entry to press
          if (COL == 1) {
            if (i == 1) pim = pent[ix(g, 0, j, k)];
            if (i == im) pip = pent[ix(g, im + 1, j, k)];
            if (j == 1) pjm = pent[ix(g, i, 0, k)];
            if (j == jm) pjp = pent[ix(g, i, jm + 1, k)];
          }
          if (k == 1) pkm = pent[ix(g, i, j, 0)];
          if (k == km) pkp = pent[ix(g, i, j, km + 1)];
        }
        const float cn = SEP ? rcp_rn((M1(m.cnax, i) + byj) + M1(m.cncz, k)) : cur.cn;  // init.f95:100-107
        const float reltmp = omega * (cn * (M1(m.cn2l, i) * pip + M1(m.cn2s, i) * pim + c3l * pjp +
                                            c3s * pjm + M1(m.cn4l, k) * pkp + M1(m.cn4s, k) * pkm -
                                            cur.r) -
                                      own);
        out[cix(s, t, j, k)] = own + reltmp;
        acc += (double)reltmp * (double)reltmp;
      }
      a_prev = cur.a;
      cur = nxt;
    }
  }
  const double sum = block_sum<NT>(acc, tid, lds);
  if (tid == 0) sb.part[cur][COL][bid] = sum;
}

// ------------------------------------------------- fused SOR iteration
// Both colour sweeps of one SOR iteration in one launch, with the same
// per-cell arithmetic (and so bit-identical iterates) as k_sweep<1> then
// k_sweep<0>.  A block owns a tile of FTX pairs x FTY rows and a chunk of
// k levels, and marches k upward with one thread per (pair, row) of the
// tile plus a two-cell ring.  At each step kr it
//   1. puts the previous black iterate of level kr in LDS (ring included),
//   2. computes the new red values of level kr on the tile and a one-cell
//      ring (the ring values are recomputed by the neighbouring block too,
//      with identical results) and puts them in LDS,
//   3. computes the new black values of level kr-1 from the red values of
//      levels kr-2..kr (in-plane neighbours from LDS, the vertical ones
//      from registers).
// LDS planes alternate between two buffers, so one barrier per step
// suffices.  HBM traffic per iteration is one read of each colour's
// iterate, one read of rhs and one write of each colour (12 B per cell);
// the ring re-reads mostly hit L2.  The periodic j ring is indexed modulo
// jm; when jm is odd the j-neighbours across the periodic boundary have
// the cell's own colour and keep their previous iterate (as in k_sweep).
constexpr int FX = 64, FY = 16, FTX = FX - 4, FTY = FY - 4, FNT = FX * FY;

__device__ __forceinline__ int jwrap(int j, int jm) {
  int r = (j - 1) % jm;
  if (r < 0) r += jm;
  return r + 1;
}

// Operands of one step, streamed three steps ahead of their use.
struct SorSlot {
  float ro, rr, cr;  // red: previous iterate, rhs, cn1 (level kr)
  float wr;          // red: periodic same-colour neighbour (jm odd, rows 1 and jm; level kr)
  float rb, cb;      // black: rhs, cn1 (level kr-1)
  float wb;          // black: periodic same-colour neighbour (level kr-1)
  float bq;          // black previous iterate (level kr+1), for the LDS ring
};

// Buffer access with a per-lane byte offset and a uniform (scalar) byte
// offset: no per-access address arithmetic.  All colour-split arrays live
// in one allocation, addressed through one descriptor; the array and the
// level enter through the scalar offset.
typedef __amdgpu_buffer_rsrc_t Rsrc;
__device__ __forceinline__ Rsrc brsrc(const float* p, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)bytes, 0x00020000);
}
// soff is the same in every lane (block- and level-derived); readfirstlane
// states so, keeping it in a scalar register
__device__ __forceinline__ float bload(Rsrc r, unsigned voff, unsigned soff) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, __builtin_amdgcn_readfirstlane(soff), 0));
}
__device__ __forceinline__ void bstore(Rsrc r, float v, unsigned voff, unsigned soff) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff,
                                        __builtin_amdgcn_readfirstlane(soff), 0);
}
// acc + x*x in FP64 (the square is exact), x taken as 0 where keep is false.
// The select stays on the FP32 value: the empty asm operand (no instruction)
// keeps the compiler from moving it after the conversion, where it takes a
// select per 32-bit half.
// omega * y, the relaxation of the SOR update (press.f95:93-104).  1 * y is
// y for every y, so with omega == 1 (OM1) the multiplication is dropped.
template <bool OM1>
__device__ __forceinline__ float relax(float omega, float y) {
  return OM1 ? y : omega * y;
}
__device__ __forceinline__ double acc_sq(double acc, float x, bool keep) {
  float z = keep ? x : 0.0f;
  asm("" : "+v"(z));
  return __builtin_fma((double)z, (double)z, acc);
}
__device__ __forceinline__ unsigned boff(const float* base, const float* p) {
  return (unsigned)((const char*)p - (const char*)base);
}

// two blocks per CU need at most 64 VGPRs per lane (8 waves per SIMD)
// ENTRY: iteration 1, which reads the ghost values p had on entry to press
// (pent); the other iterations are compiled without that code.  gate != 0:
// an ENTRY launch that runs only when ctrl->p_ghost_diff is raised (a
// launch of k_sor_iter2 for iteration 1 takes the other case).
// bid: the block's index within the launch of this domain.
template <bool SEP, bool ENTRY>
__device__ __forceinline__ void sor_iter_block(const Grid& g, const SplitGrid& s, const Metrics& m,
                                               const SplitBufs& sb, Ctrl* ctrl, int iter, int nparts, float omega,
                                               float pjuge, int ntt, int ntj, int kchunk, int remap,
                                               const float* __restrict__ pent, int nprev, int gate, int bid) {
  __shared__ float ldsB[4][FY][FX];  // previous black iterate, levels mod 4
  __shared__ float ldsR[2][FY][FX];  // new red values, levels mod 2
  __shared__ double lds[FNT / kWave];
  __shared__ int skip;
  extern __shared__ float lvm[];  // [3][kchunk + 2]: cn4l, cn4s, cncz of levels krs..kre
  const int lane = threadIdx.x;
  // a wave is one row of the block (blockDim.x == 64): row-derived values are wave-uniform
  const int row = __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int tid = row * FX + lane;
  if (tid == 0) skip = (ctrl->sor_done != 0) || (ENTRY && gate && ctrl->p_ghost_diff == 0);
  __syncthreads();
  if (skip) return;
  if (iter >= 2 && bid == 0) {  // stop test of the previous iteration (press.f95:119-121)
    const int pp = (iter - 1) & 1;
    double sum = 0.0;
    for (int b = tid; b < nprev; b += FNT) sum += sb.part[pp][1][b] + sb.part[pp][0][b];
    sum = block_sum<FNT>(sum, tid, lds);
    if (tid == 0) {
      if (iter - 1 <= kSorHist) ctrl->sor_hist[iter - 2] = sum;
      ctrl->sor_iters = iter - 1;
      if (!(sum == sum) || sum > 1.0e300) ctrl->nonfinite = 1;
      if (sum < (double)pjuge || ctrl->nonfinite) ctrl->sor_done = iter - 1;
    }
    __syncthreads();  // lds is reused below
  }
  const int prv = (iter - 1) & 1, cur = iter & 1;
  const int im = g.im, jm = g.jm, km = g.km, nt = s.nt, cp = s.cp, cpl = s.cplane;
  const float* base = sb.p[0][0];  // the first of the eight split arrays of one allocation
  const Rsrc S = brsrc(base, 8u * (unsigned)cpl * (unsigned)km * 4u);
  const unsigned oRo = boff(base, sb.p[prv][1]), oBo = boff(base, sb.p[prv][0]);
  const unsigned oRn = boff(base, sb.p[cur][1]), oBn = boff(base, sb.p[cur][0]);
  const unsigned orR = boff(base, sb.rhs[1]), orB = boff(base, sb.rhs[0]);
  const unsigned ocR = SEP ? 0u : boff(base, sb.cn1[1]), ocB = SEP ? 0u : boff(base, sb.cn1[0]);
  const int lb = remap ? xcd_contiguous(bid, nparts) : bid;
  const int bt = lb % ntt, bj = (lb / ntt) % ntj, bk = lb / (ntt * ntj);
  // a tile as wide as the domain has no ring lanes: its i edges are the
  // walls, whose neighbours are mirrored
  const bool whole = ntt == 1 && nt <= FX;
  const int t = whole ? lane : bt * FTX - 2 + lane;
  const int jl = 1 + bj * FTY - 2 + row;  // row of the tile (may lie outside 1..jm)
  const int j = jwrap(jl, jm);            // the row it stands for
  const int k0 = 1 + bk * kchunk, k1 = min(km, k0 + kchunk - 1);
  const bool tin = (t >= 0) && (t < nt);
  const bool row_red = row >= 1 && row <= FY - 2;               // wave-uniform
  const bool row_tile = row >= 2 && row <= FY - 3 && jl <= jm;  // wave-uniform
  const bool lane_tile = tin && (whole || (lane >= 2 && lane <= FX - 3));
  const bool jodd = (jm & 1) != 0;
  const bool wlo = jodd && j == 1, whi = jodd && j == jm;  // same-colour periodic neighbours
  const int wofs = (jm - 1) * cp;
  // lanes outside the pairs of the row read pair 0 of their own row, so the
  // periodic row shift stays inside the array for every lane
  const int q0 = (tin ? t : 0) + (j - 1) * cp;
  const unsigned vo = 4u * (unsigned)q0;
  const unsigned vw = 4u * (unsigned)(q0 + (wlo ? wofs : (whi ? -wofs : 0)));
  const unsigned cplb = 4u * (unsigned)cpl;
  auto soff = [&](int k) { return (unsigned)(k - 1) * cplb; };
  // metrics of the two cells of the pair (i = 2t+1, 2t+2) and of row j
  float c2l0 = 0.f, c2l1 = 0.f, c2s0 = 0.f, c2s1 = 0.f, ax0 = 0.f, ax1 = 0.f;
  if (tin) {
    c2l0 = M1(m.cn2l, 2 * t + 1);
    c2l1 = M1(m.cn2l, 2 * t + 2);
    c2s0 = M1(m.cn2s, 2 * t + 1);
    c2s1 = M1(m.cn2s, 2 * t + 2);
    if (SEP) {
      ax0 = M1(m.cnax, 2 * t + 1);
      ax1 = M1(m.cnax, 2 * t + 2);
    }
  }
  const float c3l = M1(m.cn3l, j), c3s = M1(m.cn3s, j);
  const float byj = SEP ? M1(m.cnby, j) : 0.0f;
  const int krs = max(1, k0 - 1), kre = k1 + 1;
  const int nl = kchunk + 2;
  for (int x = tid; x < nl; x += FNT) {
    const int k = krs + x;
    if (k <= km) {
      lvm[x] = M1(m.cn4l, k);
      lvm[nl + x] = M1(m.cn4s, k);
      lvm[2 * nl + x] = SEP ? M1(m.cncz, k) : 0.0f;
    }
  }
  // The streams: red operands of level L, black operands of level L-1 and
  // the black iterate of level L+1, for the step at level L.  Every lane
  // issues the same loads in every step (levels clamped into 1..km), so the
  // number of loads in flight is the same on every path and a use waits
  // only for its own load; values of clamped levels or idle lanes are never
  // used.
  auto fill = [&](SorSlot& sl, int L) {
    const unsigned sr = soff(min(L, km)), sk = soff(max(1, min(L - 1, km)));
    sl.ro = bload(S, vo, oRo + sr);
    sl.rr = bload(S, vo, orR + sr);
    if (!SEP) sl.cr = bload(S, vo, ocR + sr);
    sl.wr = bload(S, vw, oRo + sr);
    sl.rb = bload(S, vo, orB + sk);
    if (!SEP) sl.cb = bload(S, vo, ocB + sk);
    sl.wb = bload(S, vw, oBo + sk);
    sl.bq = bload(S, vo, oBo + soff(min(L + 1, km)));
  };
  SorSlot s0 = {}, s1 = {}, s2 = {};
  fill(s0, krs);
  fill(s1, krs + 1);
  fill(s2, krs + 2);
  // the black iterate of levels krs-1 and krs enters the ring directly
  if (krs >= 2) ldsB[(krs - 1) & 3][row][lane] = bload(S, vo, oBo + soff(krs - 1));
  ldsB[krs & 3][row][lane] = bload(S, vo, oBo + soff(krs));
  float rm2 = 0.f;  // new red value of level kr-2 at this (pair, row)
  double acc = 0.0;
  const bool first_t = (t == 0);          // i = 1 when d = 0
  const bool last0 = (2 * t + 1 == im);   // i = im when d = 0
  const bool last1 = (2 * t + 2 == im);   // i = im when d = 1
  const bool ok0 = lane_tile && (2 * t + 1 <= im), ok1 = lane_tile && (2 * t + 2 <= im);

  auto step = [&](int kr, SorSlot& sl) {
    const bool active = kr <= kre;  // steps past kre only pad the last group of three
    const int kb = kr - 1;
    if (active && kr + 1 <= km) ldsB[(kr + 1) & 3][row][lane] = sl.bq;
    __syncthreads();
    // The red cell of level kr and the black cell of level kr-1 in this
    // (pair, row) have the same position d in the pair: i = 2t+1+d.  The
    // other colour's array holds i-1 at pair t-1+d and i+1 at pair t+d.
    const int d = (j + kr) & 1;  // wave-uniform
    const bool w = d ? ok1 : ok0;  // this lane owns an interior cell of the tile
    const bool at_first = d ? false : first_t;
    const bool at_last = d ? last1 : last0;
    const float c2l = d ? c2l1 : c2l0, c2s = d ? c2s1 : c2s0, axd = d ? ax1 : ax0;
    float r0 = 0.f;
    if (active && kr <= km && row_red) {
      // every lane of the wave computes (lanes outside the ring compute
      // values nobody reads); the tile's cells are stored
      const float own = sl.ro;
      const float* pl = &ldsB[kr & 3][row][lane - 1 + d];
      const float lo = pl[0], hi = pl[1];
      float pim = at_first ? own : lo;
      float pip = at_last ? own : hi;
      float pjm = wlo ? sl.wr : ldsB[kr & 3][row - 1][lane];
      float pjp = whi ? sl.wr : ldsB[kr & 3][row + 1][lane];
      float pkm = (kr == 1) ? own : ldsB[(kr - 1) & 3][row][lane];
      float pkp = (kr == km) ? own : ldsB[(kr + 1) & 3][row][lane];
      if (ENTRY && tin) {  // iteration 1: the ghost values p had on entry to press
        const int i = 2 * t + 1 + d;
        if (at_first) pim = pent[ix(g, 0, j, kr)];
        if (at_last) pip = pent[ix(g, im + 1, j, kr)];
        if (j == 1) pjm = pent[ix(g, i, 0, kr)];
        if (j == jm) pjp = pent[ix(g, i, jm + 1, kr)];
        if (kr == 1) pkm = pent[ix(g, i, j, 0)];
        if (kr == km) pkp = pent[ix(g, i, j, km + 1)];
      }
      const int x = kr - krs;
      const float cn = SEP ? rcp_rn((axd + byj) + lvm[2 * nl + x]) : sl.cr;  // init.f95:100-107
      const float reltmp = omega * (cn * (c2l * pip + c2s * pim + c3l * pjp + c3s * pjm + lvm[x] * pkp +
                                          lvm[nl + x] * pkm - sl.rr) -
                                    own);
      r0 = own + reltmp;
      ldsR[kr & 1][row][lane] = r0;
      if (kr >= k0 && kr <= k1 && row_tile && w) {
        bstore(S, r0, vo, oRn + soff(kr));
        // the FP64 product of two FP32 values is exact, so the fused
        // multiply-add rounds exactly as multiply then add
        acc = __builtin_fma((double)reltmp, (double)reltmp, acc);
      }
    }
    if (active && kb >= k0 && row_tile) {
      const float own = ldsB[kb & 3][row][lane];
      const float* pl = &ldsR[kb & 1][row][lane - 1 + d];
      const float lo = pl[0], hi = pl[1];
      const float pim = at_first ? own : lo;
      const float pip = at_last ? own : hi;
      const float pjm = wlo ? sl.wb : ldsR[kb & 1][row - 1][lane];
      const float pjp = whi ? sl.wb : ldsR[kb & 1][row + 1][lane];
      float pkm = (kb == 1) ? own : rm2;
      float pkp = (kb == km) ? own : r0;
      if (ENTRY && tin) {  // iteration 1: boundp2 refreshes the k ghosts only after the black sweep
        const int i = 2 * t + 1 + d;
        if (kb == 1) pkm = pent[ix(g, i, j, 0)];
        if (kb == km) pkp = pent[ix(g, i, j, km + 1)];
      }
      const int x = kb - krs;
      const float cn = SEP ? rcp_rn((axd + byj) + lvm[2 * nl + x]) : sl.cb;
      const float reltmp = omega * (cn * (c2l * pip + c2s * pim + c3l * pjp + c3s * pjm + lvm[x] * pkp +
                                          lvm[nl + x] * pkm - sl.rb) -
                                    own);
      if (w) {
        bstore(S, own + reltmp, vo, oBn + soff(kb));
        acc = __builtin_fma((double)reltmp, (double)reltmp, acc);
      }
    }
    rm2 = ldsR[(kr - 1) & 1][row][lane];  // red of level kr-1, for the next step's black
    fill(sl, kr + 3);
  };
  // unrolled by the stream depth, with no conditional step, so that each
  // slot stays in its own registers from load to use
  const int ngroups = (kre - krs + 3) / 3;
  for (int grp = 0; grp < ngroups; grp++) {
    const int kr = krs + 3 * grp;
    step(kr, s0);
    step(kr + 1, s1);
    step(kr + 2, s2);
  }
  const double sum = block_sum<FNT>(acc, tid, lds);
  if (tid == 0) {
    sb.part[cur][0][bid] = sum;
    sb.part[cur][1][bid] = 0.0;
  }
}
template <bool SEP, bool ENTRY>
__global__ __launch_bounds__(FNT) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_sor_iter(
    Grid g, SplitGrid s, Metrics m, SplitBufs sb, Ctrl* ctrl, int iter, int nparts, float omega, float pjuge,
    int ntt, int ntj, int kchunk, int remap, const float* __restrict__ pent, int nprev, int gate) {
  sor_iter_block<SEP, ENTRY>(g, s, m, sb, ctrl, iter, nparts, omega, pjuge, ntt, ntj, kchunk, remap, pent, nprev,
                             gate, (int)blockIdx.x);
}

// ------------------------------- fused SOR iteration, two rows per thread
// The same tile (FTX pairs x FTY rows, two-cell ring), k march and per-cell
// arithmetic as k_sor_iter, so the iterates are bit-identical, with half the
// threads: wave w owns tile rows 2w and 2w+1.  For even jm the first of the
// two rows is always odd, so the pair positions of a thread's cells at step
// kr follow from the parity of kr alone: the step loop is compiled for both
// start parities and every position-dependent choice is made at compile
// time.  (Odd jm, whose periodic rows break that, uses k_sor_iter.)  A
// thread keeps what it can of its two cells in registers:
//  * the previous black iterate of its cells at levels kr-1, kr, kr+1 (a
//    six-level ring filled four steps ahead), which gives the red update's
//    k neighbours, the black update's own value and, across the thread's
//    two rows, one j neighbour of each red update;
//  * the new red values of levels kr-2..kr, for the black update's k
//    neighbours and one of its j neighbours.
// LDS carries only the in-plane neighbours owned by other threads (the
// previous black iterate of level kr, three planes; the new red values of
// level kr-1, two planes) and the per-level metrics.  The step loop is
// unrolled by six, the period of every ring, so ring slots and LDS planes
// are fixed at compile time.  Every lane issues the same loads and stores on
// every step: stores of cells the thread does not own go to the padding
// pair of their row (cp > nt), those of levels outside its chunk to the
// scratch plane at the start of the allocation; loads of levels outside
// 1..km read neighbouring planes of the same allocation and are never used.

// TY: tile rows including the ring (16 or 32); W: tile width in pairs
// including the ring (64, or 32 with 32 rows: a wave then holds two half-wave
// rows, which fits narrow domains better).  The block has TY * W / 128
// waves.  nprev: partial-sum slots written by the previous iteration's
// launch.  One block's share of the iteration: bid is the block's index
// within the launch of this domain.  gate1 != 0: this is press's iteration
// 1 launched beside the entry-ghost kernel, and runs only when
// ctrl->p_ghost_diff is clear (then it equals the entry-ghost iteration).
// Other iteration-1 launches (the inner solves of press mode 1, which start
// from e = 0 with mirrored ghosts) pass gate1 == 0 and always run.
template <bool SEP, int TY, bool OM1, int W>
__device__ __forceinline__ void sor_iter2_block(Grid g, SplitGrid s, Metrics m, const SplitBufs& sb, Ctrl* ctrl, int iter,
                                                int nparts, int nprev, float omega, float pjuge, int ntt, int ntj,
                                                int kchunk, int remap, int bid, int gate1) {
  static_assert(W == FX || (W == FX / 2 && TY == 32), "tile shapes: 16 or 32 rows x 64 pairs, 32 x 32");
  constexpr int GNT = W * TY / 2, FTYT = TY - 4, FTXW = W - 4;
  __shared__ float ldsB[3][TY][W];  // previous black iterate of level kr (planes by level mod 3)
  __shared__ float ldsR[2][TY][W];  // new red values (planes by level mod 2)
  __shared__ double lds[GNT / kWave];
  __shared__ int skip;
  extern __shared__ float4 lvm4[];  // per level from krs: cn4l, cn4s, cncz
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.y);  // blockDim.x is one wave (FX)
  const int tid = w * FX + (int)threadIdx.x;
  // lane = pair of the tile row; vw = the thread's pair of rows (wave-uniform when W == FX)
  const int lane = (W == FX) ? (int)threadIdx.x : (int)threadIdx.x % W;
  const int vw = (W == FX) ? w : w * (FX / W) + (int)threadIdx.x / W;
  if (tid == 0) skip = (ctrl->sor_done != 0) || (gate1 && iter == 1 && ctrl->p_ghost_diff != 0);
  __syncthreads();
  if (skip) return;
  if (iter >= 2 && bid == 0) {  // stop test of the previous iteration (press.f95:119-121)
    const int pp = (iter - 1) & 1;
    double sum = 0.0;
    for (int b = tid; b < nprev; b += GNT) sum += sb.part[pp][1][b] + sb.part[pp][0][b];
    sum = block_sum<GNT>(sum, tid, lds);
    if (tid == 0) {
      if (iter - 1 <= kSorHist) ctrl->sor_hist[iter - 2] = sum;
      ctrl->sor_iters = iter - 1;
      if (!(sum == sum) || sum > 1.0e300) ctrl->nonfinite = 1;
      if (sum < (double)pjuge || ctrl->nonfinite) ctrl->sor_done = iter - 1;
    }
  }
  const int prv = (iter - 1) & 1, cur = iter & 1;
  const int im = g.im, jm = g.jm, km = g.km, nt = s.nt, cp = s.cp;
  const unsigned cplb = 4u * (unsigned)s.cplane;
  const float* base = sb.base;
  const Rsrc S = brsrc(base, (unsigned)sb.nplanes * cplb);
  // byte offset of level 0 of each array: level L is at o + L * cplb
  const unsigned oRo = boff(base, sb.p[prv][1]) - cplb, oBo = boff(base, sb.p[prv][0]) - cplb;
  const unsigned oRn = boff(base, sb.p[cur][1]) - cplb, oBn = boff(base, sb.p[cur][0]) - cplb;
  const unsigned orR = boff(base, sb.rhs[1]) - cplb, orB = boff(base, sb.rhs[0]) - cplb;
  const unsigned ocR = SEP ? 0u : boff(base, sb.cn1[1]) - cplb, ocB = SEP ? 0u : boff(base, sb.cn1[0]) - cplb;
  const int lb = remap ? xcd_contiguous(bid, nparts) : bid;
  const int bt = lb % ntt, bj = (lb / ntt) % ntj, bk = lb / (ntt * ntj);
  // a tile as wide as the domain has no ring lanes: its i edges are the
  // walls, whose neighbours are mirrored
  const bool whole = ntt == 1 && nt <= W;
  const int t = whole ? lane : bt * FTXW - 2 + lane;
  const bool tin = (t >= 0) && (t < nt);
  const int rA = 2 * vw, rB = rA + 1;
  const int jlA = 1 + bj * FTYT - 2 + rA, jlB = jlA + 1;  // tile rows (may lie outside 1..jm)
  const int jA = jwrap(jlA, jm), jB = jwrap(jlB, jm);     // the rows they stand for (jA odd)
  const int k0 = 1 + bk * kchunk, k1 = min(km, k0 + kchunk - 1);
  const int krs = max(1, k0 - 1), kre = k1 + 1;
  const int tq = tin ? t : 0;
  const unsigned voA = 4u * (unsigned)(tq + (jA - 1) * cp), voB = 4u * (unsigned)(tq + (jB - 1) * cp);
  const unsigned trA = 4u * (unsigned)(nt + (jA - 1) * cp), trB = 4u * (unsigned)(nt + (jB - 1) * cp);
  const bool lane_tile = tin && (whole || (lane >= 2 && lane <= W - 3));
  const bool tileA = rA >= 2 && rA <= TY - 3 && jlA <= jm, tileB = rB >= 2 && rB <= TY - 3 && jlB <= jm;
  // per position p of the pair (i = 2t+1+p)
  float c2l[2] = {0.f, 0.f}, c2s[2] = {0.f, 0.f}, axp[2] = {0.f, 0.f};
  if (tin) {
    for (int p = 0; p < 2; p++) {
      c2l[p] = M1(m.cn2l, 2 * t + 1 + p);
      c2s[p] = M1(m.cn2s, 2 * t + 1 + p);
      if (SEP) axp[p] = M1(m.cnax, 2 * t + 1 + p);
    }
  }
  const float c3lA = M1(m.cn3l, jA), c3sA = M1(m.cn3s, jA), c3lB = M1(m.cn3l, jB), c3sB = M1(m.cn3s, jB);
  const float byA = SEP ? M1(m.cnby, jA) : 0.f, byB = SEP ? M1(m.cnby, jB) : 0.f;
  float sA[2], sB[2];
  unsigned svA[2], svB[2];
  bool okA[2], okB[2], last[2];
  for (int p = 0; p < 2; p++) {
    sA[p] = axp[p] + byA;
    sB[p] = axp[p] + byB;
    okA[p] = lane_tile && tileA && (2 * t + 1 + p <= im);
    okB[p] = lane_tile && tileB && (2 * t + 1 + p <= im);
    svA[p] = okA[p] ? voA : trA;
    svB[p] = okB[p] ? voB : trB;
    last[p] = (2 * t + 1 + p == im);
  }
  const bool first = (t == 0);  // i = 1 (position 0)
  constexpr int PL = TY * W;     // floats per LDS plane
  const int ownA = rA * W + lane;
  // the other colour's i-1 neighbour of position p is at own-1+p: bases kept
  // inside the plane (only outer-ring lanes of rows 0 and TY-1 are clamped)
  const int bsA = min(max(ownA - 1, 0), PL - 3), bsB = min(max(ownA + W - 1, 0), PL - 3);
  const int upA = (rA == 0 ? rA : rA - 1) * W + lane;       // row A-1 (row 0 reads itself, unused)
  const int dnB = (rB == TY - 1 ? rB : rB + 1) * W + lane;  // row B+1
  const int nsteps = kre - krs + 1;
  for (int x = tid; x < nsteps; x += GNT) {
    const int k = krs + x;
    float4 v = {0.f, 0.f, 0.f, 0.f};
    if (k <= km) {
      v.x = M1(m.cn4l, k);
      v.y = M1(m.cn4s, k);
      v.z = SEP ? M1(m.cncz, k) : 0.f;
    }
    lvm4[x] = v;
  }
  // streams: red operands of level L and black operands of level L-1,
  // three steps ahead (slots by step mod 3); the previous black iterate by
  // level mod 6, four levels ahead of the step
  struct Sl {
    float roA, rrA, rbA, roB, rrB, rbB;
    float crA, cbA, crB, cbB;  // !SEP
  };
  // Levels past the chunk's last use are clamped to it: the streams run
  // ahead of the march, and reloading a line just read costs no HBM traffic.
  const int lbq = min(kre + 1, km + 1);  // last previous-black level used (red of level kre)
  auto fill = [&](Sl& sl, int L) {
    const unsigned lr = (unsigned)min(L, kre) * cplb, lk = (unsigned)min(L - 1, kre - 1) * cplb;
    sl.roA = bload(S, voA, oRo + lr);
    sl.roB = bload(S, voB, oRo + lr);
    sl.rrA = bload(S, voA, orR + lr);
    sl.rrB = bload(S, voB, orR + lr);
    sl.rbA = bload(S, voA, orB + lk);
    sl.rbB = bload(S, voB, orB + lk);
    if (!SEP) {
      sl.crA = bload(S, voA, ocR + lr);
      sl.crB = bload(S, voB, ocR + lr);
      sl.cbA = bload(S, voA, ocB + lk);
      sl.cbB = bload(S, voB, ocB + lk);
    }
  };
  Sl sl[3];
  float bqA[6], bqB[6];  // previous black iterate, level L at (L - krs + 1) mod 6
  // Level km+1 is the ghost boundp2 fills from level km (press.f95:107,
  // boundp.f95:28-78), which for the red update at km is the cell itself:
  // that slot of the ring takes the red iterate of level km, and the red
  // update reads its k+1 neighbour from the ring without a test.
  auto loadbq = [&](float& a, float& b, int L) {
    const int Lc = min(L, lbq);
    const unsigned o = (Lc == km + 1) ? oRo + (unsigned)km * cplb : oBo + (unsigned)Lc * cplb;
    a = bload(S, voA, o);
    b = bload(S, voB, o);
  };
  fill(sl[0], krs);
  fill(sl[1], krs + 1);
  fill(sl[2], krs + 2);
  for (int q = 0; q < 5; q++) loadbq(bqA[q], bqB[q], krs - 1 + q);  // levels krs-1 .. krs+3
  float* const LB = &ldsB[0][0][0];
  float* const LR = &ldsR[0][0][0];
  LB[ownA] = bqA[1];  // level krs into plane 0
  LB[ownA + W] = bqB[1];
  float rdA[3] = {0.f, 0.f, 0.f}, rdB[3] = {0.f, 0.f, 0.f};  // new red values, level L at (L - krs) mod 3
  if (krs == 1) {
    // level 0 is the ghost boundp2 fills from level 1: the k-1 neighbour of
    // the red update at level 1 is the red cell itself, that of the black
    // update at level 1 the black cell itself
    bqA[0] = sl[0].roA;
    bqB[0] = sl[0].roB;
    rdA[2] = bqA[1];
    rdB[2] = bqB[1];
  }
  float mx[2] = {0.f, 0.f}, my[2] = {0.f, 0.f}, mz[2] = {0.f, 0.f};  // metrics of the step's level, by parity
  double acc = 0.0;
  __syncthreads();  // lvm4 and plane 0 are complete

  // D0: position of row A's cells at the first step (row B's is the other)
  auto march = [&](auto D0C) {
    constexpr int D0 = decltype(D0C)::value;
    auto step = [&](auto UC, int kr) {
      constexpr int U = decltype(UC)::value;
      constexpr int P = U & 1;
      constexpr int pa = D0 ^ P, pb = pa ^ 1;
      const int kb = kr - 1;
      // 1. previous black iterate of level kr+1 into its plane
      LB[((U + 1) % 3) * PL + ownA] = bqA[(U + 2) % 6];
      LB[((U + 1) % 3) * PL + ownA + W] = bqB[(U + 2) % 6];
      {
        const float4 v = lvm4[kr - krs];
        mx[P] = v.x;
        my[P] = v.y;
        mz[P] = v.z;
      }
      __syncthreads();
      const bool kinR = (unsigned)(kr - k0) <= (unsigned)(k1 - k0);
      const bool kinB = (unsigned)(kb - k0) <= (unsigned)(k1 - k0);
      const unsigned soR = kinR ? oRn + (unsigned)kr * cplb : 0u;
      const unsigned soB = kinB ? oBn + (unsigned)kb * cplb : 0u;
      const bool blast = kb == km;  // levels 0 and km+1 of the rings hold the ghost values
      Sl& q = sl[U % 3];
      const float* pb_ = LB + (U % 3) * PL;
      // 2. red, level kr (press.f95:93-104), rows A and B
      float redA, redB;
      {
        const float own = q.roA;
        const float lo = pb_[bsA + pa], hi = pb_[bsA + pa + 1];
        const float pim = (pa == 0 && first) ? own : lo;
        const float pip = last[pa] ? own : hi;
        const float pjm = pb_[upA];
        const float pjp = bqB[(U + 1) % 6];  // row B's previous black iterate, level kr
        const float pkm = bqA[U % 6];
        const float pkp = bqA[(U + 2) % 6];
        const float cn = SEP ? rcp_rn(sA[pa] + mz[P]) : q.crA;  // init.f95:100-107
        const float reltmp = relax<OM1>(omega, cn * (c2l[pa] * pip + c2s[pa] * pim + c3lA * pjp + c3sA * pjm +
                                            mx[P] * pkp + my[P] * pkm - q.rrA) -
                                      own);
        redA = own + reltmp;
        bstore(S, redA, svA[pa], soR);
        acc = acc_sq(acc, reltmp, okA[pa] && kinR);  // exact square, as k_sor_iter
      }
      {
        const float own = q.roB;
        const float lo = pb_[bsB + pb], hi = pb_[bsB + pb + 1];
        const float pim = (pb == 0 && first) ? own : lo;
        const float pip = last[pb] ? own : hi;
        const float pjm = bqA[(U + 1) % 6];  // row A's previous black iterate, level kr
        const float pjp = pb_[dnB];
        const float pkm = bqB[U % 6];
        const float pkp = bqB[(U + 2) % 6];
        const float cn = SEP ? rcp_rn(sB[pb] + mz[P]) : q.crB;
        const float reltmp = relax<OM1>(omega, cn * (c2l[pb] * pip + c2s[pb] * pim + c3lB * pjp + c3sB * pjm +
                                            mx[P] * pkp + my[P] * pkm - q.rrB) -
                                      own);
        redB = own + reltmp;
        bstore(S, redB, svB[pb], soR);
        acc = acc_sq(acc, reltmp, okB[pb] && kinR);
      }
      LR[P * PL + ownA] = redA;
      LR[P * PL + ownA + W] = redB;
      rdA[U % 3] = redA;
      rdB[U % 3] = redB;
      // 3. black, level kb = kr-1 (press.f95:93-104), from the red values of
      // levels kb-1..kb+1; its in-plane red neighbours were written last step
      const float* pr = LR + (P ^ 1) * PL;
      {
        const float own = bqA[U % 6];
        const float lo = pr[bsA + pa], hi = pr[bsA + pa + 1];
        const float pim = (pa == 0 && first) ? own : lo;
        const float pip = last[pa] ? own : hi;
        const float pjm = pr[upA];
        const float pjp = rdB[(U + 2) % 3];  // row B's red value, level kb
        const float pkm = rdA[(U + 1) % 3];
        const float pkp = blast ? own : redA;
        const float cn = SEP ? rcp_rn(sA[pa] + mz[P ^ 1]) : q.cbA;
        const float reltmp = relax<OM1>(omega, cn * (c2l[pa] * pip + c2s[pa] * pim + c3lA * pjp + c3sA * pjm +
                                            mx[P ^ 1] * pkp + my[P ^ 1] * pkm - q.rbA) -
                                      own);
        bstore(S, own + reltmp, svA[pa], soB);
        acc = acc_sq(acc, reltmp, okA[pa] && kinB);
      }
      {
        const float own = bqB[U % 6];
        const float lo = pr[bsB + pb], hi = pr[bsB + pb + 1];
        const float pim = (pb == 0 && first) ? own : lo;
        const float pip = last[pb] ? own : hi;
        const float pjm = rdA[(U + 2) % 3];  // row A's red value, level kb
        const float pjp = pr[dnB];
        const float pkm = rdB[(U + 1) % 3];
        const float pkp = blast ? own : redB;
        const float cn = SEP ? rcp_rn(sB[pb] + mz[P ^ 1]) : q.cbB;
        const float reltmp = relax<OM1>(omega, cn * (c2l[pb] * pip + c2s[pb] * pim + c3lB * pjp + c3sB * pjm +
                                            mx[P ^ 1] * pkp + my[P ^ 1] * pkm - q.rbB) -
                                      own);
        bstore(S, own + reltmp, svB[pb], soB);
        acc = acc_sq(acc, reltmp, okB[pb] && kinB);
      }
      // 4. operands of later steps
      fill(q, kr + 3);
      loadbq(bqA[(U + 5) % 6], bqB[(U + 5) % 6], kr + 4);
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using I3 = std::integral_constant<int, 3>;
    using I4 = std::integral_constant<int, 4>;
    using I5 = std::integral_constant<int, 5>;
    int kr = krs;
    for (; kr + 5 <= kre; kr += 6) {
      step(I0{}, kr);
      step(I1{}, kr + 1);
      step(I2{}, kr + 2);
      step(I3{}, kr + 3);
      step(I4{}, kr + 4);
      step(I5{}, kr + 5);
    }
    const int rest = kre - kr + 1;  // 0..5 steps left, block-uniform
    if (rest > 0) step(I0{}, kr);
    if (rest > 1) step(I1{}, kr + 1);
    if (rest > 2) step(I2{}, kr + 2);
    if (rest > 3) step(I3{}, kr + 3);
    if (rest > 4) step(I4{}, kr + 4);
  };
  // row A (jA odd) at level krs: position (jA + krs) & 1
  if (((1 + krs) & 1) == 0)
    march(std::integral_constant<int, 0>{});
  else
    march(std::integral_constant<int, 1>{});
  const double sum = block_sum<GNT>(acc, tid, lds);
  if (tid == 0) {
    sb.part[cur][0][bid] = sum;
    sb.part[cur][1][bid] = 0.0;
  }
}

template <bool SEP, int TY, bool OM1, int W>
__global__ __launch_bounds__(W * TY / 2) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_sor_iter2(
    Grid g, SplitGrid s, Metrics m, SplitBufs sb, Ctrl* ctrl, int iter, int nparts, int nprev, float omega,
    float pjuge, int ntt, int ntj, int kchunk, int remap, int gate1) {
  sor_iter2_block<SEP, TY, OM1, W>(g, s, m, sb, ctrl, iter, nparts, nprev, omega, pjuge, ntt, ntj, kchunk, remap,
                                   (int)blockIdx.x, gate1);
}

// The same iteration for several domains in one launch: each domain owns a
// range of blocks (starting at a multiple of 8, so a block's XCD is the same
// as in a launch of its own), domains past their own nmaxp return at once.
// Only k_sor_iter2<true, TY, OM1> shapes are batched, with one TY for the
// batch (OM1: every domain has omega == 1).
template <int TY, bool OM1, int W>
__global__ __launch_bounds__(W * TY / 2) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_sor_iter2_batch(
    const SorBatchEntry* __restrict__ tab, const int* __restrict__ block0, int ntab, int iter, int remap) {
  int e = 0;
  for (int q = 1; q < ntab; q++)  // block0[q] == tab[q].block0, packed
    if ((int)blockIdx.x >= block0[q]) e = q;
  const SorBatchEntry& E = tab[e];
  const int bid = (int)blockIdx.x - E.block0;
  if (bid >= E.nparts || iter > E.nmaxp) return;  // padding block, or this domain has finished its iterations
  sor_iter2_block<true, TY, OM1, W>(E.g, E.s, E.m, E.sb, E.ctrl, iter, E.nparts, iter == 2 ? E.nparts1 : E.nparts,
                                    E.omega, E.pjuge, E.ntt, E.ntj, E.kchunk, remap, bid, 0);
}

void launch_sor_batch(const SorBatchEntry* tab, const int* block0, int ntab, int nblocks, int iter, int kcmax,
                      int ty, int tw, bool om1, int remap, hipStream_t stream) {
  const size_t lds = (size_t)(kcmax + 2) * sizeof(float4);
#define SORB(TY, W)                                                                                            \
  do {                                                                                                         \
    if (om1)                                                                                                   \
      hipLaunchKernelGGL((k_sor_iter2_batch<TY, true, W>), dim3(nblocks), dim3(FX, TY * W / (2 * FX)), lds,    \
                         stream, tab, block0, ntab, iter, remap);                                              \
    else                                                                                                       \
      hipLaunchKernelGGL((k_sor_iter2_batch<TY, false, W>), dim3(nblocks), dim3(FX, TY * W / (2 * FX)), lds,   \
                         stream, tab, block0, ntab, iter, remap);                                              \
  } while (0)
  if (ty == 32 && tw == FX / 2)
    SORB(32, FX / 2);
  else if (ty == 32)
    SORB(32, FX);
  else
    SORB(FY, FX);
#undef SORB
}

// ------------------------- two SOR iterations per launch (temporal blocking)
// One launch runs iterations l and l+1 of the fused red-black iteration in a
// single k-march, so the iterates and the rhs cross the memory system once
// per two iterations.  At step kr a thread computes, for its two tile rows:
//   A  red of iteration l at level kr (from the black iterate of l-1),
//   B  black of iteration l at level kr-1 (from A's red values),
//   C  red of iteration l+1 at level kr-2 (from B's black values),
//   D  black of iteration l+1 at level kr-3 (from C's red values).
// Each stage needs the previous stage's values one cell around, so the tile
// carries a four-cell ring (32 rows of which 24 are the tile, 64 pairs of
// which 56): A is valid on the tile and three rings, B on two, C on one, D
// on the tile alone, and the outermost waves skip the stages they cannot
// complete.  The march starts three levels below the block's chunk and ends
// three above it, each stage active only on the levels its successor reads.
// In-plane neighbours owned by other threads come through LDS, one set of
// planes per stage input (the black iterate of l-1 one level ahead, then
// two planes each for A, B and C); the k neighbours and the thread's other
// row come from register rings.  The red cell of level L and the black cell
// of level L-1 share their position in the pair, and levels kr, kr-2 (red)
// and kr-1, kr-3 (black) have the same parity, so all four stages of a step
// use the same positions.  Only iteration l+1 is stored (C and D); iteration
// l lives in registers and LDS.  Every update is the expression of
// k_sor_iter2 on the same operands (the boundp1/boundp2 ghost rules included:
// i and k ghosts mirror the cell, j wraps), so the iterates are those of two
// one-iteration launches bit for bit.  A launch of iteration 1 reads the
// ghost values p had on entry to press (pent) where they differ from their
// boundp copies (ctrl->p_ghost_diff), as k_sor_iter<SEP, true>.
//
// Iterate sets rotate over three: a launch reads in_set (p_{l-1}) and writes
// out_set, leaving the set before in_set intact.  The stop test runs one
// launch behind in block 0; when iteration l, the first of a launch, met the
// threshold, every later launch exits at once and a closing fix-up launch
// recomputes iteration l alone from that launch's in_set into its out_set.
struct PairLaunch {
  int iter;         // first iteration l of the launch
  int two;          // 1: iterations l and l+1; 0: l alone
  int in_set, out_set;  // iterate sets read (p_{l-1}) and written
  int wpart;        // partial-sum slots written: sb.part[wpart][0] iteration l, [1] iteration l+1
  int prev_iter;    // the previous launch (0: none), whose stop test block 0 runs ...
  int prev_two, prev_in, prev_out, prev_nparts, rpart;  // ... its shape, sets, slots
  int fixup;        // the fix-up launch: iteration and sets from ctrl->fix_*, one iteration
  int gate;         // 0: runs; 1: runs only when no ghost of p differs from its boundp copy; 2: only when one does
  int prev_colours; // the previous launch ran one iteration in a one-iteration kernel (iteration 1 of a batch):
                    // its partials are sb.part[rpart][1] (red) and [0] (black) of that iteration
};

// The stop test (press.f95:119-121) of a launch of iteration L (and L+1 when
// two) with FP64 sums s1 (and s2) of reltmp^2: records them and, at the first
// iteration below pjuge or non-finite, stops the loop with the result in
// out_set -- for the first iteration of a pair after the fix-up recomputes it.
__device__ __forceinline__ void pair_stop(Ctrl* ctrl, int L, int two, double s1, double s2, int in_set, int out_set,
                                          float pjuge) {
  if (L <= kSorHist) ctrl->sor_hist[L - 1] = s1;
  ctrl->sor_iters = L;
  ctrl->p_final = out_set;
  if (!(s1 == s1) || s1 > 1.0e300) ctrl->nonfinite = 1;
  if (s1 < (double)pjuge || ctrl->nonfinite) {
    ctrl->sor_done = L;
    if (two) {
      ctrl->fix_iter = L;
      ctrl->fix_in = in_set;
      ctrl->fix_out = out_set;
    }
    return;
  }
  if (!two) return;
  if (L + 1 <= kSorHist) ctrl->sor_hist[L] = s2;
  ctrl->sor_iters = L + 1;
  if (!(s2 == s2) || s2 > 1.0e300) ctrl->nonfinite = 1;
  if (s2 < (double)pjuge || ctrl->nonfinite) ctrl->sor_done = L + 1;
}

constexpr int PTY = 32, PR = 4, PNT = FX * PTY / 2, PFTY = PTY - 2 * PR, PFTX = FX - 2 * PR;

// ENT: the launch may hold iteration 1 (the entry-ghost code is compiled in).
// TW: 1 two iterations, 0 one (compiled in), -1 pl.two at run time (batches).
template <bool OM1, bool ENT, int TW>
__device__ __forceinline__ void sor_pair_block(const Grid& g, const SplitGrid& s, const Metrics& m,
                                               const SplitBufs& sb, Ctrl* ctrl, const PairLaunch& pl, int nparts,
                                               float omega, float pjuge, int ntt, int ntj, int kchunk, int remap,
                                               const float* __restrict__ pent, int bid) {
  constexpr int TY = PTY, W = FX, R = PR, GNT = PNT, FTYT = PFTY, FTXW = PFTX, PL = TY * W;
  __shared__ float ldsB[3 * PL];   // black iterate of l-1, level L in plane L mod 3 (by step)
  __shared__ float ldsR[2 * PL];   // A: red of l
  __shared__ float ldsB2[2 * PL];  // B: black of l
  __shared__ float ldsR2[2 * PL];  // C: red of l+1
  __shared__ double lds[GNT / kWave];
  __shared__ int sh[5];
  extern __shared__ float4 lvm4[];  // per level from krs: cn4l, cn4s, cncz
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int lane = (int)threadIdx.x;
  const int tid = w * FX + lane;
  if (tid == 0) {
    int skip = ctrl->sor_done != 0, it = pl.iter, si = pl.in_set, so = pl.out_set;
    if (pl.fixup) {
      skip = ctrl->fix_iter == 0;
      it = ctrl->fix_iter;
      si = ctrl->fix_in;
      so = ctrl->fix_out;
    }
    sh[1] = it;
    sh[2] = si;
    sh[3] = so;
    sh[4] = (it == 1) && ctrl->p_ghost_diff != 0;
    if (pl.gate != 0) skip = skip || ((ctrl->p_ghost_diff != 0) != (pl.gate == 2));
    sh[0] = skip;
  }
  __syncthreads();
  if (sh[0]) return;
  const int in_set = sh[2], out_set = sh[3];
  const bool two = (TW < 0) ? (pl.two != 0 && !pl.fixup) : (TW == 1);
  const bool ent = ENT && sh[4] != 0;  // iteration 1 on the entry ghosts of p
  if (!pl.fixup && pl.prev_iter > 0 && bid == 0) {  // the previous launch's stop test
    double s1 = 0.0, s2 = 0.0;
    if (pl.prev_colours) {
      for (int b = tid; b < pl.prev_nparts; b += GNT) s1 += sb.part[pl.rpart][1][b] + sb.part[pl.rpart][0][b];
    } else {
      for (int b = tid; b < pl.prev_nparts; b += GNT) {
        s1 += sb.part[pl.rpart][0][b];
        s2 += sb.part[pl.rpart][1][b];
      }
    }
    s1 = block_sum<GNT>(s1, tid, lds);
    s2 = block_sum<GNT>(s2, tid, lds);
    if (tid == 0) pair_stop(ctrl, pl.prev_iter, pl.prev_two, s1, s2, pl.prev_in, pl.prev_out, pjuge);
  }
  const int im = g.im, jm = g.jm, km = g.km, nt = s.nt, cp = s.cp;
  const unsigned cplb = 4u * (unsigned)s.cplane;
  const float* base = sb.base;
  const Rsrc S = brsrc(base, (unsigned)sb.nplanes * cplb);
  // byte offset of level 0 of each array: level L is at o + L * cplb
  const unsigned oRo = boff(base, sb.p[in_set][1]) - cplb, oBo = boff(base, sb.p[in_set][0]) - cplb;
  const unsigned oRn = boff(base, sb.p[out_set][1]) - cplb, oBn = boff(base, sb.p[out_set][0]) - cplb;
  const unsigned orR = boff(base, sb.rhs[1]) - cplb, orB = boff(base, sb.rhs[0]) - cplb;
  const int lb = remap ? xcd_contiguous(bid, nparts) : bid;
  const int bt = lb % ntt, bj = (lb / ntt) % ntj, bk = lb / (ntt * ntj);
  // a tile as wide as the domain has no ring lanes: its i edges are the
  // walls, whose neighbours are mirrored
  const bool whole = ntt == 1 && nt <= W;
  const int t = whole ? lane : bt * FTXW - R + lane;
  const bool tin = (t >= 0) && (t < nt);
  const int rA = 2 * w, rB = rA + 1;
  const int jlA = 1 + bj * FTYT - R + rA, jlB = jlA + 1;  // tile rows (may lie outside 1..jm); jlA odd
  const int jA = jwrap(jlA, jm), jB = jwrap(jlB, jm);     // the rows they stand for
  const int k0 = 1 + bk * kchunk, k1 = min(km, k0 + kchunk - 1);
  // levels each stage computes: the next stage reads one level around
  const int ra = two ? 3 : 1;
  const int aLo = max(1, k0 - ra), aHi = min(km, k1 + ra);
  const int bLo = max(1, k0 - ra + 1), bHi = min(km, k1 + ra - 1);
  const int cLo = max(1, k0 - 1), cHi = min(km, k1 + 1);
  const int krs = aLo, kend = k1 + ra;  // the steps: the last stage reaches k1 at kend
  const int tq = tin ? t : 0;
  const unsigned voA = 4u * (unsigned)(tq + (jA - 1) * cp), voB = 4u * (unsigned)(tq + (jB - 1) * cp);
  const unsigned trA = 4u * (unsigned)(nt + (jA - 1) * cp), trB = 4u * (unsigned)(nt + (jB - 1) * cp);
  const bool lane_tile = tin && (whole || (lane >= R && lane <= W - 1 - R));
  const bool tileA = rA >= R && rA <= TY - 1 - R && jlA <= jm, tileB = rB >= R && rB <= TY - 1 - R && jlB <= jm;
  // waves whose rows a stage's result is needed on (wave-uniform): B on
  // rows 2..TY-3, C on 3..TY-4, D on the tile rows 4..TY-5
  const bool wB = 2 * w + 1 >= 2 && 2 * w <= TY - 3;
  const bool wC = two && 2 * w + 1 >= 3 && 2 * w <= TY - 4;
  const bool wD = two && 2 * w + 1 >= 4 && 2 * w <= TY - 5;
  float c2l[2] = {0.f, 0.f}, c2s[2] = {0.f, 0.f}, axp[2] = {0.f, 0.f};
  if (tin) {
    for (int p = 0; p < 2; p++) {
      c2l[p] = M1(m.cn2l, 2 * t + 1 + p);
      c2s[p] = M1(m.cn2s, 2 * t + 1 + p);
      axp[p] = M1(m.cnax, 2 * t + 1 + p);
    }
  }
  const float c3lA = M1(m.cn3l, jA), c3sA = M1(m.cn3s, jA), c3lB = M1(m.cn3l, jB), c3sB = M1(m.cn3s, jB);
  const float byA = M1(m.cnby, jA), byB = M1(m.cnby, jB);
  float sA[2], sB[2];
  unsigned svA[2], svB[2];
  bool okA[2], okB[2], last[2];
  for (int p = 0; p < 2; p++) {
    sA[p] = axp[p] + byA;
    sB[p] = axp[p] + byB;
    okA[p] = lane_tile && tileA && (2 * t + 1 + p <= im);
    okB[p] = lane_tile && tileB && (2 * t + 1 + p <= im);
    svA[p] = okA[p] ? voA : trA;
    svB[p] = okB[p] ? voB : trB;
    last[p] = (2 * t + 1 + p == im);
  }
  const bool first = (t == 0);  // i = 1 (position 0)
  const int ownA = rA * W + lane;
  const int bsA = min(max(ownA - 1, 0), PL - 3), bsB = min(max(ownA + W - 1, 0), PL - 3);
  const int upA = (rA == 0 ? rA : rA - 1) * W + lane;       // row A-1 (row 0 reads itself, unused)
  const int dnB = (rB == TY - 1 ? rB : rB + 1) * W + lane;  // row B+1
  const int nsteps = kend - krs + 1;
  for (int x = tid; x < nsteps; x += GNT) {
    const int k = krs + x;
    float4 v = {0.f, 0.f, 0.f, 0.f};
    if (k <= km) {
      v.x = M1(m.cn4l, k);
      v.y = M1(m.cn4s, k);
      v.z = M1(m.cncz, k);
    }
    lvm4[x] = v;
  }
  // streams, two steps ahead (slots by step mod 2): A's own red of l-1 and
  // rhs at level L, B's rhs at level L-1; levels past their last use are
  // clamped to it (a reload of a line just read costs no HBM traffic)
  struct Sl {
    float roA, rrA, rbA, roB, rrB, rbB;
  };
  auto fill = [&](Sl& sl, int L) {
    const unsigned lr = (unsigned)min(L, aHi) * cplb, lk = (unsigned)min(max(L - 1, 1), bHi) * cplb;
    sl.roA = bload(S, voA, oRo + lr);
    sl.roB = bload(S, voB, oRo + lr);
    sl.rrA = bload(S, voA, orR + lr);
    sl.rrB = bload(S, voB, orR + lr);
    sl.rbA = bload(S, voA, orB + lk);
    sl.rbB = bload(S, voB, orB + lk);
  };
  Sl sl[2];
  // rings, by level: the black iterate of l-1 at (L - krs + 1) mod 6, four
  // levels ahead; A's red, B's black and C's red at (L - krs) mod 3
  float bqA[6], bqB[6];
  float rdA[3] = {0.f, 0.f, 0.f}, rdB[3] = {0.f, 0.f, 0.f};
  float blA[3] = {0.f, 0.f, 0.f}, blB[3] = {0.f, 0.f, 0.f};
  float r2A[3] = {0.f, 0.f, 0.f}, r2B[3] = {0.f, 0.f, 0.f};
  // Level km+1 is the ghost boundp2 fills from level km: for A's red update
  // at km the cell itself, so that slot takes the red iterate of level km.
  const int lbq = min(aHi + 1, km + 1);
  auto loadbq = [&](float& a, float& b, int L) {
    const int Lc = min(L, lbq);
    const unsigned o = (Lc == km + 1) ? oRo + (unsigned)km * cplb : oBo + (unsigned)Lc * cplb;
    a = bload(S, voA, o);
    b = bload(S, voB, o);
  };
  fill(sl[0], krs);
  fill(sl[1], krs + 1);
  for (int q = 0; q < 5; q++) loadbq(bqA[q], bqB[q], krs - 1 + q);  // levels krs-1 .. krs+3
  float* const LB = ldsB;
  float* const LR = ldsR;
  float* const LB2 = ldsB2;
  float* const LR2 = ldsR2;
  LB[ownA] = bqA[1];  // level krs into plane 0
  LB[ownA + W] = bqB[1];
  if (krs == 1) {
    // level 0 is the ghost boundp2 fills from level 1: the k-1 neighbour of
    // A's red update at level 1 is the red cell itself, that of B's black
    // update at level 1 the black cell itself
    bqA[0] = sl[0].roA;
    bqB[0] = sl[0].roB;
    rdA[2] = bqA[1];
    rdB[2] = bqB[1];
  }
  double acc1 = 0.0, acc2 = 0.0;
  __syncthreads();  // lvm4 and plane 0 are complete

  auto march = [&](auto D0C) {
    constexpr int D0 = decltype(D0C)::value;
    auto step = [&](auto UC, int kr) {
      constexpr int U = decltype(UC)::value;
      constexpr int P = U & 1;
      constexpr int pa = D0 ^ P, pb = pa ^ 1;
      const int kb = kr - 1, kc = kr - 2, kd = kr - 3;
      const bool actA = kr <= aHi;
      const bool actB = kb >= bLo && kb <= bHi;
      const bool actC = two && kc >= cLo && kc <= cHi;
      const bool actD = two && kd >= k0 && kd <= k1;
      // 1. the rhs of C's and D's cells, read again (the lines were read two
      // steps earlier for A and B, so they come from the caches), and the
      // black iterate of l-1 at level kr+1 into its plane
      float hrA = 0.f, hrB = 0.f, hbA = 0.f, hbB = 0.f;
      if (two) {
        const unsigned lc = (unsigned)min(max(kc, 1), km) * cplb, ld = (unsigned)min(max(kd, 1), km) * cplb;
        hrA = bload(S, voA, orR + lc);
        hrB = bload(S, voB, orR + lc);
        hbA = bload(S, voA, orB + ld);
        hbB = bload(S, voB, orB + ld);
      }
      LB[((U + 1) % 3) * PL + ownA] = bqA[(U + 2) % 6];
      LB[((U + 1) % 3) * PL + ownA + W] = bqB[(U + 2) % 6];
      __syncthreads();
      Sl& q = sl[U % 2];
      // ---- A: red of iteration l, level kr (press.f95:93-104)
      float redA = 0.f, redB = 0.f;
      if (actA) {
        const float4 mt = lvm4[kr - krs];
        const float* pp = LB + (U % 3) * PL;
        const bool kin = kr >= k0 && kr <= k1;
        {
          const float own = q.roA;
          const float lo = pp[bsA + pa], hi = pp[bsA + pa + 1];  // read by every lane
          float pim = (pa == 0 && first) ? own : lo;
          float pip = last[pa] ? own : hi;
          float pjm = pp[upA];
          float pjp = bqB[(U + 1) % 6];
          float pkm = bqA[U % 6];
          float pkp = bqA[(U + 2) % 6];
          if (ENT && ent && tin) {  // iteration 1: the ghost values p had on entry to press
            const int i = 2 * t + 1 + pa;
            if (pa == 0 && first) pim = pent[ix(g, 0, jA, kr)];
            if (last[pa]) pip = pent[ix(g, im + 1, jA, kr)];
            if (jA == 1) pjm = pent[ix(g, i, 0, kr)];
            if (jA == jm) pjp = pent[ix(g, i, jm + 1, kr)];
            if (kr == 1) pkm = pent[ix(g, i, jA, 0)];
            if (kr == km) pkp = pent[ix(g, i, jA, km + 1)];
          }
          const float cn = rcp_rn(sA[pa] + mt.z);  // init.f95:100-107
          const float reltmp = relax<OM1>(omega, cn * (c2l[pa] * pip + c2s[pa] * pim + c3lA * pjp + c3sA * pjm +
                                                       mt.x * pkp + mt.y * pkm - q.rrA) -
                                                     own);
          redA = own + reltmp;
          acc1 = acc_sq(acc1, reltmp, okA[pa] && kin);
          if (!two) bstore(S, redA, svA[pa], kin ? oRn + (unsigned)kr * cplb : 0u);
        }
        {
          const float own = q.roB;
          const float lo = pp[bsB + pb], hi = pp[bsB + pb + 1];  // read by every lane
          float pim = (pb == 0 && first) ? own : lo;
          float pip = last[pb] ? own : hi;
          float pjm = bqA[(U + 1) % 6];
          float pjp = pp[dnB];
          float pkm = bqB[U % 6];
          float pkp = bqB[(U + 2) % 6];
          if (ENT && ent && tin) {
            const int i = 2 * t + 1 + pb;
            if (pb == 0 && first) pim = pent[ix(g, 0, jB, kr)];
            if (last[pb]) pip = pent[ix(g, im + 1, jB, kr)];
            if (jB == 1) pjm = pent[ix(g, i, 0, kr)];
            if (jB == jm) pjp = pent[ix(g, i, jm + 1, kr)];
            if (kr == 1) pkm = pent[ix(g, i, jB, 0)];
            if (kr == km) pkp = pent[ix(g, i, jB, km + 1)];
          }
          const float cn = rcp_rn(sB[pb] + mt.z);
          const float reltmp = relax<OM1>(omega, cn * (c2l[pb] * pip + c2s[pb] * pim + c3lB * pjp + c3sB * pjm +
                                                       mt.x * pkp + mt.y * pkm - q.rrB) -
                                                     own);
          redB = own + reltmp;
          acc1 = acc_sq(acc1, reltmp, okB[pb] && kin);
          if (!two) bstore(S, redB, svB[pb], kin ? oRn + (unsigned)kr * cplb : 0u);
        }
        LR[P * PL + ownA] = redA;
        LR[P * PL + ownA + W] = redB;
        rdA[U % 3] = redA;
        rdB[U % 3] = redB;
      }
      // ---- B: black of iteration l, level kb = kr-1, from A's red of levels kb-1..kb+1
      if (actB && wB) {
        const float4 mt = lvm4[kb - krs];
        const float* pr = LR + (P ^ 1) * PL;
        const bool kin = kb >= k0 && kb <= k1;
        const bool top = kb == km;  // level km+1 is the ghost boundp2 fills from level km: the cell itself
        float bA, bB;
        {
          const float own = bqA[U % 6];
          const float lo = pr[bsA + pa], hi = pr[bsA + pa + 1];  // read by every lane
          const float pim = (pa == 0 && first) ? own : lo;
          const float pip = last[pa] ? own : hi;
          const float pjm = pr[upA];
          const float pjp = rdB[(U + 2) % 3];
          float pkm = rdA[(U + 1) % 3];
          float pkp = top ? own : redA;
          if (ENT && ent && tin) {  // iteration 1: boundp2 refreshes the k ghosts only after the black sweep
            const int i = 2 * t + 1 + pa;
            if (kb == 1) pkm = pent[ix(g, i, jA, 0)];
            if (kb == km) pkp = pent[ix(g, i, jA, km + 1)];
          }
          const float cn = rcp_rn(sA[pa] + mt.z);
          const float reltmp = relax<OM1>(omega, cn * (c2l[pa] * pip + c2s[pa] * pim + c3lA * pjp + c3sA * pjm +
                                                       mt.x * pkp + mt.y * pkm - q.rbA) -
                                                     own);
          bA = own + reltmp;
          acc1 = acc_sq(acc1, reltmp, okA[pa] && kin);
          if (!two) bstore(S, bA, svA[pa], kin ? oBn + (unsigned)kb * cplb : 0u);
        }
        {
          const float own = bqB[U % 6];
          const float lo = pr[bsB + pb], hi = pr[bsB + pb + 1];  // read by every lane
          const float pim = (pb == 0 && first) ? own : lo;
          const float pip = last[pb] ? own : hi;
          const float pjm = rdA[(U + 2) % 3];
          const float pjp = pr[dnB];
          float pkm = rdB[(U + 1) % 3];
          float pkp = top ? own : redB;
          if (ENT && ent && tin) {
            const int i = 2 * t + 1 + pb;
            if (kb == 1) pkm = pent[ix(g, i, jB, 0)];
            if (kb == km) pkp = pent[ix(g, i, jB, km + 1)];
          }
          const float cn = rcp_rn(sB[pb] + mt.z);
          const float reltmp = relax<OM1>(omega, cn * (c2l[pb] * pip + c2s[pb] * pim + c3lB * pjp + c3sB * pjm +
                                                       mt.x * pkp + mt.y * pkm - q.rbB) -
                                                     own);
          bB = own + reltmp;
          acc1 = acc_sq(acc1, reltmp, okB[pb] && kin);
          if (!two) bstore(S, bB, svB[pb], kin ? oBn + (unsigned)kb * cplb : 0u);
        }
        LB2[P * PL + ownA] = bA;
        LB2[P * PL + ownA + W] = bB;
        blA[(U + 2) % 3] = bA;
        blB[(U + 2) % 3] = bB;
      }
      // ---- C: red of iteration l+1, level kc = kr-2, from B's black of levels kc-1..kc+1
      float r2a = 0.f, r2b = 0.f;
      if (actC && wC) {
        const float4 mt = lvm4[kc - krs];
        const float* pq = LB2 + (P ^ 1) * PL;
        const bool kin = kc >= k0 && kc <= k1;
        const unsigned so = kin ? oRn + (unsigned)kc * cplb : 0u;
        {
          const float own = rdA[(U + 1) % 3];
          const float lo = pq[bsA + pa], hi = pq[bsA + pa + 1];  // read by every lane
          const float pim = (pa == 0 && first) ? own : lo;
          const float pip = last[pa] ? own : hi;
          const float pjm = pq[upA];
          const float pjp = blB[(U + 1) % 3];
          const float pkm = (kc == 1) ? own : blA[U % 3];
          const float pkp = (kc == km) ? own : blA[(U + 2) % 3];
          const float cn = rcp_rn(sA[pa] + mt.z);
          const float reltmp = relax<OM1>(omega, cn * (c2l[pa] * pip + c2s[pa] * pim + c3lA * pjp + c3sA * pjm +
                                                       mt.x * pkp + mt.y * pkm - hrA) -
                                                     own);
          r2a = own + reltmp;
          acc2 = acc_sq(acc2, reltmp, okA[pa] && kin);
          bstore(S, r2a, svA[pa], so);
        }
        {
          const float own = rdB[(U + 1) % 3];
          const float lo = pq[bsB + pb], hi = pq[bsB + pb + 1];  // read by every lane
          const float pim = (pb == 0 && first) ? own : lo;
          const float pip = last[pb] ? own : hi;
          const float pjm = blA[(U + 1) % 3];
          const float pjp = pq[dnB];
          const float pkm = (kc == 1) ? own : blB[U % 3];
          const float pkp = (kc == km) ? own : blB[(U + 2) % 3];
          const float cn = rcp_rn(sB[pb] + mt.z);
          const float reltmp = relax<OM1>(omega, cn * (c2l[pb] * pip + c2s[pb] * pim + c3lB * pjp + c3sB * pjm +
                                                       mt.x * pkp + mt.y * pkm - hrB) -
                                                     own);
          r2b = own + reltmp;
          acc2 = acc_sq(acc2, reltmp, okB[pb] && kin);
          bstore(S, r2b, svB[pb], so);
        }
        LR2[P * PL + ownA] = r2a;
        LR2[P * PL + ownA + W] = r2b;
        r2A[(U + 1) % 3] = r2a;
        r2B[(U + 1) % 3] = r2b;
      }
      // ---- D: black of iteration l+1, level kd = kr-3, from C's red of levels kd-1..kd+1
      if (actD && wD) {
        const float4 mt = lvm4[kd - krs];
        const float* pr = LR2 + (P ^ 1) * PL;
        const unsigned so = oBn + (unsigned)kd * cplb;  // kd is in the chunk whenever D runs
        {
          const float own = blA[U % 3];
          const float lo = pr[bsA + pa], hi = pr[bsA + pa + 1];  // read by every lane
          const float pim = (pa == 0 && first) ? own : lo;
          const float pip = last[pa] ? own : hi;
          const float pjm = pr[upA];
          const float pjp = r2B[U % 3];
          const float pkm = (kd == 1) ? own : r2A[(U + 2) % 3];
          const float pkp = (kd == km) ? own : r2a;
          const float cn = rcp_rn(sA[pa] + mt.z);
          const float reltmp = relax<OM1>(omega, cn * (c2l[pa] * pip + c2s[pa] * pim + c3lA * pjp + c3sA * pjm +
                                                       mt.x * pkp + mt.y * pkm - hbA) -
                                                     own);
          acc2 = acc_sq(acc2, reltmp, okA[pa]);
          bstore(S, own + reltmp, svA[pa], so);
        }
        {
          const float own = blB[U % 3];
          const float lo = pr[bsB + pb], hi = pr[bsB + pb + 1];  // read by every lane
          const float pim = (pb == 0 && first) ? own : lo;
          const float pip = last[pb] ? own : hi;
          const float pjm = r2A[U % 3];
          const float pjp = pr[dnB];
          const float pkm = (kd == 1) ? own : r2B[(U + 2) % 3];
          const float pkp = (kd == km) ? own : r2b;
          const float cn = rcp_rn(sB[pb] + mt.z);
          const float reltmp = relax<OM1>(omega, cn * (c2l[pb] * pip + c2s[pb] * pim + c3lB * pjp + c3sB * pjm +
                                                       mt.x * pkp + mt.y * pkm - hbB) -
                                                     own);
          acc2 = acc_sq(acc2, reltmp, okB[pb]);
          bstore(S, own + reltmp, svB[pb], so);
        }
      }
      // 4. operands of later steps
      fill(q, kr + 2);
      loadbq(bqA[(U + 5) % 6], bqB[(U + 5) % 6], kr + 4);
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using I3 = std::integral_constant<int, 3>;
    using I4 = std::integral_constant<int, 4>;
    using I5 = std::integral_constant<int, 5>;
    int kr = krs;
    for (; kr + 5 <= kend; kr += 6) {
      step(I0{}, kr);
      step(I1{}, kr + 1);
      step(I2{}, kr + 2);
      step(I3{}, kr + 3);
      step(I4{}, kr + 4);
      step(I5{}, kr + 5);
    }
    const int rest = kend - kr + 1;  // 0..5 steps left, block-uniform
    if (rest > 0) step(I0{}, kr);
    if (rest > 1) step(I1{}, kr + 1);
    if (rest > 2) step(I2{}, kr + 2);
    if (rest > 3) step(I3{}, kr + 3);
    if (rest > 4) step(I4{}, kr + 4);
  };
  // row A (jA odd) at level krs: position (jA + krs) & 1
  if (((1 + krs) & 1) == 0)
    march(std::integral_constant<int, 0>{});
  else
    march(std::integral_constant<int, 1>{});
  const double t1 = block_sum<GNT>(acc1, tid, lds);
  const double t2 = block_sum<GNT>(acc2, tid, lds);
  if (tid == 0) {
    sb.part[pl.wpart][0][bid] = t1;
    sb.part[pl.wpart][1][bid] = t2;
  }
}

template <bool OM1, bool ENT, int TW>
__global__ __launch_bounds__(PNT) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_sor_pair(
    Grid g, SplitGrid s, Metrics m, SplitBufs sb, Ctrl* ctrl, PairLaunch pl, int nparts, float omega, float pjuge,
    int ntt, int ntj, int kchunk, int remap, const float* __restrict__ pent) {
  sor_pair_block<OM1, ENT, TW>(g, s, m, sb, ctrl, pl, nparts, omega, pjuge, ntt, ntj, kchunk, remap, pent,
                           (int)blockIdx.x);
}

// After the last two-iteration launch: its stop test, unless an earlier one
// stopped the loop.
__global__ __launch_bounds__(256) void k_sor_pair_final(SplitBufs sb, Ctrl* ctrl, PairLaunch last, int nparts,
                                                        float pjuge) {
  __shared__ double lds[4];
  if (ctrl->sor_done != 0) return;  // uniform: written by earlier launches
  double s1 = 0.0, s2 = 0.0;
  for (int b = threadIdx.x; b < nparts; b += 256) {
    s1 += sb.part[last.wpart][0][b];
    s2 += sb.part[last.wpart][1][b];
  }
  s1 = block_sum<256>(s1, threadIdx.x, lds);
  s2 = block_sum<256>(s2, threadIdx.x, lds);
  if (threadIdx.x == 0) pair_stop(ctrl, last.iter, last.two, s1, s2, last.in_set, last.out_set, pjuge);
}

// Iterations 2.. of the batched domains (k_sor_iter2_batch's table), two
// per launch: launch pl of the sequence that starts at iteration 2 from
// iterate set 1 (iteration 1, run by the one-iteration entry kernel, wrote
// set 1 and its partials to slot 1).  Each domain's blocks run the launch
// with its own shape, skip the launches past its nmaxp, and run the last one
// alone when it holds only nmaxp; the fix-up (pl.fixup) takes its state from
// each domain's control block.
template <bool OM1>
__global__ __launch_bounds__(PNT) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_sor_pair_batch(
    const SorBatchEntry* __restrict__ tab, const int* __restrict__ block0, int ntab, PairLaunch pl, int remap) {
  int e = 0;
  for (int q = 1; q < ntab; q++)  // block0[q] == tab[q].block0, packed
    if ((int)blockIdx.x >= block0[q]) e = q;
  const SorBatchEntry& E = tab[e];
  const int bid = (int)blockIdx.x - E.block0;
  if (bid >= E.nparts) return;  // padding block
  if (!pl.fixup) {
    if (pl.iter > E.nmaxp) return;  // this domain has finished its iterations
    pl.two = pl.two && pl.iter + 1 <= E.nmaxp;
    pl.prev_nparts = pl.prev_colours ? E.nparts1 : E.nparts;
  }
  sor_pair_block<OM1, false, -1>(E.g, E.s, E.m, E.sb, E.ctrl, pl, E.nparts, E.omega, E.pjuge, E.ntt, E.ntj, E.kchunk,
                             remap, nullptr, bid);
}

// After the iteration loop: when no iteration met the stop test, the last
// one (nmaxp) is summed here.  Records the iteration count and sum.
__device__ __forceinline__ void sor_final_body(const SplitBufs& sb, Ctrl* ctrl, int nmaxp, int nparts) {
  __shared__ double lds[4];
  if (ctrl->sor_done != 0) return;  // uniform: written by earlier launches
  const int pp = nmaxp & 1;
  double sum = 0.0;
  for (int b = threadIdx.x; b < nparts; b += 256) sum += sb.part[pp][1][b] + sb.part[pp][0][b];
  sum = block_sum<256>(sum, threadIdx.x, lds);
  if (threadIdx.x == 0) {
    if (nmaxp <= kSorHist) ctrl->sor_hist[nmaxp - 1] = sum;
    ctrl->sor_iters = nmaxp;
    if (!(sum == sum) || sum > 1.0e300) ctrl->nonfinite = 1;
  }
}
__global__ __launch_bounds__(256) void k_sor_final(SplitBufs sb, Ctrl* ctrl, int nmaxp, int nparts) {
  sor_final_body(sb, ctrl, nmaxp, nparts);
}

// One block per batched domain after its two-iteration launches: the stop
// test of its last launch (k_sor_pair_final's, with that launch's iteration,
// sets and slots derived from the domain's nmaxp as k_sor_pair_batch's
// sequence gives them), or, for nmaxp = 1, the closing sum of iteration 1.
__global__ __launch_bounds__(256) void k_sor_pair_batch_final(const SorBatchEntry* __restrict__ tab) {
  __shared__ double lds[4];
  const SorBatchEntry& E = tab[blockIdx.x];
  Ctrl* ctrl = E.ctrl;
  if (E.nmaxp <= 1) {
    sor_final_body(E.sb, ctrl, E.nmaxp, E.nparts1);
    return;
  }
  if (ctrl->sor_done != 0) return;  // uniform: written by earlier launches
  const int q = (E.nmaxp - 2) / 2, L = 2 + 2 * q, two = L + 1 <= E.nmaxp;
  double s1 = 0.0, s2 = 0.0;
  for (int b = threadIdx.x; b < E.nparts; b += 256) {
    s1 += E.sb.part[q & 1][0][b];
    s2 += E.sb.part[q & 1][1][b];
  }
  s1 = block_sum<256>(s1, threadIdx.x, lds);
  s2 = block_sum<256>(s2, threadIdx.x, lds);
  if (threadIdx.x == 0) pair_stop(ctrl, L, two, s1, s2, (1 + q) % 3, (2 + q) % 3, E.pjuge);
}

// pav partials: p*dx1*dy1*dzn over the SOR result (press.f95:124-133).
__device__ __forceinline__ void pav_body(const VBlk& vb, const Grid& g, const SplitGrid& s, const Metrics& m,
                                         const SplitBufs& sb, double* __restrict__ part, const Ctrl* ctrl) {
  __shared__ double lds[NT / kWave];
  const int tid = threadIdx.y * BX + threadIdx.x;
  const int bid = vlinear(vb);
  const int fs = ctrl->p_final >= 0 ? ctrl->p_final : (ctrl->sor_iters & 1);  // the set holding the result
  const float* __restrict__ red = sb.p[fs][1];
  const float* __restrict__ blk = sb.p[fs][0];
  const int i = 1 + vb.x * BX + threadIdx.x, j = 1 + vb.y * BY + wave_row();
  double acc = 0.0;
  if (i <= g.im && j <= g.jm) {
    const float dxi = M1(m.dx1, i), dyj = M1(m.dy1, j);
    const int t = (i - 1) >> 1;
    int k0, k1;
    kspan(vb, 1, g.km, k0, k1);
    for (int k = k0; k <= k1; k++) {
      const float pv = (((i + j + k) & 1) ? red : blk)[cix(s, t, j, k)];
      acc += (double)(pv * dxi * dyj * M1(m.dzn, k));
    }
  }
  const double sum = block_sum<NT>(acc, tid, lds);
  if (tid == 0) part[bid] = sum;
}
__global__ __launch_bounds__(NT) void k_pav(Grid g, SplitGrid s, Metrics m, SplitBufs sb,
                                            double* __restrict__ part, const Ctrl* ctrl) {
  pav_body(hw_blk<false>(), g, s, m, sb, part, ctrl);
}

// p = p - pav on the interior (press.f95:145-150, FP64 subtraction) and the
// ghost cells of boundp1 then boundp2 (press.f95:155-159): every ghost
// receives the adjusted value of the interior cell the copies trace it to
// (i: 0->1, im+1->im; j: 0->jm, jm+1->1; k: 0->1, km+1->km).  Covers
// i,j,k = 0..im+1, 0..jm+1, 0..km+1.
__device__ __forceinline__ void padj_body(const VBlk& vb, const Grid& g, const SplitGrid& s, const SplitBufs& sb,
                                          float* __restrict__ p, const Ctrl* ctrl) {
  const int i = vb.x * BX + threadIdx.x, j = vb.y * BY + wave_row();
  if (i > g.im + 1 || j > g.jm + 1) return;
  const int fs = ctrl->p_final >= 0 ? ctrl->p_final : (ctrl->sor_iters & 1);  // the set holding the result
  const float* __restrict__ red = sb.p[fs][1];
  const float* __restrict__ blk = sb.p[fs][0];
  const double av = ctrl->pav;
  const int ii = (i == 0) ? 1 : ((i == g.im + 1) ? g.im : i);
  const int jj = (j == 0) ? g.jm : ((j == g.jm + 1) ? 1 : j);
  const int t = (ii - 1) >> 1;
  int k0, k1;
  kspan(vb, 0, g.km + 1, k0, k1);
  for (int k = k0; k <= k1; k++) {
    const int kk = (k == 0) ? 1 : ((k == g.km + 1) ? g.km : k);
    const float v = (((ii + jj + kk) & 1) ? red : blk)[cix(s, t, jj, kk)];
    p[ix(g, i, j, k)] = (float)((double)v - av);
  }
}
__global__ __launch_bounds__(NT) void k_padj(Grid g, SplitGrid s, SplitBufs sb, float* __restrict__ p,
                                             const Ctrl* ctrl) {
  padj_body(hw_blk<false>(), g, s, sb, p, ctrl);
}

// After the SOR loop: the closing stop-test sum, the pressure mean and the
// adjusted p with its ghosts.
static void launch_press_mean(const DevDomain& d, const SplitBufs& sb, int gx, int gy, int gz, int nred) {
  const Grid& g = d.g;
  const SplitGrid& s = d.sg;
  TIMED(d, KC_PAV, hipLaunchKernelGGL(k_pav, dim3(gx, gy, gz), blk(), 0, d.stream, g, s, d.m, sb, d.partials,
                                      (const Ctrl*)d.ctrl));
  TIMED(d, KC_PAV, hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, d.stream, d.partials, nred, d.ctrl, 1,
                                      d.area));
  const int px = (g.im + 2 + BX - 1) / BX, py = (g.jm + 2 + BY - 1) / BY;
  // a few levels per thread (a copy: no value depends on the split)
  const int pz = kblocks(px, py, g.km + 2, 8192);
  TIMED(d, KC_PADJ, hipLaunchKernelGGL(k_padj, dim3(px, py, pz),
                                       blk(), 0, d.stream, g, s, sb, d.F.p, (const Ctrl*)d.ctrl));
}
static void launch_press_tail(const DevDomain& d, const SplitBufs& sb, int nparts, int gx, int gy, int gz,
                              int nred) {
  TIMED(d, KC_PAV, hipLaunchKernelGGL(k_sor_final, dim3(1), dim3(256), 0, d.stream, sb, d.ctrl, d.nmaxp, nparts));
  launch_press_mean(d, sb, gx, gy, gz, nred);
}

// ------------------------------------------- FP64 defect correction
// The precision variant of press (press mode 1): the pressure equation of
// the SOR, B(p) = rhs with B(p) = cn2l p(i+1) + cn2s p(i-1) + cn3l p(j+1) +
// cn3s p(j-1) + cn4l p(k+1) + cn4s p(k-1) - p/cn1 (press.f95:95-98 at its
// fixed point; Neumann in i and k, periodic in j as boundp1/boundp2 make
// it), is solved by defect correction: the iterate P64 and the right-hand
// side R64 = rhs - rhsav are kept in FP64; each outer step forms the defect
// d = R64 - B(P64) in FP64, solves B(e) = d approximately with FP32 SOR
// iterations from e = 0 (k_sor_iter, unchanged), and adds e to P64.  The
// defect norms are recorded per outer step; p = P64 - pav at the end.

// R64 = div(F) + div(u)/dt evaluated in FP64 (press.f95:49-60 with the
// ghosts k_rhs has just written), block partials of dV*R64 for its mean;
// P64 = p, the previous pressure, as the starting iterate.
__global__ __launch_bounds__(NT) void k_dc_init(Grid g, Metrics m, const float* __restrict__ u,
                                                const float* __restrict__ v, const float* __restrict__ w,
                                                const float* __restrict__ f, const float* __restrict__ gg,
                                                const float* __restrict__ h, const float* __restrict__ p,
                                                double* __restrict__ R64, double* __restrict__ P64,
                                                double* __restrict__ part, double dt, Ctrl* ctrl) {
  __shared__ double lds[NT / kWave];
  const int tid = threadIdx.y * BX + threadIdx.x;
  const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  const int i = 1 + blockIdx.x * BX + threadIdx.x, j = 1 + blockIdx.y * BY + threadIdx.y;
  if (bid == 0 && tid == 0) {  // a new press call: no defect-correction step taken yet
    ctrl->dc_done = 0;
    ctrl->dc_norms = 0;
  }
  double acc = 0.0;
  if (i <= g.im && j <= g.jm) {
    const double dxi = M1(m.dx1, i), dyj = M1(m.dy1, j);
    const int sj = g.sj, sk = g.sk;
    int k0, k1;
    kspan(1, g.km, k0, k1);
    for (int k = k0; k <= k1; k++) {
      const int c = ix(g, i, j, k);
      const double dzk = M1(m.dzn, k);
      const double r = (-(double)u[c - 1] + u[c]) / dxi + (-(double)v[c - sj] + v[c]) / dyj +
                       (-(double)w[c - sk] + w[c]) / dzk;
      const double rr = ((double)f[c] - f[c - 1]) / dxi + ((double)gg[c] - gg[c - sj]) / dyj +
                        ((double)h[c] - h[c - sk]) / dzk + r / dt;
      R64[c] = rr;
      P64[c] = (double)p[c];
      acc += dxi * dyj * dzk * rr;
    }
  }
  const double sum = block_sum<NT>(acc, tid, lds);
  if (tid == 0) part[bid] = sum;
}

// d = (R64 - rhsav64) - B(P64) in FP64: its FP32 copy into the split rhs
// the SOR reads, e = 0 in the split iterate of parity 0, block partials of
// (cn1*d)^2 -- the measure of the reference's stop test (press.f95:93-104,
// reltmp = -cn1*d at omega 1 for a cell updated from P64); block 0 resets
// the SOR control for the inner solve.  (The defect is not projected: K
// inner iterations from e = 0 are then, in exact arithmetic, K further
// iterations of the SOR with the inner relaxation.)  Once an earlier outer
// step met the stop threshold (ctrl->dc_done) the launch only keeps the
// inner solve from running.
template <bool SEP>
__global__ __launch_bounds__(NT) void k_dc_defect(Grid g, SplitGrid s, Metrics m, const float* __restrict__ cn1f,
                                                  const double* __restrict__ R64,
                                                  const double* __restrict__ P64, SplitBufs sb,
                                                  double* __restrict__ part, Ctrl* ctrl) {
  __shared__ double lds[NT / kWave];
  const int tid = threadIdx.y * BX + threadIdx.x;
  const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  const double rav = ctrl->rhsav64;
  const int done = ctrl->dc_done;
  __syncthreads();  // every thread has read the control block before block 0 writes it
  if (bid == 0 && tid == 0) {
    ctrl->sor_done = done ? 1 : 0;  // a finished correction: the inner launches exit at once
    ctrl->nonfinite = 0;
  }
  if (done) return;  // uniform
  const int i = 1 + blockIdx.x * BX + threadIdx.x, j = 1 + blockIdx.y * BY + threadIdx.y;
  double acc = 0.0;
  if (i <= g.im && j <= g.jm) {
    const int sj = g.sj, sk = g.sk;
    const int dip = (i == g.im) ? 0 : 1, dim_ = (i == 1) ? 0 : -1;
    const int cjp = (j == g.jm) ? (1 - j) * sj : sj, cjm = (j == 1) ? (g.jm - 1) * sj : -sj;
    const double c2l = M1(m.cn2l, i), c2s = M1(m.cn2s, i), c3l = M1(m.cn3l, j), c3s = M1(m.cn3s, j);
    const int t = (i - 1) >> 1;
    int k0, k1;
    kspan(1, g.km, k0, k1);
    for (int k = k0; k <= k1; k++) {
      const int c = ix(g, i, j, k);
      const int dkp = (k == g.km) ? 0 : sk, dkm = (k == 1) ? 0 : -sk;
      float cn;
      if (SEP)
        cn = rcp_rn((M1(m.cnax, i) + M1(m.cnby, j)) + M1(m.cncz, k));  // init.f95:100-107
      else
        cn = cn1f[c];
      const double pc = P64[c];
      const double b = c2l * P64[c + dip] + c2s * P64[c + dim_] + c3l * P64[c + cjp] + c3s * P64[c + cjm] +
                       (double)M1(m.cn4l, k) * P64[c + dkp] + (double)M1(m.cn4s, k) * P64[c + dkm] -
                       pc / (double)cn;
      const double dd = (R64[c] - rav) - b;
      const int col = (i + j + k) & 1;
      const int q = cix(s, t, j, k);
      sb.rhs[col][q] = (float)dd;
      sb.p[0][col][q] = 0.0f;
      const double rd = (double)cn * dd;
      acc += rd * rd;
    }
  }
  const double sum = block_sum<NT>(acc, tid, lds);
  if (tid == 0) part[bid] = sum;
}

// dc_hist[step] = sum (cn1*d)^2 before the outer step's correction; below
// pjuge (or non-finite) the iterate is final and the later steps skip
__global__ __launch_bounds__(256) void k_dc_norm(const double* __restrict__ part, int n, double* __restrict__ hist,
                                                 int step, Ctrl* ctrl, float pjuge) {
  __shared__ double lds[4];
  if (ctrl->dc_done) return;  // uniform
  double s = 0.0;
  for (int b = threadIdx.x; b < n; b += 256) s += part[b];
  s = block_sum<256>(s, threadIdx.x, lds);
  if (threadIdx.x == 0) {
    hist[step] = s;
    ctrl->dc_norms = step + 1;
    if (s < (double)pjuge || !(s == s) || s > 1.0e300) ctrl->dc_done = 1;
  }
}

// P64 += e (the inner solve's last iterate, in iterate set `set`), unless
// the defect met the stop threshold
__global__ __launch_bounds__(NT) void k_dc_correct(Grid g, SplitGrid s, SplitBufs sb, double* __restrict__ P64,
                                                   int set, const Ctrl* ctrl) {
  if (ctrl->dc_done) return;  // uniform
  const int i = 1 + blockIdx.x * BX + threadIdx.x, j = 1 + blockIdx.y * BY + threadIdx.y;
  if (i > g.im || j > g.jm) return;
  const int t = (i - 1) >> 1;
  int k0, k1;
  kspan(1, g.km, k0, k1);
  for (int k = k0; k <= k1; k++) {
    const int col = (i + j + k) & 1;
    P64[ix(g, i, j, k)] += (double)sb.p[set][col][cix(s, t, j, k)];
  }
}

// pav partials of P64 (press.f95:124-133 in FP64)
__global__ __launch_bounds__(NT) void k_dc_pav(Grid g, Metrics m, const double* __restrict__ P64,
                                               double* __restrict__ part) {
  __shared__ double lds[NT / kWave];
  const int tid = threadIdx.y * BX + threadIdx.x;
  const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  const int i = 1 + blockIdx.x * BX + threadIdx.x, j = 1 + blockIdx.y * BY + threadIdx.y;
  double acc = 0.0;
  if (i <= g.im && j <= g.jm) {
    const double dxy = (double)M1(m.dx1, i) * (double)M1(m.dy1, j);
    int k0, k1;
    kspan(1, g.km, k0, k1);
    for (int k = k0; k <= k1; k++) acc += P64[ix(g, i, j, k)] * dxy * (double)M1(m.dzn, k);
  }
  const double sum = block_sum<NT>(acc, tid, lds);
  if (tid == 0) part[bid] = sum;
}

// p = P64 - pav with the ghost cells of boundp1/boundp2 (as k_padj), and
// the defect history into the statistics (sor_iters = defect norms
// recorded: the outer steps taken, plus the final one below the threshold
// when the correction stopped early)
__global__ __launch_bounds__(NT) void k_dc_padj(Grid g, const double* __restrict__ P64, float* __restrict__ p,
                                                Ctrl* ctrl, const double* __restrict__ hist) {
  const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
  if (blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && threadIdx.x == 0 && threadIdx.y == 0) {
    const int nn = ctrl->dc_norms;
    for (int q = 0; q < nn; q++) ctrl->sor_hist[q] = hist[q];
    ctrl->sor_iters = nn;
    ctrl->sor_done = 0;
  }
  if (i > g.im + 1 || j > g.jm + 1) return;
  const double av = ctrl->pav;
  const int ii = (i == 0) ? 1 : ((i == g.im + 1) ? g.im : i);
  const int jj = (j == 0) ? g.jm : ((j == g.jm + 1) ? 1 : j);
  int k0, k1;
  kspan(0, g.km + 1, k0, k1);
  for (int k = k0; k <= k1; k++) {
    const int kk = (k == 0) ? 1 : ((k == g.km + 1) ? g.km : k);
    p[ix(g, i, j, k)] = (float)(P64[ix(g, ii, jj, kk)] - av);
  }
}

// Launch shape of the fused SOR iteration: kernel (rows per thread), tile
// rows and k levels per block; every shape gives the same iterates.
struct SorShape {
  int rows, ty, tw, kc, ntt, ntj, nparts;  // tw: tile width in pairs with the ring (64, or 32 with 32 rows)
};

// rows 4: two iterations per launch (k_sor_pair), 32 x 64 tiles with a
// four-cell ring (callers pass it only for separable cn1).
static SorShape sor_shape(const Grid& g, const SplitGrid& s, int rows, int ty, int kc, int tw = FX) {
  SorShape h;
  h.rows = ((g.jm & 1) != 0) ? 1 : rows;  // odd jm: one row per thread
  if (h.rows == 4) {
    h.ty = PTY;
    h.tw = FX;
    h.kc = std::max(1, std::min(kc, g.km));
    h.ntt = (s.nt <= FX) ? 1 : (s.nt + PFTX - 1) / PFTX;  // one tile without ring lanes when the width fits
    h.ntj = (g.jm + PFTY - 1) / PFTY;
    h.nparts = h.ntt * h.ntj * ((g.km + h.kc - 1) / h.kc);
    return h;
  }
  h.ty = (h.rows == 1) ? FY : ty;          // k_sor_iter has 16-row tiles
  h.tw = (h.rows == 2 && h.ty == 32 && tw == FX / 2) ? FX / 2 : FX;
  h.kc = std::max(1, std::min(kc, g.km));
  // one tile without ring lanes when the width fits
  h.ntt = (s.nt <= h.tw) ? 1 : (s.nt + (h.tw - 4) - 1) / (h.tw - 4);
  h.ntj = (g.jm + (h.ty - 4) - 1) / (h.ty - 4);
  h.nparts = h.ntt * h.ntj * ((g.km + h.kc - 1) / h.kc);
  return h;
}

// The domain's shape: its settings, or the automatic k chunk.  Two
// iterations per launch need separable cn1 (the third iterate set takes the
// storage of the split cn1) and even jm.
static SorShape sor_shape(const DevDomain& d) {
  const int rows = (d.sor_rows == 4 && !d.cn1_separable) ? 2 : d.sor_rows;
  int kc = d.sor_kchunk;
  if (kc <= 0)
    kc = (rows == 4 && (d.g.jm & 1) == 0) ? sor_pair_kchunk(d.g, d.sg, d.ncu)
                                           : sor_fused_kchunk(d.g, d.sg, d.sor_fused_blocks, d.ncu);
  return sor_shape(d.g, d.sg, rows, d.sor_tile_rows, kc, d.sor_tile_width);
}

// The shape of one iteration per launch (press mode 1's inner solves where
// two iterations per launch are not available).
static SorShape sor_shape_single(const DevDomain& d) {
  if (d.sor_rows != 4) return sor_shape(d);
  return sor_shape(d.g, d.sg, 2, d.sor_tile_rows, sor_fused_kchunk(d.g, d.sg, d.sor_fused_blocks, d.ncu),
                   d.sor_tile_width);
}

// One launch of k_sor_pair (two iterations, one, or the fix-up).
static void launch_pair(const DevDomain& d, const SplitBufs& sb, const PairLaunch& pl, const SorShape& h, float pjuge,
                        const float* pent) {
  const size_t lvm_bytes = (size_t)(h.kc + 7) * sizeof(float4);
  const bool ent = (pl.iter == 1 || pl.fixup) && pent;  // may run iteration 1 on the entry ghosts
  const int cls = (pl.iter == 1 || pl.fixup) ? KC_SOR1 : KC_SOR;
  const bool two = pl.two != 0 && !pl.fixup;  // compiled in: no branch on it in the march
#define PAIR1(OM1, ENT, TW)                                                                                     \
  TIMED(d, cls, hipLaunchKernelGGL((k_sor_pair<OM1, ENT, TW>), dim3(h.nparts), dim3(FX, PTY / 2), lvm_bytes,       \
                                   d.stream, d.g, d.sg, d.m, sb, d.ctrl, pl, h.nparts, d.omega, pjuge, h.ntt,      \
                                   h.ntj, h.kc, d.xcd_remap, pent))
#define PAIR(OM1, ENT)       \
  do {                       \
    if (two)                 \
      PAIR1(OM1, ENT, 1);    \
    else                     \
      PAIR1(OM1, ENT, 0);    \
  } while (0)
  if (d.omega == 1.0f) {
    if (ent)
      PAIR(true, true);
    else
      PAIR(true, false);
  } else {
    if (ent)
      PAIR(false, true);
    else
      PAIR(false, false);
  }
#undef PAIR
#undef PAIR1
}

// Iterations l0..l1 in launches of two (the last alone when their number is
// odd), reading iterate set in0 first and rotating over the three sets;
// returns the last launch.  pent: the entry values of p for a launch of
// iteration 1.
static PairLaunch launch_sor_pairs(const DevDomain& d, const SplitBufs& sb, const SorShape& h, float pjuge, int l0,
                                   int l1, int in0, const float* pent) {
  PairLaunch prev{};
  int in = in0, q = 0;
  for (int l = l0; l <= l1; l += 2, q++) {
    PairLaunch pl{};
    pl.iter = l;
    pl.two = (l + 1 <= l1) ? 1 : 0;
    pl.in_set = in;
    pl.out_set = (in + 1) % 3;  // the set before in_set stays intact for a fix-up
    pl.wpart = q & 1;
    pl.prev_iter = prev.iter;
    pl.prev_two = prev.two;
    pl.prev_in = prev.in_set;
    pl.prev_out = prev.out_set;
    pl.prev_nparts = h.nparts;
    pl.rpart = prev.wpart;
    pl.fixup = 0;
    if (l == 1 && pent) {
      // iteration 1: the kernel without the entry-ghost code runs when every
      // ghost of p equals its boundp copy (as after any press), the one with
      // it otherwise; both fill the same partial-sum slots
      pl.gate = 1;
      launch_pair(d, sb, pl, h, pjuge, nullptr);
      pl.gate = 2;
    }
    launch_pair(d, sb, pl, h, pjuge, pent);
    pl.gate = 0;
    prev = pl;
    in = pl.out_set;
  }
  return prev;
}

// One fused SOR iteration l with shape h; nprev is the previous launch's
// number of partial-sum slots; pent != nullptr for iteration 1 of press,
// which reads the entry ghosts of p (k_sor_iter only); gate: with pent, that
// launch runs only when a ghost of p differs from its boundp copy; without
// pent (k_sor_iter2, iteration 1 of press beside the gated entry-ghost
// launch), only when none does.
static void launch_sor_iter(const DevDomain& d, const SplitBufs& sb, int l, float pjuge, const SorShape& h, int nprev,
                            const float* pent = nullptr, int gate = 0) {
  const Grid& g = d.g;
  const SplitGrid& s = d.sg;
  if (h.rows == 1 || pent) {
    const size_t lvm_bytes = 3 * (size_t)(h.kc + 2) * sizeof(float);
#define SOR1(SEP, ENT)                                                                                          \
  TIMED(d, (ENT || l == 1) ? KC_SOR1 : KC_SOR,                                                                              \
        hipLaunchKernelGGL((k_sor_iter<SEP, ENT>), dim3(h.nparts), dim3(FX, FY), lvm_bytes, d.stream, g, s, d.m, \
                           sb, d.ctrl, l, h.nparts, d.omega, pjuge, h.ntt, h.ntj, h.kc, d.xcd_remap, pent, nprev, \
                           gate))
    if (d.cn1_separable) {
      if (pent)
        SOR1(true, true);
      else
        SOR1(true, false);
    } else {
      if (pent)
        SOR1(false, true);
      else
        SOR1(false, false);
    }
#undef SOR1
    return;
  }
  const size_t lvm_bytes = (size_t)(h.kc + 2) * sizeof(float4);
  const int cls = l == 1 ? KC_SOR1 : KC_SOR;  // iteration 1 (entry ghosts or not) is timed apart
#define SOR2(SEP, TY, W)                                                                                     \
  do {                                                                                                       \
    if (d.omega == 1.0f)                                                                                     \
      TIMED(d, cls, hipLaunchKernelGGL((k_sor_iter2<SEP, TY, true, W>), dim3(h.nparts),                   \
                                          dim3(FX, TY * W / (2 * FX)), lvm_bytes, d.stream, g, s, d.m, sb,   \
                                          d.ctrl, l, h.nparts, nprev, d.omega, pjuge, h.ntt, h.ntj, h.kc,    \
                                          d.xcd_remap, gate));                                               \
    else                                                                                                     \
      TIMED(d, cls, hipLaunchKernelGGL((k_sor_iter2<SEP, TY, false, W>), dim3(h.nparts),                  \
                                          dim3(FX, TY * W / (2 * FX)), lvm_bytes, d.stream, g, s, d.m, sb,   \
                                          d.ctrl, l, h.nparts, nprev, d.omega, pjuge, h.ntt, h.ntj, h.kc,    \
                                          d.xcd_remap, gate));                                               \
  } while (0)
  if (d.cn1_separable) {
    if (h.tw == FX / 2)
      SOR2(true, 32, FX / 2);
    else if (h.ty == 32)
      SOR2(true, 32, FX);
    else
      SOR2(true, 16, FX);
  } else {
    if (h.tw == FX / 2)
      SOR2(false, 32, FX / 2);
    else if (h.ty == 32)
      SOR2(false, 32, FX);
    else
      SOR2(false, 16, FX);
  }
#undef SOR2
}

static void launch_press_dc(const DevDomain& d, const SplitBufs& sb, int gx, int gy, int gz, int nred) {
  const Grid& g = d.g;
  const SplitGrid& s = d.sg;
  // the inner iterations two per launch where the domain's SOR runs so
  // (iterate sets rotating over three from set 0, which k_dc_defect clears),
  // else one per launch; they never stop on their own (threshold 0)
  const SorShape hp = sor_shape(d);
  const bool pairs = hp.rows == 4;
  const SorShape h = pairs ? hp : sor_shape_single(d);
  double* P64 = d.dc64;
  double* R64 = d.dc64 + d.n3;
  double* hist = d.dc64 + 2 * (size_t)d.n3;
  DevDomain di = d;  // the inner iterations run with the inner relaxation
  di.omega = d.dc_omega;
  for (int o = 0; o < d.dc_outer; o++) {
    if (d.cn1_separable)
      TIMED(d, KC_RHS, hipLaunchKernelGGL((k_dc_defect<true>), dim3(gx, gy, gz), blk(), 0, d.stream, g, s, d.m,
                                          (const float*)d.F.cn1, (const double*)R64, (const double*)P64, sb,
                                          d.partials, d.ctrl));
    else
      TIMED(d, KC_RHS, hipLaunchKernelGGL((k_dc_defect<false>), dim3(gx, gy, gz), blk(), 0, d.stream, g, s, d.m,
                                          (const float*)d.F.cn1, (const double*)R64, (const double*)P64, sb,
                                          d.partials, d.ctrl));
    TIMED(d, KC_RHS, hipLaunchKernelGGL(k_dc_norm, dim3(1), dim3(256), 0, d.stream, (const double*)d.partials, nred,
                                        hist, o, d.ctrl, d.pjuge));
    int set = d.dc_inner & 1;  // the iterate set holding the inner result
    if (pairs)
      set = launch_sor_pairs(di, sb, h, 0.0f, 1, d.dc_inner, 0, nullptr).out_set;
    else
      for (int l = 1; l <= d.dc_inner; l++) launch_sor_iter(di, sb, l, 0.0f, h, h.nparts);
    TIMED(d, KC_PAV, hipLaunchKernelGGL(k_dc_correct, dim3(gx, gy, g.km), blk(), 0, d.stream, g, s, sb, P64, set,
                                        (const Ctrl*)d.ctrl));
  }
  TIMED(d, KC_PAV, hipLaunchKernelGGL(k_dc_pav, dim3(gx, gy, gz), blk(), 0, d.stream, g, d.m, (const double*)P64,
                                      d.partials));
  TIMED(d, KC_PAV, hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, d.stream, d.partials, nred, d.ctrl, 1,
                                      d.area));
  const int px = (g.im + 2 + BX - 1) / BX, py = (g.jm + 2 + BY - 1) / BY;
  TIMED(d, KC_PADJ, hipLaunchKernelGGL(k_dc_padj, dim3(px, py, g.km + 2), blk(), 0, d.stream, g, (const double*)P64,
                                       d.F.p, d.ctrl, (const double*)hist));
}

// The right-hand side (with bondfg), its mean, and the split arrays the SOR
// works on (press.f95:47-87).
static void press_head(const DevDomain& d, float dt, const SplitBufs& sb, int gx, int gy, int gz, int nred) {
  const Grid& g = d.g;
  const SplitGrid& s = d.sg;
  TIMED(d, KC_RHS, hipLaunchKernelGGL(k_rhs, dim3(gx, gy, gz), blk(), 0, d.stream, g, d.m, d.F.u, d.F.v,
                                      d.F.w, d.F.f, d.F.g, d.F.h, d.F.rhs, d.partials, d.ctrl, dt));
  TIMED(d, KC_RHS, hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, d.stream, d.partials, nred, d.ctrl, 0,
                                      d.area));
  if (d.press_mode == 1) {
    TIMED(d, KC_RHS, hipLaunchKernelGGL(k_dc_init, dim3(gx, gy, gz), blk(), 0, d.stream, g, d.m,
                                        (const float*)d.F.u, (const float*)d.F.v, (const float*)d.F.w,
                                        (const float*)d.F.f, (const float*)d.F.g, (const float*)d.F.h,
                                        (const float*)d.F.p, d.dc64 + d.n3, d.dc64, d.partials, (double)dt,
                                        d.ctrl));
    TIMED(d, KC_RHS, hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, d.stream, d.partials, nred, d.ctrl, 2,
                                        d.area));
  }
  TIMED(d, KC_RHS, hipLaunchKernelGGL(k_rhs_split, dim3(gx, gy, g.km), blk(), 0, d.stream, g, s, d.F.rhs,
                                      (const float*)d.F.p, (const float*)d.F.cn1, sb, d.ctrl));
}

bool sor_batchable(const DevDomain& d) {
  return d.sor_fused && d.press_mode == 0 && d.cn1_separable && (d.g.jm & 1) == 0;
}

// Shapes of a batched domain's SOR: iterations 2.. (h: tile rows ty, k
// chunk kc; kc 0: the domain's own) and iteration 1 (h1, k_sor_iter).
static void batch_shapes(const DevDomain& d, int kc, int ty, int tw, SorShape* h, SorShape* h1) {
  if (kc <= 0) kc = d.sor_kchunk > 0 ? d.sor_kchunk : sor_fused_kchunk(d.g, d.sg, d.sor_fused_blocks, d.ncu);
  *h = sor_shape(d.g, d.sg, 2, ty == 32 ? 32 : FY, kc, tw);
  *h1 = sor_shape(d.g, d.sg, 1, FY, h->kc);
}

// One k chunk for all batched domains: the one that minimises (rounds of
// resident blocks) x (steps per block) over the whole batch, with two
// k_sor_iter2<true, 16> blocks (one <true, 32>) resident per CU.  A domain
// alone picks its chunk to fill the device by itself (a few levels for a
// small domain); in a batch the other domains fill it, and longer chunks
// waste fewer steps on the chunk ends.
int sor_batch_kchunk(const DevDomain* const* ds, int n, int ty, int tw, int ncu) {
  int kmax = 1;
  for (int q = 0; q < n; q++) kmax = std::max(kmax, ds[q]->g.km);
  const long slots = (ty == 32 && tw == FX ? 1L : 2L) * (ncu > 0 ? ncu : 256);
  long best = -1;
  int best_kc = kmax;
  for (int kc = 1; kc <= kmax; kc++) {
    long blocks = 0, steps = 0;
    for (int q = 0; q < n; q++) {
      const SorShape h = sor_shape(ds[q]->g, ds[q]->sg, 2, ty == 32 ? 32 : FY, kc, tw);
      blocks += h.nparts;
      steps = std::max(steps, (long)h.kc + 2);
    }
    const long cost = (blocks + slots - 1) / slots * steps;
    if (best < 0 || cost < best) {
      best = cost;
      best_kc = kc;
    }
  }
  return best_kc;
}

// The batch's k chunk for two iterations per launch: the same cost model
// with k_sor_pair's tiles, one 1024-thread block resident per CU, each block
// marching its chunk and six more levels.
int sor_pair_batch_kchunk(const DevDomain* const* ds, int n, int ncu) {
  int kmax = 1;
  for (int q = 0; q < n; q++) kmax = std::max(kmax, ds[q]->g.km);
  const long slots = ncu > 0 ? ncu : 256;
  long best = -1;
  int best_kc = kmax;
  for (int kc = 1; kc <= kmax; kc++) {
    long blocks = 0;
    for (int q = 0; q < n; q++) blocks += sor_shape(ds[q]->g, ds[q]->sg, 4, PTY, kc).nparts;
    const long cost = (blocks + slots - 1) / slots * (std::min(kc, kmax) + 6);
    if (best < 0 || cost < best) {
      best = cost;
      best_kc = kc;
    }
  }
  return best_kc;
}

// Iterations 2..lmax of the batched domains in launches of two (each domain
// stopping at its own nmaxp), the closing stop test of each domain's last
// launch and the fix-up launch.  Iteration 1 ran in the entry kernel.
void launch_sor_pair_batch(const SorBatchEntry* tab, const int* block0, int ntab, int nblocks, int lmax, int kcmax,
                           bool om1, int remap, hipStream_t stream) {
  const size_t lds = (size_t)(kcmax + 7) * sizeof(float4);
  auto launch = [&](const PairLaunch& pl) {
    if (om1)
      hipLaunchKernelGGL((k_sor_pair_batch<true>), dim3(nblocks), dim3(FX, PTY / 2), lds, stream, tab, block0, ntab,
                         pl, remap);
    else
      hipLaunchKernelGGL((k_sor_pair_batch<false>), dim3(nblocks), dim3(FX, PTY / 2), lds, stream, tab, block0, ntab,
                         pl, remap);
  };
  PairLaunch prev{};
  prev.iter = 1;  // iteration 1: set 1, partials of both colours in slot 1
  prev.two = 0;
  prev.in_set = 0;
  prev.out_set = 1;
  prev.wpart = 1;
  int q = 0;
  for (int l = 2; l <= lmax; l += 2, q++) {
    PairLaunch pl{};
    pl.iter = l;
    pl.two = (l + 1 <= lmax) ? 1 : 0;  // narrowed per domain
    pl.in_set = (1 + q) % 3;
    pl.out_set = (2 + q) % 3;
    pl.wpart = q & 1;
    pl.prev_iter = prev.iter;
    pl.prev_two = prev.two;
    pl.prev_in = prev.in_set;
    pl.prev_out = prev.out_set;
    pl.rpart = prev.wpart;
    pl.prev_colours = (q == 0) ? 1 : 0;
    launch(pl);
    prev = pl;
  }
  hipLaunchKernelGGL(k_sor_pair_batch_final, dim3((unsigned)ntab), dim3(256), 0, stream, tab);
  if (lmax >= 2) {
    PairLaunch fix{};
    fix.fixup = 1;
    launch(fix);
  }
}

void sor_batch_entry(const DevDomain& d, int kc, int ty, int tw, SorBatchEntry* e, bool pairs) {
  SorShape h, h1;
  batch_shapes(d, kc, ty, tw, &h, &h1);
  if (pairs) h = sor_shape(d.g, d.sg, 4, PTY, h.kc);  // the same k chunk as iteration 1's (h1)
  e->g = d.g;
  e->s = d.sg;
  e->m = d.m;
  e->sb = d.sb;
  e->sb.cn1[0] = e->sb.cn1[1] = nullptr;  // separable cn1 (sor_batchable)
  e->ctrl = d.ctrl;
  e->nparts = h.nparts;
  e->nparts1 = h1.nparts;
  e->ntt = h.ntt;
  e->ntj = h.ntj;
  e->kchunk = h.kc;
  e->nmaxp = d.nmaxp;
  e->block0 = 0;
  e->omega = d.omega;
  e->pjuge = d.pjuge;
}

void launch_press_part(const DevDomain& d, float dt, int part, int kc, int ty, int tw, bool pairs) {
  const Grid& g = d.g;
  const int gx = (g.im + BX - 1) / BX, gy = (g.jm + BY - 1) / BY;
  const int gz = kblocks(gx, gy, g.km, kMaxRedBlocks);
  const int nred = gx * gy * gz;
  SplitBufs sb = d.sb;
  sb.cn1[0] = sb.cn1[1] = nullptr;  // separable cn1 (sor_batchable)
  SorShape h, h1;
  batch_shapes(d, kc, ty, tw, &h, &h1);
  if (part == 0) {
    press_head(d, dt, sb, gx, gy, gz, nred);
    launch_sor_iter(d, sb, 1, d.pjuge, h1, h1.nparts, d.F.p);
  } else {
    if (pairs)  // the stop test of the last launch and the fix-up ran batched (launch_sor_pair_batch)
      launch_press_mean(d, sb, gx, gy, gz, nred);
    else
      launch_press_tail(d, sb, d.nmaxp == 1 ? h1.nparts : h.nparts, gx, gy, gz, nred);
  }
}

// ------------------------------------------------ batched press stages
// The press launches of launch_press_part for all domains of a batched step,
// one launch per stage (as the step stages in k_step.hip): each block runs
// the stage kernel's body on its own domain with that domain's grid, block
// index and arguments, so every value and every partial sum is the one of
// the domain's own launch.
#define BATCH_BLOCK(stage, xcd)                                     \
  int e;                                                            \
  VBlk vb;                                                          \
  if (!batch_block<xcd>(sp + (stage) * nd, nd, e, vb)) return;      \
  const BatchDom& B = bd[e];                                        \
  const DevDomain& d = B.d

__global__ __launch_bounds__(NT) void kb_rhs(const BatchDom* __restrict__ bd, const StageSpan* __restrict__ sp,
                                             int nd) {
  BATCH_BLOCK(BS_RHS, false);
  rhs_body(vb, d.g, d.m, d.F.u, d.F.v, d.F.w, d.F.f, d.F.g, d.F.h, d.F.rhs, d.partials, d.ctrl, B.dt);
}
// one block per domain
__global__ __launch_bounds__(256) void kb_finalize(const BatchDom* __restrict__ bd, int what) {
  const BatchDom& B = bd[blockIdx.x];
  finalize_body(B.d.partials, B.nred, B.d.ctrl, what, B.d.area);
}
__global__ __launch_bounds__(NT) void kb_rhs_split(const BatchDom* __restrict__ bd,
                                                   const StageSpan* __restrict__ sp, int nd) {
  BATCH_BLOCK(BS_SPLIT, false);
  rhs_split_body(vb, d.g, d.sg, d.F.rhs, d.F.p, d.F.cn1, B.sb, d.ctrl);
}
__global__ __launch_bounds__(FNT) __attribute__((amdgpu_waves_per_eu(8, 8))) void kb_sor_entry(
    const BatchDom* __restrict__ bd, const StageSpan* __restrict__ sp, int nd) {
  BATCH_BLOCK(BS_SOR1, false);
  sor_iter_block<true, true>(d.g, d.sg, d.m, B.sb, d.ctrl, 1, B.nparts1, d.omega, d.pjuge, B.ntt1, B.ntj1, B.kc1,
                             d.xcd_remap, d.F.p, B.nparts1, 0, vb.x);
}
// one block per domain
__global__ __launch_bounds__(256) void kb_sor_final(const BatchDom* __restrict__ bd) {
  const BatchDom& B = bd[blockIdx.x];
  sor_final_body(B.sb, B.d.ctrl, B.d.nmaxp, B.nparts_last);
}
__global__ __launch_bounds__(NT) void kb_pav(const BatchDom* __restrict__ bd, const StageSpan* __restrict__ sp,
                                             int nd) {
  BATCH_BLOCK(BS_PAV, false);
  pav_body(vb, d.g, d.sg, d.m, B.sb, d.partials, d.ctrl);
}
__global__ __launch_bounds__(NT) void kb_padj(const BatchDom* __restrict__ bd, const StageSpan* __restrict__ sp,
                                              int nd) {
  BATCH_BLOCK(BS_PADJ, false);
  padj_body(vb, d.g, d.sg, B.sb, d.F.p, d.ctrl);
}
#undef BATCH_BLOCK

void press_batch_spans(const DevDomain& d, int kc, int ty, int tw, int nzb, BatchDom* e, StageSpan* sp) {
  const Grid& g = d.g;
  const int gx = (g.im + BX - 1) / BX, gy = (g.jm + BY - 1) / BY;
  const int gz = kblocks(gx, gy, g.km, kMaxRedBlocks);
  e->nred = gx * gy * gz;
  e->sb = d.sb;
  e->sb.cn1[0] = e->sb.cn1[1] = nullptr;  // separable cn1 (sor_batchable)
  SorShape h, h1;
  batch_shapes(d, kc, ty, tw, &h, &h1);
  e->ntt1 = h1.ntt;
  e->ntj1 = h1.ntj;
  e->kc1 = h1.kc;
  e->nparts1 = h1.nparts;
  e->nparts_last = d.nmaxp == 1 ? h1.nparts : h.nparts;
  const int px = (g.im + 2 + BX - 1) / BX, py = (g.jm + 2 + BY - 1) / BY;
  // the reductions keep their own grids (their partial sums are the domain's
  // own); the copies may run longer k runs per block
  sp[BS_RHS] = {0, gx, gy, gz};
  sp[BS_SPLIT] = {0, gx, gy, nzb > 0 ? std::min(nzb, g.km) : g.km};
  sp[BS_SOR1] = {0, h1.nparts, 1, 1};
  sp[BS_PAV] = {0, gx, gy, gz};
  sp[BS_PADJ] = {0, px, py, nzb > 0 ? std::min(nzb, g.km + 2) : g.km + 2};
}

void launch_press_batch(const StageBatch& b, int part, hipStream_t stream, bool pairs) {
  const BatchDom* bd = b.doms;
  const StageSpan* sp = b.spans;
  const int nd = b.nd;
  auto nb = [&](int s) { return dim3((unsigned)b.nblocks[s]); };
  if (part == 0) {
    hipLaunchKernelGGL(kb_rhs, nb(BS_RHS), blk(), 0, stream, bd, sp, nd);
    hipLaunchKernelGGL(kb_finalize, dim3((unsigned)nd), dim3(256), 0, stream, bd, 0);
    hipLaunchKernelGGL(kb_rhs_split, nb(BS_SPLIT), blk(), 0, stream, bd, sp, nd);
    hipLaunchKernelGGL(kb_sor_entry, nb(BS_SOR1), dim3(FX, FY), 3 * (size_t)(b.kc1max + 2) * sizeof(float), stream,
                       bd, sp, nd);
  } else {
    if (!pairs)  // with pairs: launch_sor_pair_batch ran the closing stop test
      hipLaunchKernelGGL(kb_sor_final, dim3((unsigned)nd), dim3(256), 0, stream, bd);
    hipLaunchKernelGGL(kb_pav, nb(BS_PAV), blk(), 0, stream, bd, sp, nd);
    hipLaunchKernelGGL(kb_finalize, dim3((unsigned)nd), dim3(256), 0, stream, bd, 1);
    hipLaunchKernelGGL(kb_padj, nb(BS_PADJ), blk(), 0, stream, bd, sp, nd);
  }
}

void launch_press(const DevDomain& d, float dt) {
  const Grid& g = d.g;
  const SplitGrid& s = d.sg;
  const int gx = (g.im + BX - 1) / BX, gy = (g.jm + BY - 1) / BY;
  const int gz = kblocks(gx, gy, g.km, kMaxRedBlocks);
  const int nred = gx * gy * gz;
  SplitBufs sb = d.sb;
  if (d.cn1_separable) sb.cn1[0] = sb.cn1[1] = nullptr;
  if (d.press_mode == 0 && d.sor_fused && d.nmaxp >= 1 && sor_shape(d).rows == 4) {
    // two iterations per launch: the head in one pass (k_rhs_head), the mean
    // removal over the split rhs alone (k_rhs_sub); the launch of iterations
    // 1 and 2 reads the entry ghosts of p where they differ from their boundp
    // copies
    const SorShape h = sor_shape(d);
    TIMED(d, KC_RHS, hipLaunchKernelGGL(k_rhs_head, dim3(gx, gy, gz), blk(), 0, d.stream, g, s, d.m,
                                        (const float*)d.F.u, (const float*)d.F.v, (const float*)d.F.w, d.F.f,
                                        d.F.g, d.F.h, (const float*)d.F.p, sb, d.partials, nred, d.ctrl, dt));
    TIMED(d, KC_RHS, hipLaunchKernelGGL(k_finalize_head, dim3(1), dim3(256), 0, d.stream,
                                        (const double*)d.partials, nred, d.ctrl, d.area));
    const long n4 = (long)s.cplane * g.km / 2;  // both colours, four floats at a time
    TIMED(d, KC_RHS, hipLaunchKernelGGL(k_rhs_sub, dim3(kRhsSubBlocks), dim3(256), 0, d.stream,
                                        (const float4*)sb.rhsr[0], (float4*)sb.rhs[0], n4, (const Ctrl*)d.ctrl));
    const PairLaunch last = launch_sor_pairs(d, sb, h, d.pjuge, 1, d.nmaxp, 0, d.F.p);
    TIMED(d, KC_PAV, hipLaunchKernelGGL(k_sor_pair_final, dim3(1), dim3(256), 0, d.stream, sb, d.ctrl, last,
                                        h.nparts, d.pjuge));
    PairLaunch fix{};
    fix.fixup = 1;
    fix.wpart = last.wpart ^ 1;
    launch_pair(d, sb, fix, h, d.pjuge, d.F.p);
    launch_press_mean(d, sb, gx, gy, gz, nred);
    return;
  }
  press_head(d, dt, sb, gx, gy, gz, nred);
  if (d.press_mode == 1) {
    launch_press_dc(d, sb, gx, gy, gz, nred);
    return;
  }
  if (d.sor_fused) {
    const SorShape h = sor_shape(d);
    if (h.rows == 4) {  // nmaxp < 1: no SOR launch to remove the mean of rhs
      const PairLaunch last = launch_sor_pairs(d, sb, h, d.pjuge, 1, d.nmaxp, 0, d.F.p);
      TIMED(d, KC_PAV, hipLaunchKernelGGL(k_sor_pair_final, dim3(1), dim3(256), 0, d.stream, sb, d.ctrl, last,
                                          h.nparts, d.pjuge));
      PairLaunch fix{};
      fix.fixup = 1;
      fix.wpart = last.wpart ^ 1;
      launch_pair(d, sb, fix, h, d.pjuge, d.F.p);
      launch_press_mean(d, sb, gx, gy, gz, nred);
      return;
    }
    // iteration 1 reads the ghost values p had on entry (k_sor_iter, 16-row
    // tiles).  When the iteration's own kernel has the same tiles and chunks
    // (so the same partial-sum slots), both are launched: it runs iteration
    // 1 when every ghost equals its boundp copy, as after any press, and the
    // entry-ghost kernel only otherwise (ctrl->p_ghost_diff).
    const SorShape h1 = sor_shape(g, s, 1, FY, h.kc);
    if (h.rows == 2 && h.ty == FY && h.tw == FX && h.nparts == h1.nparts && d.sor_entry_fast) {
      launch_sor_iter(d, sb, 1, d.pjuge, h1, h1.nparts, d.F.p, 1);
      launch_sor_iter(d, sb, 1, d.pjuge, h, h1.nparts, nullptr, 1);
    } else {
      launch_sor_iter(d, sb, 1, d.pjuge, h1, h1.nparts, d.F.p);
    }
    for (int l = 2; l <= d.nmaxp; l++) launch_sor_iter(d, sb, l, d.pjuge, h, l == 2 ? h1.nparts : h.nparts);
    launch_press_tail(d, sb, d.nmaxp == 1 ? h1.nparts : h.nparts, gx, gy, gz, nred);
    return;
  }
  const int tx = (s.nt + BX - 1) / BX;
  const int kc = sor_kchunk(g, s, d.sor_blocks);
  const int tz = (g.km + kc - 1) / kc;
  const int nparts = tx * gy * tz;
  const dim3 sgrid(nparts);
  for (int l = 1; l <= d.nmaxp; l++) {
    if (d.cn1_separable) {
      TIMED(d, KC_SOR, hipLaunchKernelGGL((k_sweep<1, true>), sgrid, blk(), 0, d.stream, g, s, d.m, sb, d.ctrl,
                                          l, nparts, d.omega, d.pjuge, tx, gy, kc, d.xcd_remap,
                                          l == 1 ? (const float*)d.F.p : nullptr));
      TIMED(d, KC_SOR, hipLaunchKernelGGL((k_sweep<0, true>), sgrid, blk(), 0, d.stream, g, s, d.m, sb, d.ctrl,
                                          l, nparts, d.omega, d.pjuge, tx, gy, kc, d.xcd_remap,
                                          l == 1 ? (const float*)d.F.p : nullptr));
    } else {
      TIMED(d, KC_SOR, hipLaunchKernelGGL((k_sweep<1, false>), sgrid, blk(), 0, d.stream, g, s, d.m, sb, d.ctrl,
                                          l, nparts, d.omega, d.pjuge, tx, gy, kc, d.xcd_remap,
                                          l == 1 ? (const float*)d.F.p : nullptr));
      TIMED(d, KC_SOR, hipLaunchKernelGGL((k_sweep<0, false>), sgrid, blk(), 0, d.stream, g, s, d.m, sb, d.ctrl,
                                          l, nparts, d.omega, d.pjuge, tx, gy, kc, d.xcd_remap,
                                          l == 1 ? (const float*)d.F.p : nullptr));
    }
  }
  launch_press_tail(d, sb, nparts, gx, gy, gz, nred);
}

void launch_rhs_gather(const DevDomain& d) {
  const Grid& g = d.g;
  const int gx = (g.im + BX - 1) / BX, gy = (g.jm + BY - 1) / BY;
  hipLaunchKernelGGL(k_rhs_gather, dim3(gx, gy, kblocks(gx, gy, g.km, kMaxRedBlocks)), blk(), 0, d.stream, g, d.sg,
                     d.sb, d.F.rhs);
}

// k levels per SOR block so that the sweep has about `target` blocks
int sor_kchunk(const Grid& g, const SplitGrid& s, int target) {
  const int tx = (s.nt + BX - 1) / BX, gy = (g.jm + BY - 1) / BY;
  int tz = target / (tx * gy);
  if (tz > g.km) tz = g.km;
  if (tz < 1) tz = 1;
  return (g.km + tz - 1) / tz;
}

int sor_parts_needed(const Grid& g, const SplitGrid& s, int target) {
  const int tx = (s.nt + BX - 1) / BX, gy = (g.jm + BY - 1) / BY;
  const int kc = sor_kchunk(g, s, target);
  return tx * gy * ((g.km + kc - 1) / kc);
}

// k levels per block of the fused iteration.  target > 0: about `target`
// blocks.  target <= 0: the split that minimises (rounds of resident blocks)
// x (steps per block), with two 1024-thread blocks resident per CU.
int sor_fused_kchunk(const Grid& g, const SplitGrid& s, int target, int ncu) {
  const int ntt = s.nt <= FX ? 1 : (s.nt + FTX - 1) / FTX, ntj = (g.jm + FTY - 1) / FTY;
  const int tiles = ntt * ntj;
  if (target > 0) {
    int tz = target / tiles;
    if (tz > g.km) tz = g.km;
    if (tz < 1) tz = 1;
    return (g.km + tz - 1) / tz;
  }
  const long slots = 2L * (ncu > 0 ? ncu : 256);
  long best = -1;
  int best_kc = g.km;
  for (int tz = 1; tz <= g.km; tz++) {
    const int kc = (g.km + tz - 1) / tz;
    const int nz = (g.km + kc - 1) / kc;
    const long rounds = ((long)tiles * nz + slots - 1) / slots;
    const long cost = rounds * (kc + 2);
    if (best < 0 || cost < best) {
      best = cost;
      best_kc = kc;
    }
  }
  return best_kc;
}

int sor_fused_parts_needed(const Grid& g, const SplitGrid& s, int target, int ncu) {
  const int ntt = (s.nt + FTX - 1) / FTX, ntj = (g.jm + FTY - 1) / FTY;
  const int kc = sor_fused_kchunk(g, s, target, ncu);
  return ntt * ntj * ((g.km + kc - 1) / kc);
}

int sor_fused_tiles(const Grid& g, const SplitGrid& s) {
  return ((s.nt + FTX - 1) / FTX) * ((g.jm + FTY - 1) / FTY);
}

// `iters` fused SOR iterations (numbered 2.., never stopping) with the given
// rows per thread, tile rows and k chunk, on whatever the split arrays hold:
// a timing trial for the autotuner.  Writes only the split iterates, the SOR
// partial slots and the SOR fields of the control block.
void launch_sor_trial(const DevDomain& d0, int rows, int ty, int tw, int kc, int iters) {
  DevDomain d = d0;
  d.timer = nullptr;
  SplitBufs sb = d.sb;
  if (d.cn1_separable) sb.cn1[0] = sb.cn1[1] = nullptr;
  const SorShape h = sor_shape(d.g, d.sg, (rows == 4 && !d.cn1_separable) ? 2 : rows, ty, kc, tw);
  if (h.rows == 4) {  // iterations 2 .. iters+1 in launches of two (an even count: whole launches)
    launch_sor_pairs(d, sb, h, 0.0f, 2, 1 + (iters + 1) / 2 * 2, 1, nullptr);
    return;
  }
  for (int l = 2; l < 2 + iters; l++) launch_sor_iter(d, sb, l, 0.0f, h, h.nparts);
}

// k levels per block of the two-iteration launch: the split that minimises
// (rounds of resident blocks) x (steps per block), one 1024-thread block per
// CU, each block marching its chunk and six more levels.
int sor_pair_kchunk(const Grid& g, const SplitGrid& s, int ncu) {
  const int tiles = (s.nt <= FX ? 1 : (s.nt + PFTX - 1) / PFTX) * ((g.jm + PFTY - 1) / PFTY);
  const long slots = ncu > 0 ? ncu : 256;
  long best = -1;
  int best_kc = g.km;
  for (int tz = 1; tz <= g.km; tz++) {
    const int kc = (g.km + tz - 1) / tz;
    const int nz = (g.km + kc - 1) / kc;
    const long rounds = ((long)tiles * nz + slots - 1) / slots;
    const long cost = rounds * (kc + 6);
    if (best < 0 || cost < best) {
      best = cost;
      best_kc = kc;
    }
  }
  return best_kc;
}

int sor_pair_tiles(const Grid& g, const SplitGrid& s) {
  return ((s.nt + PFTX - 1) / PFTX) * ((g.jm + PFTY - 1) / PFTY);
}

int reduction_parts_needed(const Grid& g) {
  const int gx = (g.im + BX - 1) / BX, gy = (g.jm + BY - 1) / BY;
  return gx * gy * kblocks(gx, gy, g.km, kMaxRedBlocks);
}

}  // namespace les
