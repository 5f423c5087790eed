// Pressure projection of the LES step: bondfg + rhs with the volume-weighted
// mean removed, red-black SOR with its boundp ghost updates, and the
// pressure-mean removal (press.f95:47-162, bondFG.f95:18-42,
// boundp.f95:7-78).  Stencil arithmetic is FP32 in the reference's order;
// the three reductions (mean of rhs, sum of reltmp^2, mean of p) accumulate
// FP32 terms in FP64, deterministically (fixed block partition, fixed
// summation order).  The SOR stop decision is taken on the device: every
// SOR kernel first checks whether an earlier iteration already converged.
#include "les_kernels.h"

namespace les {

#define BX 64
#define BY 4
#define NT (BX * BY)

// Number of k-blocks for a reduction launch: at most npartials blocks in
// total; each block's k loop covers the remaining levels.
static inline int kblocks(const DevDomain& d, int gx, int gy, int nk) {
  int gz = d.npartials / (gx * gy);
  if (gz > nk) gz = nk;
  if (gz < 1) gz = 1;
  return gz;
}

// bondfg (bondFG.f95:18-42) folded into the divergence: f(0)=f(1),
// g(i,0)=g(i,jm), h(k=0)=h(km)=0; the ghosts are also stored so the arrays
// end as the reference leaves them.  rhs = div(F) + div(u)/dt
// (press.f95:49-60); block partials of dx1*dy1*dzn*rhs (press.f95:63-72).
__global__ __launch_bounds__(NT) void k_rhs(Grid g, Metrics m, const float* __restrict__ u,
                                            const float* __restrict__ v, const float* __restrict__ w,
                                            float* __restrict__ f, float* __restrict__ gg,
                                            float* __restrict__ h, float* __restrict__ rhs,
                                            double* partials, Ctrl* ctrl, float dt, double area) {
  __shared__ double lds[NT / kWave];
  const int tid = threadIdx.y * BX + threadIdx.x;
  const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  const int nblocks = gridDim.x * gridDim.y * gridDim.z;
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
  double total;
  if (grid_sum<NT>(s, tid, bid, nblocks, partials, &ctrl->ticket, &total, lds)) ctrl->rhsav = total / area;
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

// One colour of the SOR sweep (press.f95:92-104): nrd=0 updates the cells
// with i+j+k odd, nrd=1 the others.  Ghost values are read from memory as
// the last boundp call left them.
__global__ __launch_bounds__(NT) void k_sor(Grid g, Metrics m, const float* __restrict__ cn1,
                                            const float* __restrict__ rhs, float* __restrict__ p,
                                            double* partials, Ctrl* ctrl, int nrd, int iter,
                                            float omega, float pjuge) {
  __shared__ double lds[NT / kWave];
  const int done = ctrl->sor_done;
  if (done != 0 && done < iter) return;
  const int tid = threadIdx.y * BX + threadIdx.x;
  const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  const int nblocks = gridDim.x * gridDim.y * gridDim.z;
  const int t = blockIdx.x * BX + threadIdx.x, j = 1 + blockIdx.y * BY + threadIdx.y;
  double acc = 0.0;
  if (j <= g.jm) {
    const float c3l = M1(m.cn3l, j), c3s = M1(m.cn3s, j);
    for (int k = 1 + blockIdx.z; k <= g.km; k += gridDim.z) {
      const int i = 1 + ((k + j + nrd) & 1) + 2 * t;
      if (i > g.im) continue;
      const int c = ix(g, i, j, k);
      const float pc = p[c];
      const float reltmp =
          omega * (cn1[c] * (M1(m.cn2l, i) * p[c + 1] + M1(m.cn2s, i) * p[c - 1] + c3l * p[c + g.sj] +
                             c3s * p[c - g.sj] + M1(m.cn4l, k) * p[c + g.sk] + M1(m.cn4s, k) * p[c - g.sk] -
                             rhs[c]) -
                   pc);
      p[c] = pc + reltmp;
      acc += (double)reltmp * (double)reltmp;
    }
  }
  const double s = block_sum<NT>(acc, tid, lds);
  double total;
  if (grid_sum<NT>(s, tid, bid, nblocks, partials, &ctrl->ticket, &total, lds)) {
    if (nrd == 0) {
      ctrl->sor_first = total;
    } else {
      const double sor = ctrl->sor_first + total;
      if (iter <= kSorHist) ctrl->sor_hist[iter - 1] = sor;
      ctrl->sor_iters = iter;
      if (!(sor == sor) || sor > 1.0e300) ctrl->nonfinite = 1;
      if (sor < (double)pjuge || ctrl->nonfinite) ctrl->sor_done = iter;  // press.f95:119-121
    }
  }
}

// boundp1 (boundp.f95:44-69: i Neumann for j=0..jm+1, k=0..km+1, then
// periodic j for i=0..im+1, k=0..km+1) and, with with_k, boundp2 after it
// (boundp.f95:16-21: k Neumann for i=0..im+1, j=0..jm+1).  Each written
// ghost gets its composite value, which comes from a cell none of the
// copies writes.  iter>0: skipped once an earlier iteration converged.
__global__ __launch_bounds__(256) void k_boundp(Grid g, float* __restrict__ p, const Ctrl* ctrl,
                                                int with_k, int iter) {
  const int done = ctrl->sor_done;
  if (iter > 0 && done != 0 && done < iter) return;
  const int ni = g.im + 2, nj = g.jm + 2;
  const int nK = with_k ? 2 * ni * nj : 0;
  const int kin0 = with_k ? 1 : 0, kin1 = with_k ? g.km : g.km + 1;  // k rows handled by the side faces
  const int nkk = kin1 - kin0 + 1;
  const int nJ = 2 * ni * nkk, nI = 2 * g.jm * nkk;
  int t = blockIdx.x * 256 + threadIdx.x;
  int i, j, k;
  if (t < nK) {
    i = t % ni;
    j = (t / ni) % nj;
    k = (t / (ni * nj)) ? g.km + 1 : 0;
  } else if ((t -= nK) < nJ) {
    i = t % ni;
    j = ((t / ni) & 1) ? g.jm + 1 : 0;
    k = kin0 + t / (ni * 2);
  } else if ((t -= nJ) < nI) {
    i = (t & 1) ? g.im + 1 : 0;
    j = 1 + (t >> 1) % g.jm;
    k = kin0 + (t >> 1) / g.jm;
  } else {
    return;
  }
  int ii = i, jj = j, kk = k;
  if (i == 0) ii = 1;
  else if (i == g.im + 1) ii = g.im;
  if (j == 0) jj = g.jm;
  else if (j == g.jm + 1) jj = 1;
  if (with_k) {
    if (k == 0) kk = 1;
    else if (k == g.km + 1) kk = g.km;
  }
  p[ix(g, i, j, k)] = p[ix(g, ii, jj, kk)];
}

// pav partials: p*dx1*dy1*dzn (press.f95:124-133).
__global__ __launch_bounds__(NT) void k_pav(Grid g, Metrics m, const float* __restrict__ p,
                                            double* partials, Ctrl* ctrl, double pco) {
  __shared__ double lds[NT / kWave];
  const int tid = threadIdx.y * BX + threadIdx.x;
  const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  const int nblocks = gridDim.x * gridDim.y * gridDim.z;
  const int i = 1 + blockIdx.x * BX + threadIdx.x, j = 1 + blockIdx.y * BY + threadIdx.y;
  double acc = 0.0;
  if (i <= g.im && j <= g.jm) {
    const float dxi = M1(m.dx1, i), dyj = M1(m.dy1, j);
    for (int k = 1 + blockIdx.z; k <= g.km; k += gridDim.z)
      acc += (double)(p[ix(g, i, j, k)] * dxi * dyj * M1(m.dzn, k));
  }
  const double s = block_sum<NT>(acc, tid, lds);
  double total;
  if (grid_sum<NT>(s, tid, bid, nblocks, partials, &ctrl->ticket, &total, lds)) ctrl->pav = total / pco;
}

// p = p - pav on the interior (press.f95:145-150), in FP64.
__global__ __launch_bounds__(NT) void k_padj(Grid g, float* __restrict__ p, const Ctrl* ctrl) {
  const int i = 1 + blockIdx.x * BX + threadIdx.x, j = 1 + blockIdx.y * BY + threadIdx.y;
  if (i > g.im || j > g.jm) return;
  const double av = ctrl->pav;
  for (int k = 1 + blockIdx.z; k <= g.km; k += gridDim.z) {
    const int c = ix(g, i, j, k);
    p[c] = (float)((double)p[c] - av);
  }
}

static inline dim3 blk() { return dim3(BX, BY, 1); }

static void launch_boundp(const DevDomain& d, int with_k, int iter) {
  const Grid& g = d.g;
  const int ni = g.im + 2, nj = g.jm + 2;
  const int nkk = with_k ? g.km : g.km + 2;
  const int n = (with_k ? 2 * ni * nj : 0) + 2 * ni * nkk + 2 * g.jm * nkk;
  hipLaunchKernelGGL(k_boundp, dim3((n + 255) / 256), dim3(256), 0, d.stream, g, d.F.p,
                     (const Ctrl*)d.ctrl, with_k, iter);
}

void launch_press(const DevDomain& d, float dt) {
  const Grid& g = d.g;
  const int gx = (g.im + BX - 1) / BX, gy = (g.jm + BY - 1) / BY;
  const int gz = kblocks(d, gx, gy, g.km);
  hipLaunchKernelGGL(k_rhs, dim3(gx, gy, gz), blk(), 0, d.stream, g, d.m, d.F.u, d.F.v, d.F.w, d.F.f,
                     d.F.g, d.F.h, d.F.rhs, d.partials, d.ctrl, dt, d.area);
  hipLaunchKernelGGL(k_rhs_sub, dim3(gx, gy, g.km), blk(), 0, d.stream, g, d.F.rhs,
                     (const Ctrl*)d.ctrl);
  const int half = (g.im + 1) / 2 + 1;
  const int sx = (half + BX - 1) / BX;
  const int sz = kblocks(d, sx, gy, g.km);
  for (int l = 1; l <= d.nmaxp; l++) {
    hipLaunchKernelGGL(k_sor, dim3(sx, gy, sz), blk(), 0, d.stream, g, d.m, d.F.cn1, d.F.rhs, d.F.p,
                       d.partials, d.ctrl, 0, l, d.omega, d.pjuge);
    launch_boundp(d, 0, l);
    hipLaunchKernelGGL(k_sor, dim3(sx, gy, sz), blk(), 0, d.stream, g, d.m, d.F.cn1, d.F.rhs, d.F.p,
                       d.partials, d.ctrl, 1, l, d.omega, d.pjuge);
    launch_boundp(d, 1, l);
  }
  hipLaunchKernelGGL(k_pav, dim3(gx, gy, gz), blk(), 0, d.stream, g, d.m, d.F.p, d.partials, d.ctrl,
                     d.area);
  hipLaunchKernelGGL(k_padj, dim3(gx, gy, g.km), blk(), 0, d.stream, g, d.F.p, (const Ctrl*)d.ctrl);
  launch_boundp(d, 1, 0);
}

}  // namespace les
