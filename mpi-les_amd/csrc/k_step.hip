// Stage kernels of the LES step (velnw, bondv1, velfg+vel2, feedbf, les,
// adam) and the layout-conversion kernels.  Expressions are evaluated in the
// order of the cited Fortran lines and the library is built without FMA
// contraction, so each FP32 result equals the reference's.
#include <algorithm>
#include <type_traits>

#include "les_kernels.h"

namespace les {

#define BX 64
#define BY 4
#define NT (BX * BY)

static inline dim3 block2d() { return dim3(BX, BY, 1); }
static inline dim3 grid2d(int ni, int nj, int nk) {
  return dim3((unsigned)((ni + BX - 1) / BX), (unsigned)((nj + BY - 1) / BY), (unsigned)nk);
}
// blocks along k of an (i,j)-tiled stencil launch: one level per block
// (runs of consecutive levels per thread measured slower here -- the
// level-k+1 neighbours are found in L2/MALL either way, and one level per
// block gives the most independent waves)
static inline int kz(const Grid& g, int extra = 0) { return g.km + extra; }

// ------------------------------------------------------------------ velnw
// velnw.f95:27-59: u,v for k=1..km, w for k=1..km-1.
__device__ __forceinline__ void velnw_body(const VBlk& vb, const Grid& g, const Metrics& m,
                                           const float* __restrict__ p, const float* __restrict__ f,
                                           const float* __restrict__ gg, const float* __restrict__ h,
                                           float* __restrict__ u, float* __restrict__ v, float* __restrict__ w,
                                           float ro, float dt) {
  const int i = 1 + vb.x * BX + threadIdx.x, j = 1 + vb.y * BY + wave_row();
  if (i > g.im || j > g.jm) return;
  const float rdx = M1(m.dxs, i), rdy = M1(m.dys, j);
  int k0, k1;
  kspan(vb, 1, g.km, k0, k1);
  for (int k = k0; k <= k1; k++) {
    const int c = ix(g, i, j, k);
    const float pc = p[c];
    float pz = (-pc + p[c + 1]) / ro / rdx;
    u[c] = u[c] + dt * (f[c] - pz);
    pz = (-pc + p[c + g.sj]) / ro / rdy;
    v[c] = v[c] + dt * (gg[c] - pz);
    if (k <= g.km - 1) {
      pz = (-pc + p[c + g.sk]) / ro / M1(m.dzs, k);
      w[c] = w[c] + dt * (h[c] - pz);
    }
  }
}

__global__ __launch_bounds__(NT) void k_velnw(Grid g, Metrics m, const float* __restrict__ p,
                                              const float* __restrict__ f, const float* __restrict__ gg,
                                              const float* __restrict__ h, float* __restrict__ u,
                                              float* __restrict__ v, float* __restrict__ w, float ro, float dt) {
  velnw_body(hw_blk<true>(), g, m, p, f, gg, h, u, v, w, ro, dt);
}

void launch_velnw(const DevDomain& d, float ro, float dt) {
  const Grid& g = d.g;
  TIMED(d, KC_VELNW, hipLaunchKernelGGL(k_velnw, grid2d(g.im, g.jm, kz(g)), block2d(), 0, d.stream, g, d.m, d.F.p,
                     d.F.f, d.F.g, d.F.h, d.F.u, d.F.v, d.F.w, ro, dt));
}

// ----------------------------------------------------------------- bondv1
// Inflow planes i=0,1 (bondv1.f95:33-54) and, at n==1, the copy of column
// i=1 over i=2..im (bondv1.f95:61-88): u=u_in(j,k), v=v_in, w=w_in (zero
// unless set) for i=0..ilast.
// Also clears the outflow max/min accumulators.
__device__ __forceinline__ void bondv1_inflow_body(const VBlk& vb, const Grid& g, const Metrics& m,
                                                   float* __restrict__ u, float* __restrict__ v,
                                                   float* __restrict__ w, int ilast, Ctrl* ctrl) {
  if (vb.x == 0 && vb.y == 0 && vb.z == 0 && threadIdx.x == 0 && threadIdx.y == 0) {
    ctrl->uout_max = 0u;
    ctrl->uout_nmin = 0u;
  }
  if (ilast == 1) {  // steps n > 1: the planes i = 0, 1, one thread per (j, k) (grid of bondv1_grid)
    const int j = 1 + vb.x * BX + threadIdx.x, k = 1 + vb.y * BY + threadIdx.y;
    if (j > g.jm || k > g.km) return;
    const int q = (j - 1) + (k - 1) * g.jm;
    for (int i = 0; i <= 1; i++) {
      const int c = ix(g, i, j, k);
      u[c] = m.inflow ? m.inflow[q] : 5.0f;
      v[c] = m.inflow_v ? m.inflow_v[q] : 0.0f;
      w[c] = m.inflow_w ? m.inflow_w[q] : 0.0f;
    }
    return;
  }
  const int i = vb.x * BX + threadIdx.x, j = 1 + vb.y * BY + threadIdx.y;
  const int k = 1 + vb.z;
  if (i > ilast || j > g.jm) return;
  const int q = (j - 1) + (k - 1) * g.jm;
  const int c = ix(g, i, j, k);
  u[c] = m.inflow ? m.inflow[q] : 5.0f;
  v[c] = m.inflow_v ? m.inflow_v[q] : 0.0f;  // bondv1_data24.f95:79-87 sets v and w too
  w[c] = m.inflow_w ? m.inflow_w[q] : 0.0f;
}
__global__ __launch_bounds__(NT) void k_bondv1_inflow(Grid g, Metrics m, float* __restrict__ u,
                                                      float* __restrict__ v, float* __restrict__ w,
                                                      int ilast, Ctrl* ctrl) {
  bondv1_inflow_body(hw_blk<false>(), g, m, u, v, w, ilast, ctrl);
}

// aaa = max(0, max u(im,j,k)), bbb = min(0, min u(im,j,k)) (bondv1.f95:98-114).
// Max/min are exact in any order; non-negative floats order like their bits.
__device__ __forceinline__ void bondv1_uout_body(const VBlk& vb, const Grid& g, const float* __restrict__ u,
                                                 Ctrl* ctrl) {
  __shared__ float smax[4], smin[4];
  float mx = 0.0f, mn = 0.0f;
  const int n = g.jm * g.km;
  for (int t = vb.x * 256 + threadIdx.x; t < n; t += vb.nx * 256) {
    const int j = 1 + t % g.jm, k = 1 + t / g.jm;
    const float x = u[ix(g, g.im, j, k)];
    mx = (x > mx) ? x : mx;
    mn = (x < mn) ? x : mn;
  }
  for (int o = kWave / 2; o > 0; o >>= 1) {
    const float a = __shfl_xor(mx, o, kWave), b = __shfl_xor(mn, o, kWave);
    mx = (a > mx) ? a : mx;
    mn = (b < mn) ? b : mn;
  }
  if (threadIdx.x % kWave == 0) {
    smax[threadIdx.x / kWave] = mx;
    smin[threadIdx.x / kWave] = mn;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < 4; q++) {
      mx = (smax[q] > mx) ? smax[q] : mx;
      mn = (smin[q] < mn) ? smin[q] : mn;
    }
    atomicMax(&ctrl->uout_max, __float_as_uint(mx));
    atomicMax(&ctrl->uout_nmin, __float_as_uint(0.0f - mn));
  }
}
__global__ __launch_bounds__(256) void k_bondv1_uout(Grid g, const float* __restrict__ u, Ctrl* ctrl) {
  bondv1_uout_body(hw_blk<false>(), g, u, ctrl);
}

// Advective outflow at i=im (u) and i=im+1 (v,w), bondv1.f95:116-134.
__device__ __forceinline__ void bondv1_outflow_body(const VBlk& vb, const Grid& g, const Metrics& m,
                                                    float* __restrict__ u, float* __restrict__ v,
                                                    float* __restrict__ w, const Ctrl* ctrl, float dt) {
  const int t = vb.x * 256 + threadIdx.x;
  if (t >= g.jm * g.km) return;
  const int j = 1 + t % g.jm, k = 1 + t / g.jm;
  const float aaa = __uint_as_float(ctrl->uout_max);
  const float bbb = -__uint_as_float(ctrl->uout_nmin);
  const float uout = (aaa + bbb) / 2.0f;
  const float dxs = M1(m.dxs, g.im);
  int c = ix(g, g.im, j, k);
  u[c] = u[c] - dt * uout * (u[c] - u[c - 1]) / dxs;
  c = ix(g, g.im + 1, j, k);
  v[c] = v[c] - dt * uout * (v[c] - v[c - 1]) / dxs;
  w[c] = w[c] - dt * uout * (w[c] - w[c - 1]) / dxs;
}
__global__ __launch_bounds__(256) void k_bondv1_outflow(Grid g, Metrics m, float* __restrict__ u,
                                                        float* __restrict__ v, float* __restrict__ w,
                                                        const Ctrl* ctrl, float dt) {
  bondv1_outflow_body(hw_blk<false>(), g, m, u, v, w, ctrl, dt);
}

// Periodic j ghosts (bondv1.f95:137-156) followed by the ground/top rows
// (bondv1.f95:176-195), written as their final values:
//   part A: j=0 and j=jm+1 for k=1..km (w: k=1..km-1), i=0..im+1
//   part B: k=0 and k=km+1 rows of u,v for j=0..jm+1 (periodic image of
//           j folded in), w(k=0)=w(k=km)=0 for j=-1..jm+1.
// Every value read is an interior j row, none of which is written here.
__device__ __forceinline__ void bondv1_ghost_body(const VBlk& vb, const Grid& g, float* __restrict__ u,
                                                  float* __restrict__ v, float* __restrict__ w) {
  const int ni = g.im + 2;
  const int nA = ni * g.km, nB = ni * (g.jm + 3);
  int t = vb.x * 256 + threadIdx.x;
  if (t < nA) {
    const int i = t % ni, k = 1 + t / ni;
    const int c0 = ix(g, i, 0, k), cjm = ix(g, i, g.jm, k), c1 = ix(g, i, 1, k), cj1 = ix(g, i, g.jm + 1, k);
    u[c0] = u[cjm];
    u[cj1] = u[c1];
    v[c0] = v[cjm];
    v[cj1] = v[c1];
    if (k <= g.km - 1) {
      w[c0] = w[cjm];
      w[cj1] = w[c1];
    }
    return;
  }
  t -= nA;
  if (t >= nB) return;
  const int i = t % ni, j = -1 + t / ni;
  if (j >= 0) {
    const int cj = (j == 0) ? g.jm : ((j == g.jm + 1) ? 1 : j);
    u[ix(g, i, j, 0)] = -u[ix(g, i, cj, 1)];
    u[ix(g, i, j, g.km + 1)] = u[ix(g, i, cj, g.km)];
    v[ix(g, i, j, 0)] = -v[ix(g, i, cj, 1)];
    v[ix(g, i, j, g.km + 1)] = v[ix(g, i, cj, g.km)];
  }
  w[ix(g, i, j, 0)] = 0.0f;
  w[ix(g, i, j, g.km)] = 0.0f;
}
__global__ __launch_bounds__(256) void k_bondv1_ghost(Grid g, float* __restrict__ u, float* __restrict__ v,
                                                      float* __restrict__ w) {
  bondv1_ghost_body(hw_blk<false>(), g, u, v, w);
}

// Grids of the four bondv1 launches of step 1 and of the batched steps
// (one-dimensional but the first, which covers (i, j, k) at n == 1 and
// (j, k) for the two inflow planes after).
static dim3 bondv1_grid(const Grid& g, int ilast, int part) {
  const int nplane = g.jm * g.km;
  switch (part) {
    case 0: return ilast == 1 ? grid2d(g.jm, g.km, 1) : grid2d(ilast + 1, g.jm, g.km);
    case 1: return dim3((unsigned)std::min((nplane + 255) / 256, 256));
    case 2: return dim3((unsigned)((nplane + 255) / 256));
    default: return dim3((unsigned)(((g.im + 2) * (g.km + g.jm + 3) + 255) / 256));
  }
}

// Steps n > 1 in two launches instead of four.  The inflow planes i = 0, 1
// and the outflow column i = im (im >= 2) do not overlap, so
//   k_bondv1_in_uout writes the inflow planes (one thread per (j,k)) and, in
//     the blocks after them, the block max/min of u(im,:,:) into the
//     domain's partial slots (block b: max at b, min at 256+b, at most 256
//     blocks);
//   k_bondv1_out_ghost folds those (max and min are exact in any order) in
//     each of its outflow blocks, applies the outflow with the j and k ghost
//     copies of every cell it writes, and in the blocks after them writes
//     the ghost rows of bondv1_ghost_body for every other cell.
// Each cell gets the same expression on the same operands as in the four
// launches, and its final value: w(im+1,j,km), which the outflow computes
// and the top row then zeroes, is left to the zeroing.
__global__ __launch_bounds__(256) void k_bondv1_in_uout(Grid g, Metrics m, float* __restrict__ u,
                                                       float* __restrict__ v, float* __restrict__ w,
                                                       double* __restrict__ part, int nin) {
  const int b = blockIdx.x, tid = threadIdx.x;
  const int n = g.jm * g.km;
  if (b < nin) {  // bondv1.f95:33-54
    const int t = b * 256 + tid;
    if (t >= n) return;
    const int j = 1 + t % g.jm, k = 1 + t / g.jm;
    const int q = (j - 1) + (k - 1) * g.jm;
    for (int i = 0; i <= 1; i++) {
      const int c = ix(g, i, j, k);
      u[c] = m.inflow ? m.inflow[q] : 5.0f;
      v[c] = m.inflow_v ? m.inflow_v[q] : 0.0f;
      w[c] = m.inflow_w ? m.inflow_w[q] : 0.0f;
    }
    return;
  }
  // bondv1.f95:98-114, as bondv1_uout_body, one partial per block
  const int bu = b - nin, nu = (int)gridDim.x - nin;
  __shared__ float smax[4], smin[4];
  float mx = 0.0f, mn = 0.0f;
  for (int t = bu * 256 + tid; t < n; t += nu * 256) {
    const int j = 1 + t % g.jm, k = 1 + t / g.jm;
    const float x = u[ix(g, g.im, j, k)];
    mx = (x > mx) ? x : mx;
    mn = (x < mn) ? x : mn;
  }
  for (int o = kWave / 2; o > 0; o >>= 1) {
    const float a = __shfl_xor(mx, o, kWave), c = __shfl_xor(mn, o, kWave);
    mx = (a > mx) ? a : mx;
    mn = (c < mn) ? c : mn;
  }
  if (tid % kWave == 0) {
    smax[tid / kWave] = mx;
    smin[tid / kWave] = mn;
  }
  __syncthreads();
  if (tid == 0) {
    for (int q = 1; q < 4; q++) {
      mx = (smax[q] > mx) ? smax[q] : mx;
      mn = (smin[q] < mn) ? smin[q] : mn;
    }
    part[bu] = (double)mx;  // exact both ways
    part[256 + bu] = (double)mn;
  }
}

__global__ __launch_bounds__(256) void k_bondv1_out_ghost(Grid g, Metrics m, float* __restrict__ u,
                                                         float* __restrict__ v, float* __restrict__ w,
                                                         const double* __restrict__ part, int nu, int nout,
                                                         float dt) {
  const int b = blockIdx.x, tid = threadIdx.x;
  const int im = g.im, jm = g.jm, km = g.km;
  if (b < nout) {  // bondv1.f95:116-134 with the copies of bondv1.f95:137-156, 176-195
    __shared__ float smax[4], smin[4];
    float mx = 0.0f, mn = 0.0f;
    if (tid < nu) {
      mx = (float)part[tid];
      mn = (float)part[256 + tid];
    }
    for (int o = kWave / 2; o > 0; o >>= 1) {
      const float a = __shfl_xor(mx, o, kWave), c = __shfl_xor(mn, o, kWave);
      mx = (a > mx) ? a : mx;
      mn = (c < mn) ? c : mn;
    }
    if (tid % kWave == 0) {
      smax[tid / kWave] = mx;
      smin[tid / kWave] = mn;
    }
    __syncthreads();
    for (int q = 0; q < 4; q++) {
      mx = (smax[q] > mx) ? smax[q] : mx;
      mn = (smin[q] < mn) ? smin[q] : mn;
    }
    const int t = b * 256 + tid;
    if (t >= jm * km) return;
    const int j = 1 + t % jm, k = 1 + t / jm;
    const float aaa = mx;
    const float bbb = -(0.0f - mn);  // the sign-bit round trip of bondv1_uout_body's accumulator
    const float uout = (aaa + bbb) / 2.0f;
    const float dxs = M1(m.dxs, im);
    int c = ix(g, im, j, k);
    const float nu_ = u[c] - dt * uout * (u[c] - u[c - 1]) / dxs;
    c = ix(g, im + 1, j, k);
    const float nv = v[c] - dt * uout * (v[c] - v[c - 1]) / dxs;
    float nw = 0.0f;
    if (k <= km - 1) nw = w[c] - dt * uout * (w[c] - w[c - 1]) / dxs;
    // the rows holding this cell's value: its own, and j = 0 (the periodic
    // copy of jm) or j = jm+1 (of 1)
    const int jr1 = (j == jm) ? 0 : ((j == 1) ? jm + 1 : -2);
    for (int r = 0; r < 2; r++) {
      const int jj = r == 0 ? j : jr1;
      if (jj < 0) continue;
      u[ix(g, im, jj, k)] = nu_;
      v[ix(g, im + 1, jj, k)] = nv;
      if (k <= km - 1) w[ix(g, im + 1, jj, k)] = nw;
      if (k == 1) {
        u[ix(g, im, jj, 0)] = -nu_;
        v[ix(g, im + 1, jj, 0)] = -nv;
      }
      if (k == km) {
        u[ix(g, im, jj, km + 1)] = nu_;
        v[ix(g, im + 1, jj, km + 1)] = nv;
      }
    }
    return;
  }
  // bondv1_ghost_body for the cells the outflow blocks do not write: u at
  // i = im and v, w at i = im+1 are theirs (w's top row excepted)
  const int ni = im + 2;
  const int nA = ni * km, nB = ni * (jm + 3);
  int t = (b - nout) * 256 + tid;
  if (t < nA) {
    const int i = t % ni, k = 1 + t / ni;
    const int c0 = ix(g, i, 0, k), cjm = ix(g, i, jm, k), c1 = ix(g, i, 1, k), cj1 = ix(g, i, jm + 1, k);
    if (i != im) {
      u[c0] = u[cjm];
      u[cj1] = u[c1];
    }
    if (i != im + 1) {
      v[c0] = v[cjm];
      v[cj1] = v[c1];
      if (k <= km - 1) {
        w[c0] = w[cjm];
        w[cj1] = w[c1];
      }
    }
    return;
  }
  t -= nA;
  if (t >= nB) return;
  const int i = t % ni, j = -1 + t / ni;
  if (j >= 0) {
    const int cj = (j == 0) ? jm : ((j == jm + 1) ? 1 : j);
    if (i != im) {
      u[ix(g, i, j, 0)] = -u[ix(g, i, cj, 1)];
      u[ix(g, i, j, km + 1)] = u[ix(g, i, cj, km)];
    }
    if (i != im + 1) {
      v[ix(g, i, j, 0)] = -v[ix(g, i, cj, 1)];
      v[ix(g, i, j, km + 1)] = v[ix(g, i, cj, km)];
    }
  }
  w[ix(g, i, j, 0)] = 0.0f;
  w[ix(g, i, j, km)] = 0.0f;
}

void launch_bondv1(const DevDomain& d, int n, float dt) {
  const Grid& g = d.g;
  const int ilast = (n == 1) ? g.im : 1;
  if (ilast == 1) {  // steps n > 1: two launches
    const int nplane = g.jm * g.km;
    const int nin = (nplane + 255) / 256, nu = std::min(nin, 256);
    const int ngh = ((g.im + 2) * (g.km + g.jm + 3) + 255) / 256;
    TIMED(d, KC_BONDV1, hipLaunchKernelGGL(k_bondv1_in_uout, dim3((unsigned)(nin + nu)), dim3(256), 0, d.stream,
                                           g, d.m, d.F.u, d.F.v, d.F.w, d.partials, nin));
    TIMED(d, KC_BONDV1, hipLaunchKernelGGL(k_bondv1_out_ghost, dim3((unsigned)(nin + ngh)), dim3(256), 0, d.stream,
                                           g, d.m, d.F.u, d.F.v, d.F.w, (const double*)d.partials, nu, nin, dt));
    return;
  }
  TIMED(d, KC_BONDV1, hipLaunchKernelGGL(k_bondv1_inflow, bondv1_grid(g, ilast, 0), block2d(), 0, d.stream, g, d.m,
                     d.F.u, d.F.v, d.F.w, ilast, d.ctrl));
  TIMED(d, KC_BONDV1, hipLaunchKernelGGL(k_bondv1_uout, bondv1_grid(g, ilast, 1), dim3(256), 0, d.stream, g, d.F.u,
                                         d.ctrl));
  TIMED(d, KC_BONDV1, hipLaunchKernelGGL(k_bondv1_outflow, bondv1_grid(g, ilast, 2), dim3(256), 0, d.stream, g, d.m,
                     d.F.u, d.F.v, d.F.w, (const Ctrl*)d.ctrl, dt));
  TIMED(d, KC_BONDV1, hipLaunchKernelGGL(k_bondv1_ghost, bondv1_grid(g, ilast, 3), dim3(256), 0, d.stream, g, d.F.u,
                     d.F.v, d.F.w));
}

// ------------------------------------------------------------ vel2/velfg
struct P10 {
  float* a[10];
};

// vel2.f95:54-197: diu1..9 and cov1..9 (nou only as a temporary), each over
// the k range of its reference loop; at k=1 the wall layer of diu3/diu6
// (vel2.f95:135-141, 170-176) with |U_h| computed in place, and the final
// uspd=u/|U_h|, vspd=v/|V_h| (vel2.f95:54-65, 351-356).
__global__ __launch_bounds__(NT) void k_vel2(Grid g, Metrics m, const float* __restrict__ u,
                                             const float* __restrict__ v,
                                             const float* __restrict__ w, P10 diu, P10 cov,
                                             float* __restrict__ uspd, float* __restrict__ vspd,
                                             float wl) {
  const int i = 1 + blockIdx.x * BX + threadIdx.x, j = 1 + blockIdx.y * BY + threadIdx.y;
  if (i > g.im || j > g.jm) return;
  const float dxi = M1(m.dx1, i), dxp = M1(m.dx1, i + 1), dxm = M1(m.dx1, i - 1);
  const float dyj = M1(m.dy1, j), dyp = M1(m.dy1, j + 1), dym = M1(m.dy1, j - 1);
  const int sj = g.sj, sk = g.sk;
  int k0, k1;
  kspan(1, g.km + 1, k0, k1);
  for (int k = k0; k <= k1; k++) {
    const int c = ix(g, i, j, k);
    if (k <= g.km) {
      float nou = (u[c - 1] + u[c]) / 2.0f;  // vel2.f95:77-86
      float d = (-u[c - 1] + u[c]) / dxi;
      diu.a[1][c] = d;
      cov.a[1][c] = nou * d;
      nou = (v[c - sj] + v[c]) / 2.0f;
      d = (-v[c - sj] + v[c]) / dyj;
      diu.a[5][c] = d;
      cov.a[5][c] = nou * d;
      nou = (w[c - sk] + w[c]) / 2.0f;
      d = (-w[c - sk] + w[c]) / M1(m.dzn, k);
      diu.a[9][c] = d;
      cov.a[9][c] = nou * d;
      nou = (dxp * v[c - sj] + dxi * v[c + 1 - sj]) / (dxi + dxp);  // vel2.f95:106-108
      d = 2.0f * (-u[c - sj] + u[c]) / (dym + dyj);
      diu.a[2][c] = d;
      cov.a[2][c] = nou * d;
      nou = (dyp * u[c - 1] + dyj * u[c - 1 + sj]) / (dyj + dyp);  // vel2.f95:147-149
      d = 2.0f * (-v[c - 1] + v[c]) / (dxm + dxi);
      diu.a[4][c] = d;
      cov.a[4][c] = (nou - 0.0f) * d;
    }
    if (k >= 2) {
      const float dzsk = M1(m.dzs, k - 1);
      float nou = (dxp * w[c - sk] + dxi * w[c + 1 - sk]) / (dxi + dxp);  // vel2.f95:122-124
      float d = (-u[c - sk] + u[c]) / dzsk;
      diu.a[3][c] = d;
      cov.a[3][c] = nou * d;
      nou = (dyp * w[c - sk] + dyj * w[c + sj - sk]) / (dyj + dyp);  // vel2.f95:157-159
      d = (-v[c - sk] + v[c]) / dzsk;
      diu.a[6][c] = d;
      cov.a[6][c] = nou * d;
    } else {
      // |U_h| and |V_h| at k=1 (vel2.f95:54-65)
      float t = (0.5f * (v[c - sj] + v[c]) * dxp + 0.5f * (v[c + 1 - sj] + v[c + 1]) * dxi) / (dxi + dxp);
      const float us = sqrtf(u[c] * u[c] + t * t);
      t = (0.5f * (u[c - 1] + u[c]) * dyp + 0.5f * (u[c - 1 + sj] + u[c + sj]) * dyj) / (dyj + dyp);
      const float vs = sqrtf(v[c] * v[c] + t * t);
      const float dz1 = M1(m.dzn, 1);
      float nou = 0.5f * (dxp * w[c] + dxi * w[c + 1]) / (dxi + dxp);  // vel2.f95:137-139
      float d = us * 0.4f / wl / (0.5f * dz1) / 0.4f * u[c] / us;
      diu.a[3][c] = d;
      cov.a[3][c] = nou * d;
      nou = 0.5f * (dyp * w[c] + dyj * w[c + sj]) / (dyj + dyp);  // vel2.f95:172-174
      d = vs * 0.4f / wl / (0.5f * dz1) / 0.4f * v[c] / vs;
      diu.a[6][c] = d;
      cov.a[6][c] = nou * d;
      const int c2 = ix(g, i, j, -1);  // vel2.f95:351-356 (uspd, vspd: plane k = -1)
      uspd[c2] = u[c] / us;
      vspd[c2] = v[c] / vs;
    }
    if (k <= g.km - 1) {
      const float dzk = M1(m.dzn, k), dzp = M1(m.dzn, k + 1);
      float nou = (dzp * u[c - 1] + dzk * u[c - 1 + sk]) / (dzk + dzp);  // vel2.f95:182-184
      float d = 2.0f * (-w[c - 1] + w[c]) / (dxm + dxi);
      diu.a[7][c] = d;
      cov.a[7][c] = (nou - 0.0f) * d;
      nou = (dzp * v[c - sj] + dzk * v[c - sj + sk]) / (dzk + dzp);  // vel2.f95:192-194
      d = 2.0f * (-w[c - sj] + w[c]) / (dym + dyj);
      diu.a[8][c] = d;
      cov.a[8][c] = nou * d;
    }
  }
}

// vel2.f95:202-286: i=im+1 copies of 1,4,7 and periodic j copies of 2,5,8.
__global__ __launch_bounds__(256) void k_vel2_ghost1(Grid g, P10 diu, P10 cov) {
  const int njk = g.jm * g.km, nik = g.im * g.km;
  int t = blockIdx.x * 256 + threadIdx.x;
  if (t < njk) {
    const int j = 1 + t % g.jm, k = 1 + t / g.jm;
    const int c = ix(g, g.im + 1, j, k);
    diu.a[1][c] = diu.a[1][c - 1];
    cov.a[1][c] = cov.a[1][c - 1];
    diu.a[4][c] = diu.a[4][c - 1];
    cov.a[4][c] = cov.a[4][c - 1];
    if (k <= g.km - 1) {
      diu.a[7][c] = diu.a[7][c - 1];
      cov.a[7][c] = cov.a[7][c - 1];
    }
    return;
  }
  t -= njk;
  if (t >= nik) return;
  const int i = 1 + t % g.im, k = 1 + t / g.im;
  const int c0 = ix(g, i, 0, k), cjm = ix(g, i, g.jm, k), c1 = ix(g, i, 1, k), cj1 = ix(g, i, g.jm + 1, k);
  for (int q = 2; q <= 8; q += 3) {  // 2, 5, 8
    if (q == 8 && k > g.km - 1) break;
    diu.a[q][c0] = diu.a[q][cjm];
    cov.a[q][c0] = cov.a[q][cjm];
    diu.a[q][cj1] = diu.a[q][c1];
    cov.a[q][cj1] = cov.a[q][c1];
  }
}

// vel2.f95:299-314: diu2,diu3 at i=im+1 (j=1..jm+1, k=1..km+1), then
// diu4,diu6 at j=0 (i=1..im+1, k=1..km+1).  Reads values the first ghost
// pass produced or the reference never writes.
__global__ __launch_bounds__(256) void k_vel2_ghost2(Grid g, P10 diu) {
  const int nA = (g.jm + 1) * (g.km + 1), nB = (g.im + 1) * (g.km + 1);
  int t = blockIdx.x * 256 + threadIdx.x;
  if (t < nA) {
    const int j = 1 + t % (g.jm + 1), k = 1 + t / (g.jm + 1);
    const int c = ix(g, g.im + 1, j, k);
    diu.a[2][c] = diu.a[2][c - 1];
    diu.a[3][c] = diu.a[3][c - 1];
    return;
  }
  t -= nA;
  if (t >= nB) return;
  const int i = 1 + t % (g.im + 1), k = 1 + t / (g.im + 1);
  const int c0 = ix(g, i, 0, k), cjm = ix(g, i, g.jm, k);
  diu.a[4][c0] = diu.a[4][cjm];
  diu.a[6][c0] = diu.a[6][cjm];
}

// velFG.f95:64-110: f,g for k=1..km and h for k=1..km-1.
__global__ __launch_bounds__(NT) void k_velfg(Grid g, Metrics m, P10 diu, P10 cov,
                                              float* __restrict__ f, float* __restrict__ gg,
                                              float* __restrict__ h, float vn) {
  const int i = 1 + blockIdx.x * BX + threadIdx.x, j = 1 + blockIdx.y * BY + threadIdx.y;
  if (i > g.im || j > g.jm) return;
  const float dxi = M1(m.dx1, i), dxp = M1(m.dx1, i + 1);
  const float dyj = M1(m.dy1, j), dyp = M1(m.dy1, j + 1);
  const int sj = g.sj, sk = g.sk;
  int k0, k1;
  kspan(1, g.km, k0, k1);
  for (int k = k0; k <= k1; k++) {
    const int c = ix(g, i, j, k);
    const float dzk = M1(m.dzn, k);
    {
      const float covx1 = (dxp * cov.a[1][c] + dxi * cov.a[1][c + 1]) / (dxi + dxp);
      const float covy1 = (cov.a[2][c] + cov.a[2][c + sj]) / 2.0f;
      const float covz1 = (cov.a[3][c] + cov.a[3][c + sk]) / 2.0f;
      const float covc = covx1 + covy1 + covz1;
      const float dfu = 2.0f * (-diu.a[1][c] + diu.a[1][c + 1]) / (dxi + dxp) +
                        (-diu.a[2][c] + diu.a[2][c + sj]) / dyj + (-diu.a[3][c] + diu.a[3][c + sk]) / dzk;
      const float df = vn * dfu;
      f[c] = (-covc + df);
    }
    {
      const float covx1 = (cov.a[4][c] + cov.a[4][c + 1]) / 2.0f;
      const float covy1 = (dyp * cov.a[5][c] + dyj * cov.a[5][c + sj]) / (dyj + dyp);
      const float covz1 = (cov.a[6][c] + cov.a[6][c + sk]) / 2.0f;
      const float covc = covx1 + covy1 + covz1;
      const float dfv = (-diu.a[4][c] + diu.a[4][c + 1]) / dxi +
                        2.0f * (-diu.a[5][c] + diu.a[5][c + sj]) / (dyj + dyp) +
                        (-diu.a[6][c] + diu.a[6][c + sk]) / dzk;
      const float df = vn * dfv;
      gg[c] = (-covc + df);
    }
    if (k <= g.km - 1) {
      const float dzp = M1(m.dzn, k + 1);
      const float covx1 = (cov.a[7][c] + cov.a[7][c + 1]) / 2.0f;
      const float covy1 = (cov.a[8][c] + cov.a[8][c + sj]) / 2.0f;
      const float covz1 = (dzp * cov.a[9][c] + dzk * cov.a[9][c + sk]) / (dzk + dzp);
      const float covc = covx1 + covy1 + covz1;
      const float dfw = (-diu.a[7][c] + diu.a[7][c + 1]) / dxi + (-diu.a[8][c] + diu.a[8][c + sj]) / dyj +
                        (-diu.a[9][c] + diu.a[9][c + sk]) / M1(m.dzs, k);
      const float df = vn * dfw;
      h[c] = (-covc + df);
    }
  }
}

__host__ __device__ inline P10 p10(float* const* a) {
  P10 r;
  for (int q = 0; q < 10; q++) r.a[q] = a[q];
  return r;
}

void launch_velfg(const DevDomain& d, float vn) {
  const Grid& g = d.g;
  const P10 diu = p10(d.F.diu), cov = p10(d.F.cov);
  TIMED(d, KC_VEL2, hipLaunchKernelGGL(k_vel2, grid2d(g.im, g.jm, kz(g, 1)), block2d(), 0, d.stream, g, d.m, d.F.u,
                     d.F.v, d.F.w, diu, cov, d.F.uspd, d.F.vspd, d.wl));
  const int n1 = g.jm * g.km + g.im * g.km;
  TIMED(d, KC_VEL2_GHOST, hipLaunchKernelGGL(k_vel2_ghost1, dim3((n1 + 255) / 256), dim3(256), 0, d.stream, g, diu, cov));
  const int n2 = (g.jm + 1) * (g.km + 1) + (g.im + 1) * (g.km + 1);
  TIMED(d, KC_VEL2_GHOST, hipLaunchKernelGGL(k_vel2_ghost2, dim3((n2 + 255) / 256), dim3(256), 0, d.stream, g, diu));
  TIMED(d, KC_VELFG, hipLaunchKernelGGL(k_velfg, grid2d(g.im, g.jm, kz(g)), block2d(), 0, d.stream, g, d.m, diu, cov,
                     d.F.f, d.F.g, d.F.h, vn));
}

// boundsm.f95:18-64 for the ghost cells whose composite copy (boundsm_body)
// reads interior cell (i,j,k), whose new value is val: its images on the
// i, j and k faces it touches, negated at k = 0.  The threads of row j = 1
// also write the ghosts of row j = -1, whose copies read row j = -1 itself,
// which nothing writes.
__device__ __forceinline__ void sm_ghosts(const Grid& g, int i, int j, int k, float val, float* __restrict__ sm) {
  constexpr int kNone = -1000;
  const int ie = (i == 1) ? 0 : ((i == g.im) ? g.im + 1 : kNone);
  const int je = (j == 1) ? 0 : ((j == g.jm) ? g.jm + 1 : kNone);
  const int ke = (k == 1) ? 0 : ((k == g.km) ? g.km + 1 : kNone);
  if (ie == kNone && je == kNone && ke == kNone) return;
  for (int a = 0; a < 2; a++) {
    const int gi = a ? ie : i;
    if (gi == kNone) continue;
    for (int c = 0; c < 2; c++) {
      const int gk = c ? ke : k;
      if (gk == kNone) continue;
      for (int b = 0; b < 2; b++) {
        const int gj = b ? je : j;
        if (gj == kNone || (a | b | c) == 0) continue;
        sm[ix(g, gi, gj, gk)] = (gk == 0) ? -val : val;
      }
      if (j == 1 && (a | c) != 0) {
        const float s = sm[ix(g, i, -1, k)];
        sm[ix(g, gi, -1, gk)] = (gk == 0) ? -s : s;
      }
    }
  }
}

// sm of one cell (les.f95:44-79) from the diu values around it, in the
// reference's order of operations: a = 0.1*delx1(k); the diu2, diu4, diu6,
// diu8 (x, y, z) averages of the four values at the cell's corners.
__device__ __forceinline__ float sm_value(float dlx, float d1c, float d2im, float d2imjp, float d2c, float d2jp,
                                          float d3im, float d3imkp, float d3c, float d3kp, float d4c, float d4jm,
                                          float d4ip, float d4ipjm, float d5c, float d6jm, float d6jmkp, float d6c,
                                          float d6kp, float d7c, float d7km, float d7ip, float d7ipkm, float d8c,
                                          float d8km, float d8jp, float d8jpkm, float d9c) {
  const float dudxx1 = d1c;
  const float dudyx1 = (d2im + d2imjp + d2c + d2jp) * 0.25f;
  const float dudzx1 = (d3im + d3imkp + d3c + d3kp) * 0.25f;
  const float dvdxx1 = (d4c + d4jm + d4ip + d4ipjm) * 0.25f;
  const float dvdyx1 = d5c;
  const float dvdzx1 = (d6jm + d6jmkp + d6c + d6kp) * 0.25f;
  const float dwdxx1 = (d7c + d7km + d7ip + d7ipkm) * 0.25f;
  const float dwdyx1 = (d8c + d8km + d8jp + d8jpkm) * 0.25f;
  const float dwdzx1 = d9c;
  const float a = 0.1f * dlx;
  const float s = 2.0f * (dudxx1 * dudxx1 + dvdyx1 * dvdyx1 + dwdzx1 * dwdzx1) +
                  (dudyx1 + dvdxx1) * (dudyx1 + dvdxx1) + (dwdyx1 + dvdzx1) * (dwdyx1 + dvdzx1) +
                  (dudzx1 + dwdxx1) * (dudzx1 + dwdxx1);
  return a * a * sqrtf(s);
}

// ------------------------------------------- fused vel2 + velfg + feedbf
// One pass of vel2 (vel2.f95:54-356), velfg (velFG.f95:64-110) and feedbf
// (feedbf.f95:35-61) for the whole-step path.  A thread owns the column
// (i,j) and marches its k levels.  The products cov1..cov9 (dead after
// velfg) are never stored; diu1..diu9, which les reads, are stored together
// with every ghost cell vel2's copies leave (vel2.f95:202-314), each written
// by the thread that owns its source.  The diu/cov values velfg needs at
// i+1 and j+1 are recomputed (at i=im and j=jm the ghost copies: the value
// at im, and the periodic image at j=1), those at k+1 carried from level to
// level.  Every value is computed by the expression of the unfused kernels,
// so results are bit-identical to velfg followed by feedbf.
struct DC {
  float d, c;  // diu, cov
};

struct VelCtx {
  Grid g;
  Metrics m;
  const float* __restrict__ u;
  const float* __restrict__ v;
  const float* __restrict__ w;
  float wl;
  __device__ float U(int i, int j, int k) const { return u[ix(g, i, j, k)]; }
  __device__ float V(int i, int j, int k) const { return v[ix(g, i, j, k)]; }
  __device__ float W(int i, int j, int k) const { return w[ix(g, i, j, k)]; }
  __device__ float dx(int i) const { return M1(m.dx1, i); }
  __device__ float dy(int j) const { return M1(m.dy1, j); }
  // vel2.f95:77-86
  __device__ DC q1(int i, int j, int k) const {
    const float a = U(i - 1, j, k), b = U(i, j, k);
    const float nou = (a + b) / 2.0f, d = (-a + b) / dx(i);
    return {d, nou * d};
  }
  __device__ DC q5(int i, int j, int k) const {
    const float a = V(i, j - 1, k), b = V(i, j, k);
    const float nou = (a + b) / 2.0f, d = (-a + b) / dy(j);
    return {d, nou * d};
  }
  __device__ DC q9(int i, int j, int k) const {
    const float a = W(i, j, k - 1), b = W(i, j, k);
    const float nou = (a + b) / 2.0f, d = (-a + b) / M1(m.dzn, k);
    return {d, nou * d};
  }
  // vel2.f95:103-111
  __device__ DC q2(int i, int j, int k) const {
    const float dxi = dx(i), dxp = dx(i + 1);
    const float nou = (dxp * V(i, j - 1, k) + dxi * V(i + 1, j - 1, k)) / (dxi + dxp);
    const float d = 2.0f * (-U(i, j - 1, k) + U(i, j, k)) / (dy(j - 1) + dy(j));
    return {d, nou * d};
  }
  // vel2.f95:144-152
  __device__ DC q4(int i, int j, int k) const {
    const float dyj = dy(j), dyp = dy(j + 1);
    const float nou = (dyp * U(i - 1, j, k) + dyj * U(i - 1, j + 1, k)) / (dyj + dyp);
    const float d = 2.0f * (-V(i - 1, j, k) + V(i, j, k)) / (dx(i - 1) + dx(i));
    return {d, (nou - 0.0f) * d};
  }
  // vel2.f95:119-124 (k >= 2)
  __device__ DC q3(int i, int j, int k) const {
    const float dxi = dx(i), dxp = dx(i + 1);
    const float nou = (dxp * W(i, j, k - 1) + dxi * W(i + 1, j, k - 1)) / (dxi + dxp);
    const float d = (-U(i, j, k - 1) + U(i, j, k)) / M1(m.dzs, k - 1);
    return {d, nou * d};
  }
  // vel2.f95:154-159 (k >= 2)
  __device__ DC q6(int i, int j, int k) const {
    const float dyj = dy(j), dyp = dy(j + 1);
    const float nou = (dyp * W(i, j, k - 1) + dyj * W(i, j + 1, k - 1)) / (dyj + dyp);
    const float d = (-V(i, j, k - 1) + V(i, j, k)) / M1(m.dzs, k - 1);
    return {d, nou * d};
  }
  // vel2.f95:179-186 (k <= km-1)
  __device__ DC q7(int i, int j, int k) const {
    const float dzk = M1(m.dzn, k), dzp = M1(m.dzn, k + 1);
    const float nou = (dzp * U(i - 1, j, k) + dzk * U(i - 1, j, k + 1)) / (dzk + dzp);
    const float d = 2.0f * (-W(i - 1, j, k) + W(i, j, k)) / (dx(i - 1) + dx(i));
    return {d, (nou - 0.0f) * d};
  }
  // vel2.f95:189-197 (k <= km-1)
  __device__ DC q8(int i, int j, int k) const {
    const float dzk = M1(m.dzn, k), dzp = M1(m.dzn, k + 1);
    const float nou = (dzp * V(i, j - 1, k) + dzk * V(i, j - 1, k + 1)) / (dzk + dzp);
    const float d = 2.0f * (-W(i, j - 1, k) + W(i, j, k)) / (dy(j - 1) + dy(j));
    return {d, nou * d};
  }
  // the wall layer k=1 of diu3/diu6 with |U_h|, |V_h| (vel2.f95:54-65,
  // 135-141, 170-176) and uspd, vspd (vel2.f95:351-356)
  __device__ void wall(int i, int j, DC& r3, DC& r6, float& us_out, float& vs_out) const {
    const float dxi = dx(i), dxp = dx(i + 1), dyj = dy(j), dyp = dy(j + 1);
    const float uc = U(i, j, 1), vc = V(i, j, 1);
    float t = (0.5f * (V(i, j - 1, 1) + vc) * dxp + 0.5f * (V(i + 1, j - 1, 1) + V(i + 1, j, 1)) * dxi) / (dxi + dxp);
    const float us = sqrtf(uc * uc + t * t);
    t = (0.5f * (U(i - 1, j, 1) + uc) * dyp + 0.5f * (U(i - 1, j + 1, 1) + U(i, j + 1, 1)) * dyj) / (dyj + dyp);
    const float vs = sqrtf(vc * vc + t * t);
    const float dz1 = M1(m.dzn, 1);
    float nou = 0.5f * (dxp * W(i, j, 1) + dxi * W(i + 1, j, 1)) / (dxi + dxp);
    float d = us * 0.4f / wl / (0.5f * dz1) / 0.4f * uc / us;
    r3 = {d, nou * d};
    nou = 0.5f * (dyp * W(i, j, 1) + dyj * W(i, j + 1, 1)) / (dyj + dyp);
    d = vs * 0.4f / wl / (0.5f * dz1) / 0.4f * vc / vs;
    r6 = {d, nou * d};
    us_out = uc / us;
    vs_out = vc / vs;
  }
};

// SM: also sm with its ghost cells (les.f95:44-79, boundsm.f95:18-64), the
// whole-step path's k_les_sm.  The diu values sm reads around the cell that
// other threads store in this launch are recomputed (at i-1, j-1 and j+1,
// the j=0 and j=jm+1 ghosts as the periodic images they copy), or carried
// from level to level (k-1, k+1); those no thread stores (at i=0, and diu7,
// diu8 at k=0 and k=km) are read from the fields.
template <bool SM>
__device__ __forceinline__ void velfg_feedbf_body(
    const VBlk& vb, const VelCtx& X, const P10& diu, float* __restrict__ f, float* __restrict__ gg,
    float* __restrict__ h, float* __restrict__ usum, float* __restrict__ vsum, float* __restrict__ wsum,
    const float* __restrict__ bm, const float* __restrict__ cm, const float* __restrict__ dm,
    const float* __restrict__ ms, float* __restrict__ uspd, float* __restrict__ vspd, float vn, float alpha,
    float dt, float beta, int kchunk, const float* __restrict__ delx1, float* __restrict__ sm) {
  const Grid& g = X.g;
  const Metrics& m = X.m;
  // a wave is one row of the block (blockDim.x == 64): j is wave-uniform
  const int i = 1 + vb.x * BX + threadIdx.x, j = 1 + vb.y * BY + wave_row();
  if (i > g.im || j > g.jm) return;
  const int im = g.im, jm = g.jm, km = g.km;
  const int k0 = 1 + vb.z * kchunk, k1 = min(km, k0 + kchunk - 1);
  if (k0 > km) return;
  const int jn = (j == jm) ? 1 : j + 1;  // the row velfg reads at j+1 (periodic ghost at jm)
  const bool ilast = (i == im);
  const float dxi = X.dx(i), dxp = X.dx(i + 1);
  const float dyj = X.dy(j), dyp = X.dy(j + 1);
  // Masks of buildings standing on the ground (ms: a value per column at
  // (i-1)+(j-1)*im, then each level's threshold): the three are equal and
  // are 1 exactly where the column's value exceeds the level's threshold --
  // zbm(i,j) and z2(k)+0.5*dzn(k) for masks built from the height map
  // (k_masks' expression), the top solid level K(i,j) and k - 1/2 for given
  // masks of that form -- so the three mask fields need not be read.  ms is
  // null for other masks.
  const float* const mtop = ms ? ms + (size_t)im * jm : nullptr;
  const float zb = ms ? ms[(i - 1) + (j - 1) * im] : 0.0f;
  float* const* D = diu.a;
  auto put = [&](int q, int ii, int jj, int kk, float val) { D[q][ix(g, ii, jj, kk)] = val; };
  auto cpy = [&](int q, int ii, int jj, int kk, int si, int sj, int sk) {
    D[q][ix(g, ii, jj, kk)] = D[q][ix(g, si, sj, sk)];
  };
  // diu3, diu6 (and diu9) of level k; the march carries them from k to k+1
  DC z3, z6, z9 = {0.f, 0.f};
  if (k0 == 1) {
    float us, vs;
    X.wall(i, j, z3, z6, us, vs);
    const int c2 = ix(g, i, j, -1);  // uspd, vspd: plane k = -1
    uspd[c2] = us;
    vspd[c2] = vs;
  } else {
    z3 = X.q3(i, j, k0);
    z6 = X.q6(i, j, k0);
  }
  z9 = X.q9(i, j, k0);
  // SM: jp is the row whose diu4/diu6 the cell reads at j-1 (the periodic
  // image jm at j=1); s3m, s6m are diu3(i-1,j,k), diu6(i,j-1,k); p7c, p7p,
  // p8c, p8j diu7 at (i,j),(i+1,j) and diu8 at (i,j),(i,j+1) of level k-1
  const int jp = (j == 1) ? jm : j - 1;
  const bool ifirst = (i == 1);
  float s3m = 0.f, s6m = 0.f, p7c = 0.f, p7p = 0.f, p8c = 0.f, p8j = 0.f;
  if (SM) {
    const float m3 = D[3][ix(g, 0, j, k0)];  // i=0: never stored
    if (k0 == 1) {
      DC r3, r6, t3, t6;
      float us, vs;
      X.wall(i - 1, j, r3, t6, us, vs);
      X.wall(i, jp, t3, r6, us, vs);
      s3m = ifirst ? m3 : r3.d;
      s6m = r6.d;
      p7c = D[7][ix(g, i, j, 0)];  // k=0: never stored
      p7p = D[7][ix(g, i + 1, j, 0)];
      p8c = D[8][ix(g, i, j, 0)];
      p8j = D[8][ix(g, i, j + 1, 0)];
    } else {
      const float r3 = X.q3(i - 1, j, k0).d;
      s3m = ifirst ? m3 : r3;
      s6m = X.q6(i, jp, k0).d;
      const DC e7 = X.q7(i, j, k0 - 1), f7 = X.q7(i + 1, j, k0 - 1);
      p7c = e7.d;
      p7p = ilast ? e7.d : f7.d;
      p8c = X.q8(i, j, k0 - 1).d;
      p8j = X.q8(i, jn, k0 - 1).d;
    }
  }
  // One level.  HK: k <= km-1, where h, diu7, diu8 and diu9's k+1 term
  // exist; the levels below km run the HK body, which has no branch on k,
  // and the top level km its own.
  auto level = [&](auto HKC, int k) {
    constexpr bool hk = decltype(HKC)::value;
    const int c = ix(g, i, j, k);
    const float dzk = M1(m.dzn, k);
    const DC a1 = X.q1(i, j, k), a2 = X.q2(i, j, k), a4 = X.q4(i, j, k), a5 = X.q5(i, j, k);
    const DC a3 = z3, a6 = z6, a9 = z9;
    DC a7 = {0.f, 0.f}, a8 = {0.f, 0.f};
    if (hk) {
      a7 = X.q7(i, j, k);
      a8 = X.q8(i, j, k);
    }
    // next level (k+1 <= km+1 for diu3/diu6, <= km for diu9)
    const DC n3 = X.q3(i, j, k + 1), n6 = X.q6(i, j, k + 1);
    const DC n9 = hk ? X.q9(i, j, k + 1) : DC{0.f, 0.f};
    // neighbours at i+1 (the copy of im at i=im) and j+1 (row 1 at j=jm)
    // computed by every lane (at i = im from the ghost cells, then replaced):
    // no lane-dependent branch around the loads
    const DC c1 = X.q1(i + 1, j, k), c4 = X.q4(i + 1, j, k);
    const DC b1 = ilast ? a1 : c1;
    const DC b4 = ilast ? a4 : c4;
    const DC b2 = X.q2(i, jn, k);
    const DC b5 = X.q5(i, jn, k);
    // ---- velFG.f95:64-77
    float fv, gv;
    {
      const float covx1 = (dxp * a1.c + dxi * b1.c) / (dxi + dxp);
      const float covy1 = (a2.c + b2.c) / 2.0f;
      const float covz1 = (a3.c + n3.c) / 2.0f;
      const float covc = covx1 + covy1 + covz1;
      const float dfu = 2.0f * (-a1.d + b1.d) / (dxi + dxp) + (-a2.d + b2.d) / dyj + (-a3.d + n3.d) / dzk;
      fv = (-covc + vn * dfu);
    }
    // ---- velFG.f95:80-93
    {
      const float covx1 = (a4.c + b4.c) / 2.0f;
      const float covy1 = (dyp * a5.c + dyj * b5.c) / (dyj + dyp);
      const float covz1 = (a6.c + n6.c) / 2.0f;
      const float covc = covx1 + covy1 + covz1;
      const float dfv = (-a4.d + b4.d) / dxi + 2.0f * (-a5.d + b5.d) / (dyj + dyp) + (-a6.d + n6.d) / dzk;
      gv = (-covc + vn * dfv);
    }
    // ---- velFG.f95:97-110 (k <= km-1; h(km) keeps its value)
    float hv;
    DC b7 = {0.f, 0.f}, b8 = {0.f, 0.f};
    if (hk) {
      const DC c7 = X.q7(i + 1, j, k);
      b7 = ilast ? a7 : c7;
      b8 = X.q8(i, jn, k);
      const float dzp = M1(m.dzn, k + 1);
      const float covx1 = (a7.c + b7.c) / 2.0f;
      const float covy1 = (a8.c + b8.c) / 2.0f;
      const float covz1 = (dzp * a9.c + dzk * n9.c) / (dzk + dzp);
      const float covc = covx1 + covy1 + covz1;
      const float dfw = (-a7.d + b7.d) / dxi + (-a8.d + b8.d) / dyj + (-a9.d + n9.d) / M1(m.dzs, k);
      hv = (-covc + vn * dfw);
    } else {
      hv = h[c];
    }
    // ---- feedbf.f95:35-61
    {
      const float uc = X.u[c], vc = X.v[c], wc = X.w[c];
      float b, cc, dd;
      if (ms) {
        b = (zb > mtop[k - 1]) ? 1.0f : 0.0f;
        cc = b;
        dd = b;
      } else {
        b = bm[c];
        cc = cm[c];
        dd = dm[c];
      }
      const float su = (usum[c] + uc) * b, sv = (vsum[c] + vc) * cc, sw = (wsum[c] + wc) * dd;
      usum[c] = su;
      vsum[c] = sv;
      wsum[c] = sw;
      f[c] = fv + (alpha * su * dt + beta * uc * b);
      gg[c] = gv + (alpha * sv * dt + beta * vc * cc);
      h[c] = hv + (alpha * sw * dt + beta * wc * dd);
    }
    // ---- diu of level k and the ghost cells vel2.f95:202-314 derives from them
    put(1, i, j, k, a1.d);
    put(2, i, j, k, a2.d);
    put(3, i, j, k, a3.d);
    put(4, i, j, k, a4.d);
    put(5, i, j, k, a5.d);
    put(6, i, j, k, a6.d);
    put(9, i, j, k, a9.d);
    if (hk) {
      put(7, i, j, k, a7.d);
      put(8, i, j, k, a8.d);
    }
    if (ilast) {  // i=im+1 copies of 1,4,7 (vel2.f95:208-226) and of 2,3 (vel2.f95:299-306)
      put(1, im + 1, j, k, a1.d);
      put(4, im + 1, j, k, a4.d);
      if (hk) put(7, im + 1, j, k, a7.d);
      put(2, im + 1, j, k, a2.d);
      put(3, im + 1, j, k, a3.d);
    }
    if (j == jm) {  // j=0 copies of 2,5,8 (vel2.f95:231-286) and of 4,6 (vel2.f95:308-314)
      put(2, i, 0, k, a2.d);
      put(5, i, 0, k, a5.d);
      if (hk) put(8, i, 0, k, a8.d);
      put(4, i, 0, k, a4.d);
      put(6, i, 0, k, a6.d);
      if (ilast) {
        put(4, im + 1, 0, k, a4.d);               // diu4(im+1,jm) is the i-copy of diu4(im,jm)
        cpy(6, im + 1, 0, k, im + 1, jm, k);      // diu6(im+1,jm) is never written
      }
    }
    if (j == 1) {  // j=jm+1 copies of 2,5,8
      put(2, i, jm + 1, k, a2.d);
      put(5, i, jm + 1, k, a5.d);
      if (hk) put(8, i, jm + 1, k, a8.d);
      if (ilast) {
        put(2, im + 1, jm + 1, k, a2.d);          // diu2(im,jm+1) is the j-copy of diu2(im,1)
        cpy(3, im + 1, jm + 1, k, im, jm + 1, k); // diu3(im,jm+1) is never written
      }
    }
    if (SM) {  // les.f95:44-79 and boundsm.f95:18-64
      const float m2a = D[2][ix(g, 0, j, k)], m2b = D[2][ix(g, 0, j + 1, k)];  // i=0: never stored
      const float m3 = D[3][ix(g, 0, j, k + 1)];
      const float r2a = X.q2(i - 1, j, k).d, r2b = X.q2(i - 1, jn, k).d, r3 = X.q3(i - 1, j, k + 1).d;
      const float d4jm = X.q4(i, jp, k).d, r4 = X.q4(i + 1, jp, k).d;
      const float n6m = X.q6(i, jp, k + 1).d;
      const float n3m = ifirst ? m3 : r3;
      float c7c, c7p, c8c, c8j;
      if (hk) {
        c7c = a7.d;
        c7p = b7.d;
        c8c = a8.d;
        c8j = b8.d;
      } else {  // diu7, diu8 at k=km: never stored
        c7c = D[7][c];
        c7p = D[7][c + 1];
        c8c = D[8][c];
        c8j = D[8][c + g.sj];
      }
      const float val = sm_value(delx1[k], a1.d, ifirst ? m2a : r2a, ifirst ? m2b : r2b, a2.d, b2.d, s3m, n3m, a3.d,
                                 n3.d, a4.d, d4jm, b4.d, ilast ? d4jm : r4, a5.d, s6m, n6m, a6.d, n6.d, c7c, p7c, c7p,
                                 p7p, c8c, p8c, c8j, p8j, a9.d);
      sm[c] = val;
      sm_ghosts(g, i, j, k, val, sm);
      s3m = n3m;
      s6m = n6m;
      p7c = c7c;
      p7p = c7p;
      p8c = c8c;
      p8j = c8j;
    }
    z3 = n3;
    z6 = n6;
    z9 = n9;
  };
  const int kh = min(k1, km - 1);
  for (int k = k0; k <= kh; k++) level(std::true_type{}, k);
  if (k1 == km) level(std::false_type{}, km);
  if (k1 == km) {  // level km+1: diu3, diu6 (vel2.f95:119-176) and the copies from it
    const int k = km + 1;
    put(3, i, j, k, z3.d);
    put(6, i, j, k, z6.d);
    if (ilast) {
      put(3, im + 1, j, k, z3.d);
      cpy(2, im + 1, j, k, im, j, k);
    }
    if (j == jm) {
      cpy(4, i, 0, k, i, jm, k);
      put(6, i, 0, k, z6.d);
      if (ilast) {
        cpy(4, im + 1, 0, k, im + 1, jm, k);
        cpy(6, im + 1, 0, k, im + 1, jm, k);
      }
    }
    if (j == 1 && ilast) {
      cpy(2, im + 1, jm + 1, k, im, jm + 1, k);
      cpy(3, im + 1, jm + 1, k, im, jm + 1, k);
    }
  }
}
template <bool SM>
__global__ __launch_bounds__(NT) void k_velfg_feedbf(VelCtx X, P10 diu, float* __restrict__ f,
                                                     float* __restrict__ gg, float* __restrict__ h,
                                                     float* __restrict__ usum, float* __restrict__ vsum,
                                                     float* __restrict__ wsum, const float* __restrict__ bm,
                                                     const float* __restrict__ cm, const float* __restrict__ dm,
                                                     const float* __restrict__ ms, float* __restrict__ uspd,
                                                     float* __restrict__ vspd, float vn, float alpha, float dt,
                                                     float beta, int kchunk, const float* __restrict__ delx1,
                                                     float* __restrict__ sm) {
  velfg_feedbf_body<SM>(hw_blk<false>(), X, diu, f, gg, h, usum, vsum, wsum, bm, cm, dm, ms, uspd, vspd, vn, alpha,
                        dt, beta, kchunk, delx1, sm);
}

// Grid of k_velfg_feedbf and its k levels per thread.
static dim3 velfg_grid(const DevDomain& d, int* kchunk) {
  const Grid& g = d.g;
  const int gx = (g.im + BX - 1) / BX, gy = (g.jm + BY - 1) / BY;
  // about 36 blocks per compute unit: short k runs measured fastest (at
  // 300x300x90, 4 levels per thread: 310 us against 335 us for 18 levels)
  int nz = 36 * (d.ncu > 0 ? d.ncu : 256) / (gx * gy);
  if (nz < 1) nz = 1;
  if (nz > g.km) nz = g.km;
  int kc = (g.km + nz - 1) / nz;
  if (d.velfg_kchunk > 0) kc = d.velfg_kchunk < g.km ? d.velfg_kchunk : g.km;
  nz = (g.km + kc - 1) / kc;
  *kchunk = kc;
  return dim3((unsigned)gx, (unsigned)gy, (unsigned)nz);
}

void launch_velfg_feedbf(const DevDomain& d, float vn, float alpha, float dt, float beta) {
  int kc = 0;
  const dim3 grid = velfg_grid(d, &kc);
  VelCtx X{d.g, d.m, d.F.u, d.F.v, d.F.w, d.wl};
  if (d.fuse_sm)
    TIMED(d, KC_VEL2, hipLaunchKernelGGL(k_velfg_feedbf<true>, grid, block2d(), 0, d.stream, X, p10(d.F.diu),
                                          d.F.f, d.F.g, d.F.h, d.F.usum, d.F.vsum, d.F.wsum, d.F.bmask, d.F.cmask,
                                          d.F.dmask, d.mask_src, d.F.uspd, d.F.vspd, vn, alpha, dt, beta, kc,
                                          d.m.delx1 + 1, d.F.sm));
  else
    TIMED(d, KC_VEL2, hipLaunchKernelGGL(k_velfg_feedbf<false>, grid, block2d(), 0, d.stream, X, p10(d.F.diu),
                                          d.F.f, d.F.g, d.F.h, d.F.usum, d.F.vsum, d.F.wsum, d.F.bmask, d.F.cmask,
                                          d.F.dmask, d.mask_src, d.F.uspd, d.F.vspd, vn, alpha, dt, beta, kc,
                                          nullptr, nullptr));
}

// ----------------------------------------------------------------- feedbf
// feedbf.f95:35-61 (fx,fy,fz are dead outputs and are not stored).
__global__ __launch_bounds__(NT) void k_feedbf(Grid g, const float* __restrict__ u,
                                               const float* __restrict__ v,
                                               const float* __restrict__ w,
                                               const float* __restrict__ bm,
                                               const float* __restrict__ cm,
                                               const float* __restrict__ dm, float* __restrict__ us,
                                               float* __restrict__ vs, float* __restrict__ ws,
                                               float* __restrict__ f, float* __restrict__ gg,
                                               float* __restrict__ h, float alpha, float dt,
                                               float beta) {
  const int i = 1 + blockIdx.x * BX + threadIdx.x, j = 1 + blockIdx.y * BY + threadIdx.y;
  if (i > g.im || j > g.jm) return;
  int k0, k1;
  kspan(1, g.km, k0, k1);
  for (int k = k0; k <= k1; k++) {
    const int c = ix(g, i, j, k);
    const float uc = u[c], vc = v[c], wc = w[c], b = bm[c], cc = cm[c], dd = dm[c];
    const float su = (us[c] + uc) * b, sv = (vs[c] + vc) * cc, sw = (ws[c] + wc) * dd;
    us[c] = su;
    vs[c] = sv;
    ws[c] = sw;
    const float fx = alpha * su * dt + beta * uc * b;
    const float fy = alpha * sv * dt + beta * vc * cc;
    const float fz = alpha * sw * dt + beta * wc * dd;
    f[c] = f[c] + fx;
    gg[c] = gg[c] + fy;
    h[c] = h[c] + fz;
  }
}

void launch_feedbf(const DevDomain& d, float alpha, float dt, float beta) {
  const Grid& g = d.g;
  TIMED(d, KC_FEEDBF, hipLaunchKernelGGL(k_feedbf, grid2d(g.im, g.jm, kz(g)), block2d(), 0, d.stream, g, d.F.u, d.F.v,
                     d.F.w, d.F.bmask, d.F.cmask, d.F.dmask, d.F.usum, d.F.vsum, d.F.wsum, d.F.f,
                     d.F.g, d.F.h, alpha, dt, beta));
}

// -------------------------------------------------------------------- les
// Eddy viscosity sm (les.f95:45-65); delx1 arrives precomputed per level.

// GHOSTS: also write the ghost cells boundsm sets (k_les_sm of a domain's
// own step; the batched step keeps its boundsm launch).
template <bool GHOSTS>
__device__ __forceinline__ void les_sm_body(const VBlk& vb, const Grid& g, const P10& diu,
                                            const float* __restrict__ delx1, float* __restrict__ sm) {
  const int i = 1 + vb.x * BX + threadIdx.x, j = 1 + vb.y * BY + wave_row();
  if (i > g.im || j > g.jm) return;
  const int sj = g.sj, sk = g.sk;
  const float* d2 = diu.a[2];
  const float* d3 = diu.a[3];
  const float* d4 = diu.a[4];
  const float* d6 = diu.a[6];
  const float* d7 = diu.a[7];
  const float* d8 = diu.a[8];
  int k0, k1;
  kspan(vb, 1, g.km, k0, k1);
  if (k0 > k1) return;
  // A thread walks its run of levels keeping what level k+1 shares with
  // level k: diu3 and diu6 at k+1 become the next step's level-k values,
  // diu7 and diu8 at k become its level-(k-1) values.
  int c = ix(g, i, j, k0);
  float h3m = d3[c - 1], h3c = d3[c], h6m = d6[c - sj], h6c = d6[c];  // diu3, diu6 at level k
  float p7c = d7[c - sk], p7p = d7[c + 1 - sk], p8c = d8[c - sk], p8j = d8[c + sj - sk];  // diu7, diu8 at k-1
  for (int k = k0; k <= k1; k++, c += sk) {
    const float l3m = h3m, l3c = h3c, l6m = h6m, l6c = h6c;
    h3m = d3[c - 1 + sk];
    h3c = d3[c + sk];
    h6m = d6[c - sj + sk];
    h6c = d6[c + sk];
    const float c7c = d7[c], c7p = d7[c + 1], c8c = d8[c], c8j = d8[c + sj];
    const float val = sm_value(delx1[k], diu.a[1][c], d2[c - 1], d2[c - 1 + sj], d2[c], d2[c + sj], l3m, h3m, l3c,
                               h3c, d4[c], d4[c - sj], d4[c + 1], d4[c + 1 - sj], diu.a[5][c], l6m, h6m, l6c, h6c,
                               c7c, p7c, c7p, p7p, c8c, p8c, c8j, p8j, diu.a[9][c]);
    p7c = c7c;
    p7p = c7p;
    p8c = c8c;
    p8j = c8j;
    sm[c] = val;
    if (GHOSTS) sm_ghosts(g, i, j, k, val, sm);
  }
}
// les.f95:44-79 with boundsm.f95:18-64
__global__ __launch_bounds__(NT) void k_les_sm(Grid g, P10 diu, const float* __restrict__ delx1,
                                               float* __restrict__ sm) {
  les_sm_body<true>(hw_blk<true>(), g, diu, delx1, sm);
}
// Grid of a stencil kernel whose threads walk runs of kchunk levels
// (0: automatic, about 36 blocks per compute unit, as velfg).
static dim3 krun_grid(const DevDomain& d, int kchunk) {
  const Grid& g = d.g;
  const int gx = (g.im + BX - 1) / BX, gy = (g.jm + BY - 1) / BY;
  int kc = kchunk;
  if (kc <= 0) {
    int nz = 36 * (d.ncu > 0 ? d.ncu : 256) / (gx * gy);
    nz = std::max(1, std::min(nz, g.km));
    kc = (g.km + nz - 1) / nz;
  }
  kc = std::min(kc, g.km);
  return grid2d(g.im, g.jm, (g.km + kc - 1) / kc);
}

// boundsm.f95:18-64 as one pass: the serial build runs the i-Neumann
// copies (j=-1..jm+1, k=0..km+1), then the j-Neumann copies (i=0..im+1,
// k=0..km+1), then sm(k=0)=-sm(1), sm(km+1)=sm(km) (j=-1..jm+1,
// i=0..im+1).  Each written cell receives the composite of those copies,
// whose source is always a cell none of them writes.
__device__ __forceinline__ bool sm_written(const Grid& g, int i, int j, int k) {
  const bool ib = (i == 0 || i == g.im + 1), jb = (j == 0 || j == g.jm + 1), kb = (k == 0 || k == g.km + 1);
  const bool iin = (i >= 0 && i <= g.im + 1), jin = (j >= -1 && j <= g.jm + 1), kin = (k >= 0 && k <= g.km + 1);
  return (ib && jin && kin) || (jb && iin && kin) || (kb && jin && iin);
}

__device__ __forceinline__ void boundsm_body(const VBlk& vb, const Grid& g, float* __restrict__ sm) {
  // enumerate the box i=0..im+1, j=-1..jm+1, k=0..km+1 face by face:
  // k faces (2 planes), then j faces (rows j=-1,0,jm+1 for k=1..km), then
  // i faces (i=0,im+1 for j=1..jm, k=1..km)
  const int ni = g.im + 2, nj = g.jm + 3;
  const int nK = 2 * ni * nj, nJ = 3 * ni * g.km, nI = 2 * g.jm * g.km;
  int t = vb.x * 256 + threadIdx.x;
  int i, j, k;
  if (t < nK) {
    i = t % ni;
    j = -1 + (t / ni) % nj;
    k = (t / (ni * nj)) ? g.km + 1 : 0;
  } else if ((t -= nK) < nJ) {
    i = t % ni;
    const int r = (t / ni) % 3;
    j = (r == 0) ? -1 : ((r == 1) ? 0 : g.jm + 1);
    k = 1 + t / (ni * 3);
  } else if ((t -= nJ) < nI) {
    i = (t & 1) ? g.im + 1 : 0;
    j = 1 + (t >> 1) % g.jm;
    k = 1 + (t >> 1) / g.jm;
  } else {
    return;
  }
  if (!sm_written(g, i, j, k)) return;
  float sgn = 1.0f;
  int kk = k;
  if (k == 0) {
    kk = 1;
    sgn = -1.0f;
  } else if (k == g.km + 1) {
    kk = g.km;
  }
  int jj = j;
  if (j == 0) jj = 1;
  else if (j == g.jm + 1) jj = g.jm;
  int ii = i;
  if (i == 0) ii = 1;
  else if (i == g.im + 1) ii = g.im;
  const float s = sm[ix(g, ii, jj, kk)];
  sm[ix(g, i, j, k)] = (sgn < 0.0f) ? -s : s;
}

// Viscous terms added to f (k=2..km, wall k=1), g (same) and h (k=1..km),
// les.f95:83-242.
// ADAM: the whole-step path also applies adam (adam.f95:30-44) to the new
// f,g,h, which is what k_adam after this kernel computes.
template <bool ADAM>
__device__ __forceinline__ void les_visc_body(const VBlk& vb, const Grid& g, const Metrics& m, const P10& diu,
                                              const float* __restrict__ sm, const float* __restrict__ uspd,
                                              const float* __restrict__ vspd, float* __restrict__ f,
                                              float* __restrict__ gg, float* __restrict__ h, float* __restrict__ fo,
                                              float* __restrict__ go, float* __restrict__ ho, float wl) {
  const int i = 1 + vb.x * BX + threadIdx.x, j = 1 + vb.y * BY + wave_row();
  if (i > g.im || j > g.jm) return;
  const int sj = g.sj, sk = g.sk;
  const float dxi = M1(m.dx1, i), dxp = M1(m.dx1, i + 1), dxm = M1(m.dx1, i - 1);
  const float dyj = M1(m.dy1, j), dyp = M1(m.dy1, j + 1), dym = M1(m.dy1, j - 1);
  const float dxs = M1(m.dxs, i), dys = M1(m.dys, j);
  const float* d1 = diu.a[1];
  const float* d2 = diu.a[2];
  const float* d3 = diu.a[3];
  const float* d4 = diu.a[4];
  const float* d5 = diu.a[5];
  const float* d6 = diu.a[6];
  const float* d7 = diu.a[7];
  const float* d8 = diu.a[8];
  const float* d9 = diu.a[9];
#define SMX(c) ((dxp * sm[(c)] + dxi * sm[(c) + 1]) / (dxi + dxp))      // x-interp at i+1/2
#define SMXM(c) ((dxi * sm[(c) - 1] + dxm * sm[(c)]) / (dxm + dxi))     // x-interp at i-1/2
#define SMY(c) ((dyp * sm[(c)] + dyj * sm[(c) + sj]) / (dyj + dyp))     // y-interp at j+1/2
#define SMYM(c) ((dyj * sm[(c) - sj] + dym * sm[(c)]) / (dym + dyj))    // y-interp at j-1/2
  int k0, k1;
  kspan(vb, 1, g.km, k0, k1);
  if (k0 > k1) return;
  // A thread walks its run of levels.  What level k shares with level k-1
  // is carried in registers, each the same expression on the same operands
  // as its recomputation would be: the sm interpolations at k (computed as
  // the k+1 ones of the level below) and the k-1/2 face products visuz1,
  // visvz1, viswz1 (= visuz2, visvz2, viswz2 of the level below, whose evsz2
  // is this level's evsz1).
  int c = ix(g, i, j, k0);
  float cX0 = SMX(c), cXM0 = SMXM(c), cY0 = SMY(c), cYM0 = SMYM(c);  // SMX, SMXM, SMY, SMYM at level k
  float cUz1 = 0.f, cVz1 = 0.f;                                      // visuz1, visvz1 (k >= 2)
  float cWz1 = (sm[c]) * 2.0f * d9[c];                                // viswz1
  if (k0 >= 2) {
    const float dzk = M1(m.dzn, k0), dzm = M1(m.dzn, k0 - 1);
    const float evsz1 = (dzk * SMX(c - sk) + dzm * cX0) / (dzm + dzk);
    cUz1 = (evsz1) * (d3[c] + d7[c + 1 - sk]);
    cVz1 = (evsz1) * (d6[c] + d8[c + sj - sk]);
  }
  for (int k = k0; k <= k1; k++, c += sk) {
    const float dzk = M1(m.dzn, k), dzp = M1(m.dzn, k + 1);
    const float Xp = SMX(c + sk), XMp = SMXM(c + sk), Yp = SMY(c + sk), YMp = SMYM(c + sk);  // level k+1
    const float evsz2 = (dzp * cX0 + dzk * Xp) / (dzk + dzp);
    float fv, gv, hv;
    float visuz2, visvz2, viswz2;
    // ---- f (les.f95:83-114; wall les.f95:118-139)
    {
      const float evsx2 = sm[c + 1];
      const float evsx1 = sm[c];
      const float evsy2 = (dyp * cX0 + dyj * SMX(c + sj)) / (dyj + dyp);
      const float evsy1 = (dyp * SMX(c - sj) + dyj * cX0) / (dyj + dyp);
      const float visux2 = (evsx2) * 2.0f * d1[c + 1];
      const float visux1 = (evsx1) * 2.0f * d1[c];
      const float visuy2 = (evsy2) * (d2[c + sj] + d4[c + 1]);
      const float visuy1 = (evsy1) * (d2[c] + d4[c + 1 - sj]);
      visuz2 = (evsz2) * (d3[c + sk] + d7[c + 1]);
      float vfu;
      if (k >= 2) {
        const float visuz1 = cUz1;
        vfu = (visux2 - visux1) / dxs + (visuy2 - visuy1) / dyj + (visuz2 - visuz1) / dzk;
      } else {
        const float t = 0.4f * uspd[ix(g, i, j, -1)] / wl;
        const float visuz1 = t * t * uspd[ix(g, i, j, -1)];
        vfu = (visux2 - visux1) / dxs + (visuy2 - visuy1) / dyj + (visuz2 - visuz1) / dzk +
              (visuy2 - visuy1) / dyj + (visuz2 - visuz1) / dzk;
      }
      fv = (f[c] + vfu);
    }
    // ---- g (les.f95:143-174; wall les.f95:178-201)
    {
      const float evsy2 = sm[c + sj];
      const float evsy1 = sm[c];
      const float evsx2 = (dyp * cX0 + dyj * SMX(c + sj)) / (dyj + dyp);
      const float evsx1 = (dyp * cXM0 + dyj * SMXM(c + sj)) / (dyj + dyp);
      const float visvx2 = (evsx2) * (d2[c + sj] + d4[c + 1]);
      const float visvx1 = (evsx1) * (d2[c - 1 + sj] + d4[c]);
      const float visvy2 = (evsy2) * 2.0f * d5[c + sj];
      const float visvy1 = (evsy1) * 2.0f * d5[c];
      visvz2 = (evsz2) * (d6[c + sk] + d8[c + sj]);
      float visvz1;
      if (k >= 2) {
        visvz1 = cVz1;
      } else {
        const float t = 0.4f * vspd[ix(g, i, j, -1)] / wl;
        visvz1 = t * t * vspd[ix(g, i, j, -1)];
      }
      const float vfv = (visvx2 - visvx1) / dxi + (visvy2 - visvy1) / dys + (visvz2 - visvz1) / dzk;
      gv = (gg[c] + vfv);
    }
    // ---- h (les.f95:206-242)
    {
      const float hevsz2 = sm[c + sk];
      const float evsx2 = evsz2;  // (dzp * SMX(c) + dzk * SMX(c + sk)) / (dzk + dzp)
      const float evsx1 = (dzp * cXM0 + dzk * XMp) / (dzk + dzp);
      const float evsy2 = (dzp * cY0 + dzk * Yp) / (dzk + dzp);
      const float evsy1 = (dzp * cYM0 + dzk * YMp) / (dzk + dzp);
      const float viswx2 = (evsx2) * (d3[c + sk] + d7[c + 1]);
      const float viswx1 = (evsx1) * (d3[c - 1 + sk] + d7[c]);
      const float viswy2 = (evsy2) * (d6[c + sk] + d8[c + sj]);
      const float viswy1 = (evsy1) * (d6[c - sj + sk] + d8[c]);
      viswz2 = (hevsz2) * 2.0f * d9[c + sk];
      const float viswz1 = cWz1;
      const float vfw = (viswx2 - viswx1) / dxi + (viswy2 - viswy1) / dyj + (viswz2 - viswz1) / dzk;
      hv = (h[c] + vfw);
    }
    // carry to level k+1
    cX0 = Xp;
    cXM0 = XMp;
    cY0 = Yp;
    cYM0 = YMp;
    cUz1 = visuz2;
    cVz1 = visvz2;
    cWz1 = viswz2;
    if (ADAM) {
      f[c] = 1.5f * fv - 0.5f * fo[c];
      gg[c] = 1.5f * gv - 0.5f * go[c];
      h[c] = 1.5f * hv - 0.5f * ho[c];
      fo[c] = fv;
      go[c] = gv;
      ho[c] = hv;
    } else {
      f[c] = fv;
      gg[c] = gv;
      h[c] = hv;
    }
  }
#undef SMX
#undef SMXM
#undef SMY
#undef SMYM
}
template <bool ADAM>
__global__ __launch_bounds__(NT) void k_les_visc(Grid g, Metrics m, P10 diu, const float* __restrict__ sm,
                                                 const float* __restrict__ uspd, const float* __restrict__ vspd,
                                                 float* __restrict__ f, float* __restrict__ gg,
                                                 float* __restrict__ h, float* __restrict__ fo,
                                                 float* __restrict__ go, float* __restrict__ ho, float wl) {
  les_visc_body<ADAM>(hw_blk<true>(), g, m, diu, sm, uspd, vspd, f, gg, h, fo, go, ho, wl);
}

static int boundsm_blocks(const Grid& g) {
  const int ni = g.im + 2, nj = g.jm + 3;
  return (2 * ni * nj + 3 * ni * g.km + 2 * g.jm * g.km + 255) / 256;
}

void launch_les(const DevDomain& d) {
  const Grid& g = d.g;
  const P10 diu = p10(d.F.diu);
  TIMED(d, KC_LES_SM, hipLaunchKernelGGL(k_les_sm, krun_grid(d, d.sm_kchunk), block2d(), 0, d.stream, g, diu,
                     d.m.delx1 + 1, d.F.sm));
  TIMED(d, KC_LES_VISC, hipLaunchKernelGGL(k_les_visc<false>, krun_grid(d, d.visc_kchunk), block2d(), 0, d.stream, g,
                     d.m, diu, d.F.sm, d.F.uspd, d.F.vspd, d.F.f, d.F.g, d.F.h, nullptr, nullptr, nullptr, d.wl));
}

// les followed by adam (whole-step path); with fuse_sm the fused velfg
// kernel has computed sm
void launch_les_adam(const DevDomain& d) {
  const Grid& g = d.g;
  const P10 diu = p10(d.F.diu);
  if (!d.fuse_sm)
    TIMED(d, KC_LES_SM, hipLaunchKernelGGL(k_les_sm, krun_grid(d, d.sm_kchunk), block2d(), 0, d.stream, g, diu,
                       d.m.delx1 + 1, d.F.sm));
  TIMED(d, KC_LES_VISC, hipLaunchKernelGGL(k_les_visc<true>, krun_grid(d, d.visc_kchunk), block2d(), 0, d.stream, g,
                     d.m, diu, d.F.sm, d.F.uspd, d.F.vspd, d.F.f, d.F.g, d.F.h, d.F.fold, d.F.gold, d.F.hold,
                     d.wl));
}

// ------------------------------------------------------------------- adam
// adam.f95:30-44.
__global__ __launch_bounds__(NT) void k_adam(Grid g, float* __restrict__ f, float* __restrict__ gg,
                                             float* __restrict__ h, float* __restrict__ fo,
                                             float* __restrict__ go, float* __restrict__ ho) {
  const int i = 1 + blockIdx.x * BX + threadIdx.x, j = 1 + blockIdx.y * BY + threadIdx.y;
  if (i > g.im || j > g.jm) return;
  int k0, k1;
  kspan(1, g.km, k0, k1);
  for (int k = k0; k <= k1; k++) {
    const int c = ix(g, i, j, k);
    const float fd = f[c], gd = gg[c], hd = h[c];
    f[c] = 1.5f * fd - 0.5f * fo[c];
    gg[c] = 1.5f * gd - 0.5f * go[c];
    h[c] = 1.5f * hd - 0.5f * ho[c];
    fo[c] = fd;
    go[c] = gd;
    ho[c] = hd;
  }
}

void launch_adam(const DevDomain& d) {
  const Grid& g = d.g;
  TIMED(d, KC_ADAM, hipLaunchKernelGGL(k_adam, grid2d(g.im, g.jm, kz(g)), block2d(), 0, d.stream, g, d.F.f, d.F.g,
                     d.F.h, d.F.fold, d.F.gold, d.F.hold));
}

// ---------------------------------------------------------- building masks
// feedbfm.f95:87-116 on the device: amask1=0 where zbm(i,j) > z2(k)+0.5*dzn(k)
// (top[k-1] holds that sum, evaluated in FP32 on the host), and
// b/c/dmask1 = 1 exactly there; ghost cells stay 0.
__global__ __launch_bounds__(NT) void k_masks(Grid g, const float* __restrict__ zbm, const float* __restrict__ top,
                                              float* __restrict__ bm, float* __restrict__ cm,
                                              float* __restrict__ dm) {
  const int i = 1 + blockIdx.x * BX + threadIdx.x, j = 1 + blockIdx.y * BY + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > g.im || j > g.jm) return;
  const float v = (zbm[(i - 1) + (j - 1) * g.im] > top[k - 1]) ? 1.0f : 0.0f;
  const int c = ix(g, i, j, k);
  bm[c] = v;
  cm[c] = v;
  dm[c] = v;
}

void launch_masks(const DevDomain& d, const float* zbm, const float* top) {
  const Grid& g = d.g;
  hipLaunchKernelGGL(k_masks, grid2d(g.im, g.jm, g.km), block2d(), 0, d.stream, g, zbm, top, d.F.bmask, d.F.cmask,
                     d.F.dmask);
}

// ------------------------------------------------------------- self-test
// Counts the x with bit patterns in [lo, hi) whose rcp_rn(x) differs from
// the correctly rounded 1.0f/x (one vector atomic per block with a miss).
__global__ __launch_bounds__(256) void k_rcp_check(unsigned lo, unsigned hi, unsigned* __restrict__ miss,
                                                  unsigned* __restrict__ first) {
  __shared__ unsigned cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  const unsigned long long stride = (unsigned long long)gridDim.x * 256;
  unsigned c = 0, f = 0xffffffffu;
  for (unsigned long long b = lo + (unsigned long long)blockIdx.x * 256 + threadIdx.x; b < hi; b += stride) {
    const float x = __uint_as_float((unsigned)b);
    const float a = rcp_rn(x), e = 1.0f / x;
    if (__float_as_uint(a) != __float_as_uint(e)) {
      c++;
      if ((unsigned)b < f) f = (unsigned)b;
    }
  }
  if (c) {
    atomicAdd(&cnt, c);
    atomicMin(first, f);
  }
  __syncthreads();
  if (threadIdx.x == 0 && cnt) atomicAdd(miss, cnt);
}

int selftest_reciprocal(unsigned lo, unsigned hi, unsigned long long* misses, unsigned* first_bits) {
  unsigned* buf = nullptr;
  if (hipMalloc(&buf, 2 * sizeof(unsigned)) != hipSuccess) return 1;
  const unsigned init[2] = {0u, 0xffffffffu};
  hipError_t e = hipMemcpy(buf, init, sizeof(init), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_rcp_check, dim3(8192), dim3(256), 0, 0, lo, hi, buf, buf + 1);
    e = hipGetLastError();
  }
  unsigned out[2] = {0u, 0u};
  if (e == hipSuccess) e = hipMemcpy(out, buf, sizeof(out), hipMemcpyDeviceToHost);
  hipFree(buf);
  if (e != hipSuccess) return 1;
  *misses = out[0];
  *first_bits = out[1];
  return 0;
}

// ------------------------------------------------------ layout conversion
__global__ __launch_bounds__(256) void k_unpack(Grid g, const float* __restrict__ src,
                                                float* __restrict__ dst, int li, int lj, int lk,
                                                int ni, int nj, int nk) {
  const long n = (long)ni * nj * nk;
  for (long t = blockIdx.x * 256L + threadIdx.x; t < n; t += (long)gridDim.x * 256) {
    const int a = (int)(t % ni), b = (int)((t / ni) % nj), c = (int)(t / ((long)ni * nj));
    dst[ix(g, li + a, lj + b, lk + c)] = src[t];
  }
}

__global__ __launch_bounds__(256) void k_pack(Grid g, const float* __restrict__ src,
                                              float* __restrict__ dst, int li, int lj, int lk,
                                              int ni, int nj, int nk) {
  const long n = (long)ni * nj * nk;
  for (long t = blockIdx.x * 256L + threadIdx.x; t < n; t += (long)gridDim.x * 256) {
    const int a = (int)(t % ni), b = (int)((t / ni) % nj), c = (int)(t / ((long)ni * nj));
    dst[t] = src[ix(g, li + a, lj + b, lk + c)];
  }
}

void launch_unpack(const DevDomain& d, const float* src, float* dst, int li, int lj, int lk, int ni,
                   int nj, int nk) {
  const long n = (long)ni * nj * nk;
  long nb = (n + 255) / 256;
  if (nb > 65536) nb = 65536;
  TIMED(d, KC_COPY, hipLaunchKernelGGL(k_unpack, dim3((unsigned)nb), dim3(256), 0, d.stream, d.g, src, dst, li, lj, lk,
                     ni, nj, nk));
}

void launch_pack(const DevDomain& d, const float* src, float* dst, int li, int lj, int lk, int ni,
                 int nj, int nk) {
  const long n = (long)ni * nj * nk;
  long nb = (n + 255) / 256;
  if (nb > 65536) nb = 65536;
  TIMED(d, KC_COPY, hipLaunchKernelGGL(k_pack, dim3((unsigned)nb), dim3(256), 0, d.stream, d.g, src, dst, li, lj, lk,
                     ni, nj, nk));
}

// ------------------------------------------------ batched step stages
// One launch per stage for all domains of a batched step (steps n > 1,
// velfg+feedbf and les+adam fused).  Each block finds its domain and its
// position in that domain's own grid, then runs the stage kernel's body with
// that domain's arguments: the work of every block, and so every value, is
// the same as in the domain's own launch.
#define BATCH_BLOCK(stage, xcd)                                     \
  int e;                                                            \
  VBlk vb;                                                          \
  if (!batch_block<xcd>(sp + (stage) * nd, nd, e, vb)) return;      \
  const BatchDom& B = bd[e];                                        \
  const DevDomain& d = B.d

__global__ __launch_bounds__(NT) void kb_velnw(const BatchDom* __restrict__ bd, const StageSpan* __restrict__ sp,
                                               int nd) {
  BATCH_BLOCK(BS_VELNW, true);
  velnw_body(vb, d.g, d.m, d.F.p, d.F.f, d.F.g, d.F.h, d.F.u, d.F.v, d.F.w, B.ro, B.dt);
}
__global__ __launch_bounds__(NT) void kb_bondv1_inflow(const BatchDom* __restrict__ bd,
                                                       const StageSpan* __restrict__ sp, int nd) {
  BATCH_BLOCK(BS_INFLOW, false);
  bondv1_inflow_body(vb, d.g, d.m, d.F.u, d.F.v, d.F.w, B.ilast, d.ctrl);
}
__global__ __launch_bounds__(256) void kb_bondv1_uout(const BatchDom* __restrict__ bd,
                                                      const StageSpan* __restrict__ sp, int nd) {
  BATCH_BLOCK(BS_UOUT, false);
  bondv1_uout_body(vb, d.g, d.F.u, d.ctrl);
}
__global__ __launch_bounds__(256) void kb_bondv1_outflow(const BatchDom* __restrict__ bd,
                                                         const StageSpan* __restrict__ sp, int nd) {
  BATCH_BLOCK(BS_OUTFLOW, false);
  bondv1_outflow_body(vb, d.g, d.m, d.F.u, d.F.v, d.F.w, d.ctrl, B.dt);
}
__global__ __launch_bounds__(256) void kb_bondv1_ghost(const BatchDom* __restrict__ bd,
                                                       const StageSpan* __restrict__ sp, int nd) {
  BATCH_BLOCK(BS_GHOST, false);
  bondv1_ghost_body(vb, d.g, d.F.u, d.F.v, d.F.w);
}
template <bool SM>
__global__ __launch_bounds__(NT) void kb_velfg_feedbf(const BatchDom* __restrict__ bd,
                                                      const StageSpan* __restrict__ sp, int nd) {
  BATCH_BLOCK(BS_VELFG, false);
  const VelCtx X{d.g, d.m, d.F.u, d.F.v, d.F.w, d.wl};
  velfg_feedbf_body<SM>(vb, X, p10(d.F.diu), d.F.f, d.F.g, d.F.h, d.F.usum, d.F.vsum, d.F.wsum, d.F.bmask,
                        d.F.cmask, d.F.dmask, d.mask_src, d.F.uspd, d.F.vspd, B.vn, B.alpha, B.dt, B.beta,
                        B.velfg_kc, d.m.delx1 + 1, d.F.sm);
}
__global__ __launch_bounds__(NT) void kb_les_sm(const BatchDom* __restrict__ bd, const StageSpan* __restrict__ sp,
                                                int nd) {
  BATCH_BLOCK(BS_LES_SM, true);
  les_sm_body<false>(vb, d.g, p10(d.F.diu), d.m.delx1 + 1, d.F.sm);
}
__global__ __launch_bounds__(256) void kb_boundsm(const BatchDom* __restrict__ bd, const StageSpan* __restrict__ sp,
                                                  int nd) {
  BATCH_BLOCK(BS_BOUNDSM, false);
  boundsm_body(vb, d.g, d.F.sm);
}
__global__ __launch_bounds__(NT) void kb_les_visc_adam(const BatchDom* __restrict__ bd,
                                                       const StageSpan* __restrict__ sp, int nd) {
  BATCH_BLOCK(BS_LES_VISC, true);
  les_visc_body<true>(vb, d.g, d.m, p10(d.F.diu), d.F.sm, d.F.uspd, d.F.vspd, d.F.f, d.F.g, d.F.h, d.F.fold,
                      d.F.gold, d.F.hold, d.wl);
}
#undef BATCH_BLOCK

static StageSpan span_of(dim3 g) { return {0, (int)g.x, (int)g.y, (int)g.z}; }

void step_batch_spans(const DevDomain& d, int nzb, BatchDom* e, StageSpan* sp) {
  const Grid& g = d.g;
  // velfg, les_sm and les_visc (whose threads walk runs of levels, carrying
  // what neighbouring levels share) take the batch's k split nzb; velnw keeps
  // one level per block
  const int nz = nzb > 0 ? std::min(nzb, g.km) : kz(g);
  e->ilast = 1;
  sp[BS_VELNW] = span_of(grid2d(g.im, g.jm, kz(g)));
  sp[BS_INFLOW] = span_of(bondv1_grid(g, 1, 0));
  sp[BS_UOUT] = span_of(bondv1_grid(g, 1, 1));
  sp[BS_OUTFLOW] = span_of(bondv1_grid(g, 1, 2));
  sp[BS_GHOST] = span_of(bondv1_grid(g, 1, 3));
  if (nzb > 0) {
    e->velfg_kc = (g.km + nz - 1) / nz;
    sp[BS_VELFG] = span_of(grid2d(g.im, g.jm, (g.km + e->velfg_kc - 1) / e->velfg_kc));
  } else {
    sp[BS_VELFG] = span_of(velfg_grid(d, &e->velfg_kc));
  }
  sp[BS_LES_SM] = span_of(nzb > 0 ? grid2d(g.im, g.jm, nz) : krun_grid(d, d.sm_kchunk));
  sp[BS_BOUNDSM] = span_of(dim3((unsigned)boundsm_blocks(g)));
  sp[BS_LES_VISC] = span_of(nzb > 0 ? grid2d(g.im, g.jm, nz) : krun_grid(d, d.visc_kchunk));
}

void launch_step_batch(const StageBatch& b, hipStream_t stream) {
  const BatchDom* bd = b.doms;
  const StageSpan* sp = b.spans;
  const int nd = b.nd;
  auto nb = [&](int s) { return dim3((unsigned)b.nblocks[s]); };
  hipLaunchKernelGGL(kb_velnw, nb(BS_VELNW), block2d(), 0, stream, bd, sp, nd);
  hipLaunchKernelGGL(kb_bondv1_inflow, nb(BS_INFLOW), block2d(), 0, stream, bd, sp, nd);
  hipLaunchKernelGGL(kb_bondv1_uout, nb(BS_UOUT), dim3(256), 0, stream, bd, sp, nd);
  hipLaunchKernelGGL(kb_bondv1_outflow, nb(BS_OUTFLOW), dim3(256), 0, stream, bd, sp, nd);
  hipLaunchKernelGGL(kb_bondv1_ghost, nb(BS_GHOST), dim3(256), 0, stream, bd, sp, nd);
  if (b.fuse_sm) {
    hipLaunchKernelGGL(kb_velfg_feedbf<true>, nb(BS_VELFG), block2d(), 0, stream, bd, sp, nd);
  } else {
    hipLaunchKernelGGL(kb_velfg_feedbf<false>, nb(BS_VELFG), block2d(), 0, stream, bd, sp, nd);
    hipLaunchKernelGGL(kb_les_sm, nb(BS_LES_SM), block2d(), 0, stream, bd, sp, nd);
    hipLaunchKernelGGL(kb_boundsm, nb(BS_BOUNDSM), dim3(256), 0, stream, bd, sp, nd);
  }
  hipLaunchKernelGGL(kb_les_visc_adam, nb(BS_LES_VISC), block2d(), 0, stream, bd, sp, nd);
}

}  // namespace les
