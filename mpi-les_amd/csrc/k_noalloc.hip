// press (press.f95:9-220) on caller-owned device arrays, allocating nothing:
// les_press_noalloc.  For callers that cannot allocate device memory (code
// under HIP graph capture, serving loops that forbid hipMalloc).  Every
// array is a device pointer in the reference layout (the bounds of the
// Fortran dummies, extents from im, jm, km); the partial sums of the three
// reductions and the control block live in `scratch`, a static device
// array of fixed size (kNaScratchBytes) that each GPU holds until the
// process ends.  Calls that overlap on one GPU share it and race: the caller
// orders them (one stream, or events).
//
// The arithmetic is that of les_press, term for term: bondfg and the
// right-hand side with its FP64 volume mean (the same block partition and
// summation order as k_rhs and k_finalize, so rhsav has the same bits
// wherever the FP64 volume sum is exact, as on every grid of power-of-two
// dx1*dy1 with FP32 dzn), the red-black SOR in place with boundp1 after each
// colour and boundp2 after both (press.f95:89-122, boundp.f95:7-78; a
// colour's update reads only the other colour and ghosts, so in place gives
// the iterates of the reference bit for bit), the stop test on
// sum(reltmp^2) in FP64 after each iteration, then the pressure mean and its
// removal with boundp1/boundp2 (press.f95:124-162).
#include <mutex>
#include <string>

#include "../../include/les_amd.h"
#include "les_kernels.h"

namespace les {

int report_error(const std::string& msg);  // runtime.hip: les_last_error's message
int press_config_values(int* nmaxp, float* pjuge, float* omega);

namespace {

constexpr int kNaBX = 64, kNaBY = 4, kNaNT = kNaBX * kNaBY;
constexpr int kNaMaxParts = 8192;     // block partials per reduction launch, at most
constexpr int kNaMaxRedBlocks = 2048;  // the k split of k_rhs and k_pav (kMaxRedBlocks)
constexpr size_t kNaCtrlBytes = (sizeof(Ctrl) + 255) / 256 * 256;
// control block, then four partial arrays: value, volume, red, black
constexpr size_t kNaScratchBytes = kNaCtrlBytes + 4 * kNaMaxParts * sizeof(double);

static_assert(kNaScratchBytes == 270592, "the size stated in include/les_amd.h");

static __device__ __align__(256) unsigned char scratch[kNaScratchBytes];

// Reference layouts (main.f95:63-189): lower bounds and extents per axis.
struct RefA {
  int li, lj, lk, ni, nj;
  __device__ __forceinline__ int operator()(int i, int j, int k) const {
    return (i - li) + ni * ((j - lj) + nj * (k - lk));
  }
};
struct NaArgs {
  int im, jm, km;
  RefA U, W, P, F, R, I;
  float *rhs, *f, *g, *h, *p;
  const float *u, *v, *w, *dx1, *dy1, *dzn, *cn1, *cn2l, *cn2s, *cn3l, *cn3s, *cn4l, *cn4s;
};

__device__ __forceinline__ void na_kspan(int lo, int hi, int& k0, int& k1) {
  const int n = hi - lo + 1;
  const int c = (n + (int)gridDim.z - 1) / (int)gridDim.z;
  k0 = lo + (int)blockIdx.z * c;
  k1 = (k0 + c - 1 < hi) ? k0 + c - 1 : hi;
}
__device__ __forceinline__ int na_bid() { return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z); }

// bondfg (bondFG.f95:18-42) folded into the divergence, rhs (press.f95:
// 49-60), block partials of dx1*dy1*dzn*rhs and of dx1*dy1*dzn
// (press.f95:63-72).  Resets the control block.
__global__ __launch_bounds__(kNaNT) void knr_rhs(NaArgs a, Ctrl* ctrl, double* part, double* vol, float dt) {
  __shared__ double lds[kNaNT / kWave];
  const int tid = threadIdx.y * kNaBX + threadIdx.x;
  const int bid = na_bid();
  if (bid == 0 && tid == 0) {
    ctrl->sor_done = 0;
    ctrl->sor_iters = 0;
    ctrl->nonfinite = 0;
  }
  const int i = 1 + blockIdx.x * kNaBX + threadIdx.x, j = 1 + blockIdx.y * kNaBY + threadIdx.y;
  double acc = 0.0, av = 0.0;
  if (i <= a.im && j <= a.jm) {
    const float dxi = a.dx1[i + 1], dyj = a.dy1[j];
    int k0, k1;
    na_kspan(1, a.km, k0, k1);
    for (int k = k0; k <= k1; k++) {
      const float dzk = a.dzn[k + 1];
      const float fc = a.f[a.F(i, j, k)], gc = a.g[a.F(i, j, k)];
      const float fm = (i == 1) ? fc : a.f[a.F(i - 1, j, k)];
      const float gm = (j == 1) ? a.g[a.F(i, a.jm, k)] : a.g[a.F(i, j - 1, k)];
      const float hc = (k == a.km) ? 0.0f : a.h[a.F(i, j, k)];
      const float hm = (k == 1) ? 0.0f : a.h[a.F(i, j, k - 1)];
      if (i == 1) a.f[a.F(0, j, k)] = fc;
      if (j == 1) a.g[a.F(i, 0, k)] = gm;
      if (k == 1) a.h[a.F(i, j, 0)] = 0.0f;
      if (k == a.km) a.h[a.F(i, j, a.km)] = 0.0f;
      const float r = (-a.u[a.U(i - 1, j, k)] + a.u[a.U(i, j, k)]) / dxi +
                      (-a.v[a.U(i, j - 1, k)] + a.v[a.U(i, j, k)]) / dyj +
                      (-a.w[a.W(i, j, k - 1)] + a.w[a.W(i, j, k)]) / dzk;
      const float rr = (fc - fm) / dxi + (gc - gm) / dyj + (hc - hm) / dzk + r / dt;
      a.rhs[a.R(i, j, k)] = rr;
      acc += (double)(dxi * dyj * dzk * rr);
      av += (double)(dxi * dyj * dzk);
    }
  }
  const double s = block_sum<kNaNT>(acc, tid, lds);
  const double sv = block_sum<kNaNT>(av, tid, lds);
  if (tid == 0) {
    part[bid] = s;
    vol[bid] = sv;
  }
}

// One block: the volume mean of the n partials (each total in k_finalize's
// order).  what 0: rhsav; 1: pav, and the SOR statistics into stats.
__global__ __launch_bounds__(256) void knr_mean(const double* part, const double* vol, int n, Ctrl* ctrl, int what,
                                                int* stats) {
  __shared__ double lds[4];
  double s = 0.0, v = 0.0;
  for (int b = threadIdx.x; b < n; b += 256) s += part[b];
  s = block_sum<256>(s, threadIdx.x, lds);
  for (int b = threadIdx.x; b < n; b += 256) v += vol[b];
  v = block_sum<256>(v, threadIdx.x, lds);
  if (threadIdx.x == 0) {
    if (what == 0) {
      ctrl->rhsav = s / v;
    } else {
      ctrl->pav = s / v;
      if (stats) {
        stats[0] = ctrl->sor_iters;
        stats[1] = ctrl->nonfinite;
      }
    }
  }
}

// rhs = rhs - rhsav on the interior (press.f95:81-87; FP64 subtraction).
__global__ __launch_bounds__(kNaNT) void knr_rhs_sub(NaArgs a, const Ctrl* ctrl) {
  const int i = 1 + blockIdx.x * kNaBX + threadIdx.x, j = 1 + blockIdx.y * kNaBY + threadIdx.y;
  if (i > a.im || j > a.jm) return;
  const double av = ctrl->rhsav;
  int k0, k1;
  na_kspan(1, a.km, k0, k1);
  for (int k = k0; k <= k1; k++) {
    float* r = &a.rhs[a.R(i, j, k)];
    *r = (float)((double)*r - av);
  }
}

// True when SOR iteration `iter` is not to run: an earlier iteration met
// the stop test (press.f95:119-121).
__device__ __forceinline__ bool na_stopped(const Ctrl* ctrl, int iter) {
  const int done = ctrl->sor_done;
  return done != 0 && done < iter;
}

// One colour sweep of SOR iteration `iter` in place (press.f95:92-104).
// COL 1: the red cells (i+j+k odd, nrd = 0, swept first); 0: black.  Thread
// (t, j) updates the cell of the colour in pair t (i = 2t+1, 2t+2) at every
// level of its k run; block partials of reltmp^2 in FP64 (the squares are
// exact).
template <int COL>
__global__ __launch_bounds__(kNaNT) void knr_sweep(NaArgs a, const Ctrl* ctrl, double* part, int iter,
                                                   float omega) {
  __shared__ double lds[kNaNT / kWave];
  if (na_stopped(ctrl, iter)) return;
  const int tid = threadIdx.y * kNaBX + threadIdx.x;
  const int t = blockIdx.x * kNaBX + threadIdx.x, j = 1 + blockIdx.y * kNaBY + threadIdx.y;
  double acc = 0.0;
  if (t < (a.im + 1) / 2 && j <= a.jm) {
    const float c3l = a.cn3l[j - 1], c3s = a.cn3s[j - 1];
    int k0, k1;
    na_kspan(1, a.km, k0, k1);
    for (int k = k0; k <= k1; k++) {
      // red: i+j+k odd; black: even
      const int i = 2 * t + 1 + ((j + k + (COL ? 0 : 1)) & 1);
      if (i > a.im) continue;
      const int c = a.P(i, j, k);
      const float pc = a.p[c];
      const float reltmp =
          omega * (a.cn1[a.I(i, j, k)] * (a.cn2l[i - 1] * a.p[a.P(i + 1, j, k)] + a.cn2s[i - 1] * a.p[a.P(i - 1, j, k)] +
                                         c3l * a.p[a.P(i, j + 1, k)] + c3s * a.p[a.P(i, j - 1, k)] +
                                         a.cn4l[k - 1] * a.p[a.P(i, j, k + 1)] +
                                         a.cn4s[k - 1] * a.p[a.P(i, j, k - 1)] - a.rhs[a.R(i, j, k)]) -
                   pc);
      a.p[c] = pc + reltmp;
      acc = __builtin_fma((double)reltmp, (double)reltmp, acc);
    }
  }
  const double s = block_sum<kNaNT>(acc, tid, lds);
  if (tid == 0) part[na_bid()] = s;
}

// The ghost cells of p, each from the interior cell the reference's copies
// trace it to (i: 0 -> 1, im+1 -> im; j: 0 -> jm, jm+1 -> 1; k: 0 -> 1,
// km+1 -> km).  FULL false: boundp1 (boundp.f95:28-78), the cells with a
// ghost i or j at k = 0..km+1, from the same level; FULL true: boundp1 then
// boundp2 (boundp.f95:7-26), every ghost cell.  Sources are never written
// by the launch.  iter > 0: skipped once the SOR has stopped before iter.
template <bool FULL>
__global__ __launch_bounds__(kNaNT) void knr_bound(NaArgs a, const Ctrl* ctrl, int iter) {
  if (iter > 0 && na_stopped(ctrl, iter)) return;
  const int i = blockIdx.x * kNaBX + threadIdx.x, j = blockIdx.y * kNaBY + threadIdx.y;
  if (i > a.im + 1 || j > a.jm + 1) return;
  const bool gi = (i == 0 || i == a.im + 1), gj = (j == 0 || j == a.jm + 1);
  if (!FULL && !gi && !gj) return;
  const int ii = (i == 0) ? 1 : ((i == a.im + 1) ? a.im : i);
  const int jj = (j == 0) ? a.jm : ((j == a.jm + 1) ? 1 : j);
  int k0, k1;
  na_kspan(0, a.km + 1, k0, k1);
  for (int k = k0; k <= k1; k++) {
    const bool gk = (k == 0 || k == a.km + 1);
    if (!(gi || gj || (FULL && gk))) continue;
    const int kk = FULL ? ((k == 0) ? 1 : ((k == a.km + 1) ? a.km : k)) : k;
    a.p[a.P(i, j, k)] = a.p[a.P(ii, jj, kk)];
  }
}

// The stop test after iteration iter (press.f95:119-121): sum(reltmp^2) of
// both colours in FP64, recorded; a non-finite sum stops the loop too.
__global__ __launch_bounds__(256) void knr_stop(Ctrl* ctrl, const double* red, const double* black, int n,
                                                int iter, float pjuge) {
  __shared__ double lds[4];
  if (na_stopped(ctrl, iter)) return;
  double s = 0.0;
  for (int b = threadIdx.x; b < n; b += 256) s += red[b] + black[b];
  s = block_sum<256>(s, threadIdx.x, lds);
  if (threadIdx.x == 0) {
    if (iter <= kSorHist) ctrl->sor_hist[iter - 1] = s;
    ctrl->sor_iters = iter;
    if (!(s == s) || s > 1.0e300) ctrl->nonfinite = 1;
    if (s < (double)pjuge || ctrl->nonfinite) ctrl->sor_done = iter;
  }
}

// pav partials (press.f95:124-133): p*dx1*dy1*dzn and the volume, in
// k_pav's partition.
__global__ __launch_bounds__(kNaNT) void knr_pav(NaArgs a, double* part, double* vol) {
  __shared__ double lds[kNaNT / kWave];
  const int tid = threadIdx.y * kNaBX + threadIdx.x;
  const int i = 1 + blockIdx.x * kNaBX + threadIdx.x, j = 1 + blockIdx.y * kNaBY + threadIdx.y;
  double acc = 0.0, av = 0.0;
  if (i <= a.im && j <= a.jm) {
    const float dxi = a.dx1[i + 1], dyj = a.dy1[j];
    int k0, k1;
    na_kspan(1, a.km, k0, k1);
    for (int k = k0; k <= k1; k++) {
      const float dzk = a.dzn[k + 1];
      acc += (double)(a.p[a.P(i, j, k)] * dxi * dyj * dzk);
      av += (double)(dxi * dyj * dzk);
    }
  }
  const double s = block_sum<kNaNT>(acc, tid, lds);
  const double sv = block_sum<kNaNT>(av, tid, lds);
  if (tid == 0) {
    part[na_bid()] = s;
    vol[na_bid()] = sv;
  }
}

// p = p - pav on the interior (press.f95:145-150, FP64 subtraction).
__global__ __launch_bounds__(kNaNT) void knr_padj(NaArgs a, const Ctrl* ctrl) {
  const int i = 1 + blockIdx.x * kNaBX + threadIdx.x, j = 1 + blockIdx.y * kNaBY + threadIdx.y;
  if (i > a.im || j > a.jm) return;
  const double av = ctrl->pav;
  int k0, k1;
  na_kspan(1, a.km, k0, k1);
  for (int k = k0; k <= k1; k++) {
    float* q = &a.p[a.P(i, j, k)];
    *q = (float)((double)*q - av);
  }
}

int na_kblocks(int gx, int gy, int nk, int cap) {
  int gz = cap / (gx * gy);
  if (gz > nk) gz = nk;
  if (gz < 1) gz = 1;
  return gz;
}

// The device address of `scratch`, per device, taken once with that device
// current (the array's name is a host-side stand-in, never a device address).
std::mutex g_na_mu;
unsigned char* g_na_ptr[64] = {};

int na_scratch(unsigned char** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return report_error("les_press_noalloc: no device");
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (!g_na_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return report_error(std::string("les_press_noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_na_ptr[dev] = (unsigned char*)p;
  }
  *out = g_na_ptr[dev];
  return 0;
}

}  // namespace
}  // namespace les

using namespace les;

extern "C" int les_press_noalloc(int km, int jm, int im, float* rhs, const float* u, const float* dx1, const float* v,
                                 const float* dy1, const float* w, const float* dzn, float* f, float* g, float* h,
                                 float dt, const float* cn1, const float* cn2l, float* p, const float* cn2s,
                                 const float* cn3l, const float* cn3s, const float* cn4l, const float* cn4s, int n,
                                 int nmax, const float* usum, const float* vsum, const float* wsum, void* stream,
                                 int* stats) {
  (void)n;
  (void)nmax;
  (void)usum;
  (void)vsum;
  (void)wsum;
  if (im < 1 || jm < 1 || km < 1) return report_error("les_press_noalloc: extents must be positive");
  if (!rhs || !u || !dx1 || !v || !dy1 || !w || !dzn || !f || !g || !h || !cn1 || !cn2l || !p || !cn2s || !cn3l ||
      !cn3s || !cn4l || !cn4s)
    return report_error("les_press_noalloc: null array");
  const int gx = (im + kNaBX - 1) / kNaBX, gy = (jm + kNaBY - 1) / kNaBY;
  const int sx = ((im + 1) / 2 + kNaBX - 1) / kNaBX;
  const long cells_p = (long)(im + 3) * (jm + 3) * (km + 2);
  if ((long)gx * gy > kNaMaxParts || cells_p >= (1L << 31))
    return report_error("les_press_noalloc: domain too large for the fixed scratch (at most " +
                        std::to_string(kNaMaxParts) + " blocks of 64 x 4 columns, i.e. ceil(im/64)*ceil(jm/4) <= " +
                        std::to_string(kNaMaxParts) + ")");
  int nmaxp = 50;
  float pjuge = 1e-4f, omega = 1.0f;
  press_config_values(&nmaxp, &pjuge, &omega);
  unsigned char* base = nullptr;
  if (na_scratch(&base)) return 1;
  Ctrl* ctrl = (Ctrl*)base;
  double* part = (double*)(base + kNaCtrlBytes);
  double* vol = part + kNaMaxParts;
  double* pred = vol + kNaMaxParts;
  double* pblk = pred + kNaMaxParts;

  NaArgs a;
  a.im = im;
  a.jm = jm;
  a.km = km;
  a.U = {0, -1, 0, im + 2, jm + 3};
  a.W = {0, -1, -1, im + 2, jm + 3};
  a.P = {0, 0, 0, im + 3, jm + 3};
  a.F = {0, 0, 0, im + 1, jm + 1};
  a.R = {0, 0, 0, im + 2, jm + 2};
  a.I = {1, 1, 1, im, jm};
  a.rhs = rhs;
  a.f = f;
  a.g = g;
  a.h = h;
  a.p = p;
  a.u = u;
  a.v = v;
  a.w = w;
  a.dx1 = dx1;
  a.dy1 = dy1;
  a.dzn = dzn;
  a.cn1 = cn1;
  a.cn2l = cn2l;
  a.cn2s = cn2s;
  a.cn3l = cn3l;
  a.cn3s = cn3s;
  a.cn4l = cn4l;
  a.cn4s = cn4s;

  hipStream_t st = (hipStream_t)stream;
  const dim3 blk(kNaBX, kNaBY);
  const int gz = na_kblocks(gx, gy, km, kNaMaxRedBlocks);
  const int nred = gx * gy * gz;
  const int sz = na_kblocks(sx, gy, km, kNaMaxRedBlocks);
  const int nsw = sx * gy * sz;
  const int bx = (im + 2 + kNaBX - 1) / kNaBX, by = (jm + 2 + kNaBY - 1) / kNaBY;
  const int bz = std::min(km + 2, std::max(1, kNaMaxRedBlocks / (bx * by)));
  hipLaunchKernelGGL(knr_rhs, dim3(gx, gy, gz), blk, 0, st, a, ctrl, part, vol, dt);
  hipLaunchKernelGGL(knr_mean, dim3(1), dim3(256), 0, st, (const double*)part, (const double*)vol, nred, ctrl, 0,
                     (int*)nullptr);
  hipLaunchKernelGGL(knr_rhs_sub, dim3(gx, gy, gz), blk, 0, st, a, (const Ctrl*)ctrl);
  for (int l = 1; l <= nmaxp; l++) {
    hipLaunchKernelGGL((knr_sweep<1>), dim3(sx, gy, sz), blk, 0, st, a, (const Ctrl*)ctrl, pred, l, omega);
    hipLaunchKernelGGL((knr_bound<false>), dim3(bx, by, bz), blk, 0, st, a, (const Ctrl*)ctrl, l);
    hipLaunchKernelGGL((knr_sweep<0>), dim3(sx, gy, sz), blk, 0, st, a, (const Ctrl*)ctrl, pblk, l, omega);
    hipLaunchKernelGGL((knr_bound<true>), dim3(bx, by, bz), blk, 0, st, a, (const Ctrl*)ctrl, l);
    hipLaunchKernelGGL(knr_stop, dim3(1), dim3(256), 0, st, ctrl, (const double*)pred, (const double*)pblk, nsw, l,
                       pjuge);
  }
  hipLaunchKernelGGL(knr_pav, dim3(gx, gy, gz), blk, 0, st, a, part, vol);
  hipLaunchKernelGGL(knr_mean, dim3(1), dim3(256), 0, st, (const double*)part, (const double*)vol, nred, ctrl, 1,
                     stats);
  hipLaunchKernelGGL(knr_padj, dim3(gx, gy, gz), blk, 0, st, a, (const Ctrl*)ctrl);
  hipLaunchKernelGGL((knr_bound<true>), dim3(bx, by, bz), blk, 0, st, a, (const Ctrl*)ctrl, 0);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return report_error(std::string("les_press_noalloc: launch: ") + hipGetErrorString(e));
  return 0;
}
