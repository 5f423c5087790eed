// Output-side kernels that keep the reference's per-step consumers of the
// fields on the device: the time statistics of aveflow (aveflow.f95).
#include "les_kernels.h"

namespace les {

namespace {

constexpr int AX = 64, AY = 4;

// One pass over i=1..im+1, j=1..jm, k=0..km.  With `first` every sum
// starts from 0.0 (aveflow.f95:80-113) before this step's value is added,
// so the result is 0.0 + x, as in the reference.
__global__ __launch_bounds__(AX* AY) void k_ave_acc(Grid g, DevFields F, AveBufs A, int first) {
  const int i = 1 + blockIdx.x * AX + threadIdx.x, j = 1 + blockIdx.y * AY + threadIdx.y;
  const int k = blockIdx.z;
  if (i > g.im + 1 || j > g.jm) return;
  const int c = ix(g, i, j, k);
  auto acc = [first](float s, float x) { return (first ? 0.0f : s) + x; };
  A.avew[c] = acc(A.avew[c], F.w[c]);  // aveflow.f95:141-147, i=1..im+1, k=0..km
  if (i > g.im) return;
  A.aveu[c] = acc(A.aveu[c], F.u[c]);  // aveflow.f95:134-140, k=0..km
  if (k == 0) return;
  // aveflow.f95:117-131, k=1..km
  const float u = F.u[c], v = F.v[c], w = F.w[c], s = F.sm[c];
  A.avev[c] = acc(A.avev[c], v);
  A.avep[c] = acc(A.avep[c], F.p[c]);
  A.aveuu[c] = acc(A.aveuu[c], u * u);
  A.avevv[c] = acc(A.avevv[c], v * v);
  A.aveww[c] = acc(A.aveww[c], w * w);
  A.avesm[c] = acc(A.avesm[c], s);
  A.avesmsm[c] = acc(A.avesmsm[c], s * s);
  const int km1 = c - g.sk;
  A.uwfx[c] = acc(A.uwfx[c], 0.5f * (F.u[km1] + u) * 0.5f * (F.w[km1] + F.w[km1 + 1]));
  if (first) A.avel[c] = 0.0f;  // zeroed, never accumulated (aveflow.f95:86)
}

// Sums to means: x / float(nmax-n1+1) (aveflow.f95:151-182).
__global__ __launch_bounds__(AX* AY) void k_ave_div(Grid g, AveBufs A, float cnt) {
  const int i = 1 + blockIdx.x * AX + threadIdx.x, j = 1 + blockIdx.y * AY + threadIdx.y;
  const int k = blockIdx.z;
  if (i > g.im + 1 || j > g.jm) return;
  const int c = ix(g, i, j, k);
  A.avew[c] = A.avew[c] / cnt;
  if (i > g.im) return;
  A.aveu[c] = A.aveu[c] / cnt;
  if (k == 0) return;
  A.avev[c] = A.avev[c] / cnt;
  A.avep[c] = A.avep[c] / cnt;
  A.avel[c] = A.avel[c] / cnt;
  A.aveuu[c] = A.aveuu[c] / cnt;
  A.avevv[c] = A.avevv[c] / cnt;
  A.aveww[c] = A.aveww[c] / cnt;
  A.avesm[c] = A.avesm[c] / cnt;
  A.avesmsm[c] = A.avesmsm[c] / cnt;
}

// u-w flux minus the product of the means (aveflow.f95:185-192) and the
// standard deviations sqrt(|<x^2> - <x>^2|) (aveflow.f95:196-203).  Reads
// the means of aveu/avev/avew, which this kernel does not write.
__global__ __launch_bounds__(AX* AY) void k_ave_fin(Grid g, AveBufs A, float cnt) {
  const int i = 1 + blockIdx.x * AX + threadIdx.x, j = 1 + blockIdx.y * AY + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > g.im || j > g.jm) return;
  const int c = ix(g, i, j, k), km1 = c - g.sk;
  A.uwfx[c] = A.uwfx[c] / cnt - 0.5f * (A.aveu[km1] + A.aveu[c]) * 0.5f * (A.avew[km1] + A.avew[km1 + 1]);
  const float au = A.aveu[c], av = A.avev[c], aw = A.avew[c];
  A.aveuu[c] = sqrtf(fabsf(A.aveuu[c] - au * au));
  A.avevv[c] = sqrtf(fabsf(A.avevv[c] - av * av));
  A.aveww[c] = sqrtf(fabsf(A.aveww[c] - aw * aw));
}

}  // namespace

void launch_aveflow(const DevDomain& d, const AveBufs& a, int n, int n1, int nmax) {
  const Grid& g = d.g;
  const dim3 blk(AX, AY);
  const dim3 grd((g.im + 1 + AX - 1) / AX, (g.jm + AY - 1) / AY, g.km + 1);
  if (n >= n1) hipLaunchKernelGGL(k_ave_acc, grd, blk, 0, d.stream, g, d.F, a, (int)(n == n1));
  if (n == nmax) {
    const float cnt = (float)(nmax - n1 + 1);
    hipLaunchKernelGGL(k_ave_div, grd, blk, 0, d.stream, g, a, cnt);
    hipLaunchKernelGGL(k_ave_fin, dim3((g.im + AX - 1) / AX, (g.jm + AY - 1) / AY, g.km), blk, 0, d.stream, g, a,
                       cnt);
  }
}

}  // namespace les
