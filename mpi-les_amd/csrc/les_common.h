// Shared device-side definitions of the LES step library (gfx950).
//
// Device layout: every 3-D field of a domain uses one layout that covers
// i,j,k = -1 .. n+2 in each direction, which contains every reference
// extent (u,v: (0:ip+1,-1:jp+1,0:kp+1); w: (0:ip+1,-1:jp+1,-1:kp+1);
// p: (0:ip+2,0:jp+2,0:kp+1); diu1: (-1:ip+2,0:jp+2,0:kp+2); sm:
// (-1:ip+1,-1:jp+1,0:kp+1); ... main.f95:63-189).  i is fastest; i = 1
// starts a 256-byte segment of its row (i = 0 and -1 end the segment before)
// and the row pitch sj is a multiple of 64 floats (256 B), so the 64 lanes of
// a wave on i = 1..64 (and on every later run of 64) touch two 128-byte lines,
// not three.  Cells a reference array does not have are never read by the
// kernels.  1-D metrics are stored with index x at slot x+1.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace les {

struct Grid {
  int im, jm, km;
  int sj;  // floats between consecutive j rows
  int sk;  // floats between consecutive k planes
};

constexpr int kIOrg = 63;  // row offset of i = 0

// Row pitch of an im-wide domain: i runs to im+2.
__host__ __device__ constexpr long row_pitch(long im) { return (im + 2 + kIOrg + 1 + 63) / 64 * 64; }

__host__ __device__ __forceinline__ int ix(const Grid& g, int i, int j, int k) {
  return (i + kIOrg) + (j + 1) * g.sj + (k + 1) * g.sk;
}

// 1-D metrics (slot x+1 holds index x).  Pointers into one device buffer.
struct Metrics {
  const float* dx1;     // dx1(-1:ip+1)
  const float* dy1;     // dy1(0:jp+1)
  const float* dzn;     // dzn(-1:kp+2)
  const float* dzs;     // dzs(-1:kp+2)
  const float* dxs;     // dxs(0:ip)
  const float* dys;     // dys(0:jp)
  const float* cn2l;    // cn2l(1:ip)
  const float* cn2s;
  const float* cn3l;    // cn3l(1:jp)
  const float* cn3s;
  const float* cn4l;    // cn4l(1:kp)
  const float* cn4s;
  const float* delx1;   // delx1(1:kp), (dx1(0)*dy1(0)*dzn(k))**(1/3), les.f95:38-42
  const float* cnax;    // 2/(dxs(i-1)*dxs(i)), the three terms of cn1 (init.f95:100-107)
  const float* cnby;    // 2/(dys(j-1)*dys(j))
  const float* cncz;    // 2/(dzs(k-1)*dzs(k))
  const float* inflow;    // u_in(j,k) at (j-1)+(k-1)*jm, nullptr = 5 m/s
  const float* inflow_v;  // v_in(j,k), nullptr = 0 (bondv1_data24.f95:79-87 reads all three)
  const float* inflow_w;  // w_in(j,k), nullptr = 0
};
#define M1(a, x) ((a)[(x) + 1])

constexpr int kSorHist = 1024;
constexpr int kDcMaxOuter = 64;  // defect-correction steps per press call, at most

// Per-domain device scalars of the reductions and the SOR loop.
struct Ctrl {
  unsigned int ticket;     // last-workgroup counter of the running reduction
  unsigned int uout_max;   // bits of max(0, max u(im,:,:))   (>= +0)
  unsigned int uout_nmin;  // bits of -min(0, min u(im,:,:))  (>= +0)
  int sor_done;            // SOR iteration (1-based) whose residual fell below pjuge, 0 = none
  int sor_iters;           // iterations executed in the current press call
  int nonfinite;           // a non-finite residual was met
  int p_ghost_diff;        // a ghost cell of p differs from the copy boundp would make (set by the rhs split)
  int p_final;             // iterate set holding the SOR result (-1: the parity of sor_iters)
  int fix_iter;            // two-iteration launches: iteration to recompute alone (0: none) ...
  int fix_in, fix_out;     // ... from iterate set fix_in into set fix_out
  int dc_done;             // press mode 1: the defect met the stop threshold (later outer steps skip)
  int dc_norms;            // press mode 1: defect norms recorded in this press call
  double rhsav;            // volume-weighted mean of rhs
  double pav;              // volume-weighted mean of p
  double sor_first;        // sum of reltmp^2 of the current iteration's first colour
  double rhsav64;          // volume-weighted mean of the FP64 rhs (press mode 1)
  double sor_hist[kSorHist];
};

constexpr int kWave = 64;

// The row of a block whose rows are one wave wide (blockDim.x == 64): the
// same for every lane of a wave, so the compiler can keep it, and whatever
// derives from it, in scalar registers.
__device__ __forceinline__ int wave_row() { return __builtin_amdgcn_readfirstlane(threadIdx.y); }

// 1/x correctly rounded for positive normal x whose reciprocal is normal:
// the hardware reciprocal (within 1 ulp) refined by one Newton step with
// fused multiply-adds (the residual 1 - x*r is exact in FMA).  It equals
// the correctly rounded division 1.0f/x bit for bit over the whole range
// [2^-125, 2^125); les_selftest_reciprocal checks every such x on the
// device.
__device__ __forceinline__ float rcp_rn(float x) {
  const float r = __builtin_amdgcn_rcpf(x);
  const float e = __builtin_fmaf(-x, r, 1.0f);
  return __builtin_fmaf(r, e, r);
}

// The levels lo..hi are split into gridDim.z contiguous runs; block z gets
// [k0, k1] (empty when k0 > k1).  A thread walking consecutive levels finds
// the level-(k-1) and level-(k+1) neighbours of level k in its own CU's
// caches.
__device__ __forceinline__ void kspan(int lo, int hi, int& k0, int& k1) {
  const int n = hi - lo + 1;
  const int c = (n + (int)gridDim.z - 1) / (int)gridDim.z;
  k0 = lo + (int)blockIdx.z * c;
  k1 = (k0 + c - 1 < hi) ? k0 + c - 1 : hi;
}

// A block's position (x, y, z) in its domain's grid (nx, ny, nz).  The stage
// kernels of a whole step take it in place of blockIdx/gridDim, so that one
// launch can run the grids of several domains (each domain keeps the grid it
// has in a launch of its own, hence the same per-block work and partials).
struct VBlk {
  int x, y, z, nx, ny, nz;
};
// XCD = false: the hardware position.  XCD = true: workgroups go to the 8
// XCDs round-robin (workgroup b of the dispatch order x, y, z to XCD b % 8)
// and each XCD has its own L2, so the positions are renumbered to give each
// XCD one contiguous range of the grid (x fastest, then y, then z): a tile's
// neighbours in j and k then mostly run on the same XCD and share its L2.
// Measured at 300x300x90: velnw 78 -> 67 us, les_visc 214 -> 195 us,
// les_sm 95 -> 92 us, the rhs kernels the same time with a quarter less
// traffic; padj 23 -> 36 us and velfg unchanged, so those two keep the
// hardware order.  The work of a position, and so every value and partial
// sum (partials are indexed by position), is the same either way.
template <bool XCD>
__device__ __forceinline__ VBlk hw_blk() {
  const int nx = gridDim.x, ny = gridDim.y, nz = gridDim.z;
  if (!XCD) return {(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z, nx, ny, nz};
  const int n = nx * ny * nz;
  const int b = blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z);
  const int x8 = b % 8, l = b / 8, base = n / 8, rem = n % 8;
  const int L = x8 * base + (x8 < rem ? x8 : rem) + l;
  return {L % nx, (L / nx) % ny, L / (nx * ny), nx, ny, nz};
}
__device__ __forceinline__ int vlinear(const VBlk& b) { return b.x + b.nx * (b.y + b.ny * b.z); }
__device__ __forceinline__ void kspan(const VBlk& b, int lo, int hi, int& k0, int& k1) {
  const int n = hi - lo + 1;
  const int c = (n + b.nz - 1) / b.nz;
  k0 = lo + b.z * c;
  k1 = (k0 + c - 1 < hi) ? k0 + c - 1 : hi;
}

// Deterministic sum over a block of NT threads (tid = linear thread id):
// wavefront butterfly, then wave totals through LDS added by thread 0 in
// wave order.  The result is valid in thread 0.
template <int NT>
__device__ __forceinline__ double block_sum(double v, int tid, double* lds) {
  static_assert(NT % kWave == 0, "block must be whole wavefronts");
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  if (tid % kWave == 0) lds[tid / kWave] = v;
  __syncthreads();
  double s = 0.0;
  if (tid == 0)
    for (int w = 0; w < NT / kWave; w++) s += lds[w];
  __syncthreads();
  return s;
}

// Finish a grid-wide deterministic sum: every block stores its partial,
// the last block to arrive (ticket) adds all partials in index order.
// Returns true in thread 0 of the last block, with *total set.
template <int NT>
__device__ __forceinline__ bool grid_sum(double block_total, int tid, int bid, int nblocks,
                                         double* partials, unsigned int* ticket, double* total,
                                         double* lds) {
  __shared__ int last;
  if (tid == 0) {
    __hip_atomic_store(&partials[bid], block_total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned int t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    last = (t == (unsigned int)nblocks - 1u);
  }
  __syncthreads();
  if (!last) return false;
  double s = 0.0;
  for (int b = tid; b < nblocks; b += NT)
    s += __hip_atomic_load(&partials[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  s = block_sum<NT>(s, tid, lds);
  if (tid == 0) {
    *total = s;
    __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return tid == 0;
}

}  // namespace les
