// Device-resident domain descriptor and the stage launchers of the LES
// step library.  Each launcher enqueues its kernels on the domain's stream;
// none of them synchronises with the host.
#pragma once
#include "les_common.h"

namespace les {

struct DevFields {
  float *u, *v, *w, *p;           // prognostic
  float *f, *g, *h;               // tendencies
  float *fold, *gold, *hold;      // previous tendencies (AB2)
  float *usum, *vsum, *wsum;      // IBM integral feedback
  float *rhs, *sm;                // Poisson rhs, eddy viscosity
  float *uspd, *vspd;             // 2-D wall-function speeds (0:ip+1,0:jp+1)
  float* diu[10];                 // diu1..diu9 (index 0 unused)
  float* cov[10];                 // cov1..cov9 (index 0 unused)
  float *bmask, *cmask, *dmask;   // IBM masks
  float* cn1;                     // Poisson diagonal (1:ip,1:jp,1:kp)
};

// Colour-split SOR storage: pair t of row (j,k) holds i = 2t+1, 2t+2; one
// array per colour, index t + (j-1)*cp + (k-1)*cplane.
struct SplitGrid {
  int nt;      // pairs per row, (im+1)/2
  int cp;      // row pitch (floats)
  int cplane;  // plane pitch (floats)
};
struct SplitBufs {
  // [iterate set][colour: 0 black, 1 red].  One iteration per launch uses
  // sets 0 and 1 by iteration parity; two iterations per launch rotate over
  // all three (set 2 shares the storage of cn1, which is then recomputed).
  float* p[3][2];
  float* rhs[2];        // [colour] the rhs the SOR reads (its mean removed)
  float* cn1[2];        // [colour]; unused while cn1 is recomputed
  float* rhsr[2];       // [colour] two iterations per launch: rhs before its mean is removed
  double* part[2][2];   // [iterate parity][colour] per-block sums of reltmp^2
  float* base;          // start of the allocation: a scratch plane, the ten arrays, a scratch plane
  int nplanes;          // planes in the allocation (10 km + 2)
};

// Kernel classes for the optional per-launch timing (les_domain_set_timing).
enum KClass {
  KC_VELNW, KC_BONDV1, KC_VEL2, KC_VEL2_GHOST, KC_VELFG, KC_FEEDBF, KC_LES_SM, KC_BOUNDSM,
  KC_LES_VISC, KC_ADAM, KC_RHS, KC_SOR, KC_BOUNDP, KC_PAV, KC_PADJ, KC_COPY,
  KC_SOR1,  // SOR iteration 1 of a press call (either kernel; KC_SOR: iterations 2..nmaxp)
  KC_N
};

struct DevDomain {
  Grid g;
  Metrics m;
  DevFields F;
  Ctrl* ctrl;
  double* partials;  // per-block partials of the reductions outside the SOR: sums, then per-block flags
  int npartials;     // slots before the SOR's partial slots (twice the largest reduction grid, at least 512)
  hipStream_t stream;
  float wl;          // logf(0.5*dzn(1)/0.1), the wall-function log (vel2.f95:138, les.f95:133)
  double area;       // sum of dx1*dy1*dzn over the interior (press.f95:63-72)
  int nmaxp;         // SOR iteration cap (press.f95:44)
  float pjuge;       // SOR stop threshold on sum(reltmp^2) (press.f95:43)
  float omega;       // SOR relaxation (press.f95:45)
  void* timer;       // host-side event recorder, nullptr when timing is off
  int cn1_separable; // cn1 equals 1/(cnax+cnby+cncz) bit for bit: recompute instead of reading it
  SplitGrid sg;
  SplitBufs sb;
  int sor_blocks;    // target number of blocks of one SOR sweep
  int xcd_remap;     // give each XCD a contiguous range of SOR tiles
  int sor_fused;     // one launch per SOR iteration (k_sor_iter) instead of one per colour
  int fuse_velfg;    // whole steps run vel2+velfg+feedbf as one kernel
  int velfg_kchunk;  // k levels per thread of that kernel (0: automatic)
  int sm_kchunk;     // k levels per thread of k_les_sm (0: automatic)
  int visc_kchunk;   // k levels per thread of k_les_visc (0: automatic)
  int fuse_les;      // whole steps apply adam inside the les viscous kernel
  int fuse_sm;       // whole steps compute sm (and its ghosts) inside the fused velfg kernel
  int press_mode;    // 0: the reference's SOR; 1: FP64 defect correction around FP32 SOR
  int dc_outer;      // defect-correction steps per press call
  int dc_inner;      // FP32 SOR iterations per defect-correction step
  float dc_omega;    // relaxation of those inner iterations (1: the reference's)
  int n3;            // doubles per padded field (FP64 buffers of press mode 1)
  double* dc64;      // P64, R64 (n3 each) and the defect history (kDcMaxOuter)
  int sor_fused_blocks;  // target number of blocks of one fused SOR iteration (0: automatic)
  int ncu;           // compute units of the device
  int sor_rows;      // fused SOR kernel: 1 k_sor_iter, 2 k_sor_iter2 (one iteration per launch), 4 k_sor_pair (two)
  int sor_kchunk;    // k levels per block of the fused SOR iteration (0: from sor_fused_blocks)
  int sor_tile_rows; // tile rows (with the ring) of k_sor_iter2: 16 or 32
  int sor_tile_width;  // tile width in pairs (with the ring) of k_sor_iter2: 64, or 32 with 32 rows
  int sor_entry_fast;  // SOR iteration 1 on k_sor_iter2 when p's ghosts equal their boundp copies
  // masks of buildings standing on the ground, as a value per column at
  // (i-1)+(j-1)*im and km level thresholds (mask = value > threshold): the
  // height map zbm for masks built from it, the top solid level for given
  // masks of that form; nullptr for other masks (the fields are read)
  const float* mask_src;
};

int sor_kchunk(const Grid& g, const SplitGrid& s, int target);
int sor_parts_needed(const Grid& g, const SplitGrid& s, int target);
int sor_fused_kchunk(const Grid& g, const SplitGrid& s, int target, int ncu);
int sor_fused_parts_needed(const Grid& g, const SplitGrid& s, int target, int ncu);
int sor_fused_tiles(const Grid& g, const SplitGrid& s);
// the two-iterations-per-launch SOR (32 x 64 tiles with a four-cell ring)
int sor_pair_kchunk(const Grid& g, const SplitGrid& s, int ncu);
int sor_pair_tiles(const Grid& g, const SplitGrid& s);

// One domain of a batched SOR iteration (k_sor_iter2_batch): its arguments
// of k_sor_iter2<true, 16> and its range of blocks in the launch.
struct SorBatchEntry {
  Grid g;
  SplitGrid s;
  Metrics m;
  SplitBufs sb;  // with the cn1 pointers cleared (separable cn1)
  Ctrl* ctrl;
  int nparts;    // blocks of this domain
  int nparts1;   // blocks of its iteration 1 (the slots iteration 2's stop test reads)
  int ntt, ntj, kchunk, nmaxp;
  int block0;    // first block of this domain in the launch (a multiple of 8)
  float omega, pjuge;
};
// block0: the entries' block0 values, packed (the block-to-domain lookup);
// ty, tw: tile rows and width of the batch (16 x 64, 32 x 64 or 32 x 32);
// om1: every entry has omega == 1.
void launch_sor_batch(const SorBatchEntry* tab, const int* block0, int ntab, int nblocks, int iter, int kcmax,
                      int ty, int tw, bool om1, int remap, hipStream_t stream);
// The press stage in two parts around iterations 2..nmaxp of the SOR, for
// domains whose iterations run batched: part 0 is the right-hand side and
// iteration 1, part 1 the stop test of the last iteration and the mean
// removal.  kc, ty, tw: the SOR k chunk (0: the domain's own), tile rows
// and tile width of the batch.
void launch_press_part(const DevDomain& d, float dt, int part, int kc, int ty, int tw, bool pairs = false);
// The domain's entry in a batched SOR launch with k chunk kc, tile rows ty
// and tile width tw (block0 left to the caller).
void sor_batch_entry(const DevDomain& d, int kc, int ty, int tw, SorBatchEntry* e, bool pairs = false);
// The k chunk of a batch of domains' SOR launches with tile rows ty and
// tile width tw.
int sor_batch_kchunk(const DevDomain* const* ds, int n, int ty, int tw, int ncu);
// The batch's k chunk with two SOR iterations per launch (k_sor_pair tiles).
int sor_pair_batch_kchunk(const DevDomain* const* ds, int n, int ncu);
// Iterations 2..lmax of the batched domains two per launch (each domain up to
// its nmaxp; tables as built with sor_batch_entry(..., pairs = true)), then
// each domain's closing stop test and the fix-up launch.
void launch_sor_pair_batch(const SorBatchEntry* tab, const int* block0, int ntab, int nblocks, int lmax, int kcmax,
                           bool om1, int remap, hipStream_t stream);
// Whether a domain's SOR iterations can run in a batch (fused reference SOR,
// separable cn1, even jm).
bool sor_batchable(const DevDomain& d);

// The launches of a whole step (velfg+feedbf and les+adam fused, steps
// n > 1) that a batched step makes once for all its domains, in step order.
// One-block launches (the reduction finalizers, the SOR's closing sum) run
// one block per domain and need no span.
enum BatchStage {
  BS_VELNW, BS_INFLOW, BS_UOUT, BS_OUTFLOW, BS_GHOST, BS_VELFG, BS_LES_SM, BS_BOUNDSM, BS_LES_VISC,
  BS_RHS, BS_SPLIT, BS_SOR1, BS_PAV, BS_PADJ, BS_N
};
// A domain's grid in one stage and its first block in the batched launch (a
// multiple of 8, so every block runs on the XCD it has in a launch of its own).
struct StageSpan {
  int b0, gx, gy, gz;
};
// A domain of a batched step: its description and the scalars and shapes its
// own launches would use.
struct BatchDom {
  DevDomain d;
  SplitBufs sb;    // the split arrays with the cn1 pointers cleared (separable cn1)
  float ro, dt, vn, alpha, beta;
  int nred;        // partial-sum slots of the rhs and pav reductions
  int velfg_kc;    // k levels per thread of k_velfg_feedbf
  int ilast;       // last column bondv1 sets from the inflow (1 for n > 1)
  int ntt1, ntj1, kc1, nparts1;  // shape of SOR iteration 1 (k_sor_iter<true, true>)
  int nparts_last;               // partial-sum slots of the launch of iteration nmaxp
};
struct StageBatch {
  const BatchDom* doms;    // device, nd entries
  const StageSpan* spans;  // device, [BS_N][nd]
  int nd;
  int nblocks[BS_N];
  int kc1max;              // largest iteration-1 k chunk (dynamic LDS of that launch)
  int remap;
  int fuse_sm;             // sm and its ghost cells inside kb_velfg_feedbf (no kb_les_sm, kb_boundsm)
};
// This block's domain e and position in that domain's grid, in a batched
// launch over the spans sp[0..nd) (b0 ascending); false for the padding
// blocks after a domain's grid.  Block-uniform.  XCD: the positions are
// renumbered as hw_blk<true> does (b0 is a multiple of 8, so block l of a
// domain runs on XCD l % 8, as in a launch of its own).
template <bool XCD>
__device__ __forceinline__ bool batch_block(const StageSpan* __restrict__ sp, int nd, int& e, VBlk& vb) {
  const int b = (int)blockIdx.x;
  int q = 0;
  for (int r = 1; r < nd; r++)
    if (b >= sp[r].b0) q = r;
  const StageSpan s = sp[q];
  const int n = s.gx * s.gy * s.gz;
  int l = b - s.b0;
  if (l >= n) return false;
  if (XCD) {
    const int x8 = l % 8, base = n / 8, rem = n % 8;
    l = x8 * base + (x8 < rem ? x8 : rem) + l / 8;
  }
  vb = {l % s.gx, (l / s.gx) % s.gy, l / (s.gx * s.gy), s.gx, s.gy, s.gz};
  e = q;
  return true;
}
// A domain's shapes and grids (b0 left to the caller): the step stages, then
// the press stages.  nzb > 0: at most nzb blocks along k for velfg and the
// copies into and out of the split arrays (longer k runs per block; no
// value depends on the split); kc: the SOR k chunk of the batch.
void step_batch_spans(const DevDomain& d, int nzb, BatchDom* e, StageSpan* sp);
void press_batch_spans(const DevDomain& d, int kc, int ty, int tw, int nzb, BatchDom* e, StageSpan* sp);
// velnw .. les+adam of every domain, one launch per stage.
void launch_step_batch(const StageBatch& b, hipStream_t stream);
// part 0: rhs, its mean, the split arrays and SOR iteration 1; part 1 (after
// iterations 2..nmaxp): the closing stop-test sum, pav and the adjusted p.
void launch_press_batch(const StageBatch& b, int part, hipStream_t stream, bool pairs = false);
void launch_sor_trial(const DevDomain& d, int rows, int ty, int tw, int kc, int iters);
int reduction_parts_needed(const Grid& g);

// Event pairs around launches, recorded only when d.timer is set.
void timer_begin(const DevDomain& d, int cls);
void timer_end(const DevDomain& d, int cls);
#define TIMED(d, cls, ...)      \
  do {                          \
    timer_begin((d), (cls));    \
    __VA_ARGS__;                \
    timer_end((d), (cls));      \
  } while (0)

void launch_velnw(const DevDomain& d, float ro, float dt);
void launch_bondv1(const DevDomain& d, int n, float dt);
void launch_velfg(const DevDomain& d, float vn);
void launch_feedbf(const DevDomain& d, float alpha, float dt, float beta);
void launch_velfg_feedbf(const DevDomain& d, float vn, float alpha, float dt, float beta);
void launch_les_adam(const DevDomain& d);
int selftest_reciprocal(unsigned lo, unsigned hi, unsigned long long* misses, unsigned* first_bits);
void launch_masks(const DevDomain& d, const float* zbm, const float* top);
void launch_les(const DevDomain& d);
void launch_adam(const DevDomain& d);
void launch_press(const DevDomain& d, float dt);
// Copy the split rhs (its mean removed) into the reference-layout rhs, which
// press with two SOR iterations per launch does not write.
void launch_rhs_gather(const DevDomain& d);

// The time statistics of aveflow (aveflow.f95:8-23), device layout.
struct AveBufs {
  float *aveu, *avev, *avew, *avep, *avel;
  float *aveuu, *avevv, *aveww, *avesm, *avesmsm, *uwfx;
};
// One aveflow call (aveflow.f95:82-197): zero at n == n1, accumulate for
// n >= n1, turn sums into means, deviations and the u-w flux at n == nmax.
void launch_aveflow(const DevDomain& d, const AveBufs& a, int n, int n1, int nmax);

// layout conversion between a packed reference-layout staging buffer and
// the device layout: the reference array covers [lo, lo+n) per axis
void launch_unpack(const DevDomain& d, const float* src, float* dst, int li, int lj, int lk, int ni,
                   int nj, int nk);
void launch_pack(const DevDomain& d, const float* src, float* dst, int li, int lj, int lk, int ni,
                 int nj, int nk);

}  // namespace les
