/*
 * les_amd.h -- C ABI of the MI355X (gfx950) LES step library.
 *
 * The library runs the per-time-step hot path of MPI-LES on one GPU:
 *   velnw -> bondv1 -> velfg (vel2) -> feedbf -> les (boundsm) -> adam ->
 *   press (bondfg, rhs, red-black SOR + boundp1/boundp2, mean removal)
 * with every field resident in device memory across steps.
 *
 * Two layers:
 *
 * 1. Reference-signature entry points (les_velnw ... les_press).  Each takes
 *    the argument list of the Fortran subroutine it replaces, in the same
 *    order, with every array passed as a pointer to its lower-bound corner
 *    in the reference's layout (Fortran order, explicit lower bounds,
 *    extents from ip,jp,kp == im,jm,km; ifdata.f95:238-241) and scalars by
 *    value.  A Fortran iso_c_binding shim module with the reference's module
 *    and subroutine names forwards to them unchanged.  The library keeps a
 *    device mirror of every array it is handed: the first call that sees a
 *    host address uploads its contents; after that the device copy is
 *    authoritative and the host array is stale until les_sync_host().
 *    Arrays the reference writes but never reads again (cov1..9, nou1..9,
 *    dfu1/dfv1/dfw1, fx/fy/fz) are accepted and left untouched.
 *
 * 2. Domain API (les_domain_*): create a domain from its set-up arrays,
 *    advance it by whole steps, read and write fields by reference name.
 *
 * Every function returns 0 on success and a nonzero status otherwise;
 * les_last_error() describes the last failure of the calling thread.
 * Calls are asynchronous on the domain's stream unless stated otherwise.
 */
#ifndef LES_AMD_H
#define LES_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* ----------------------------------------------------------- errors */
const char* les_last_error(void);

/* ------------------------------------- reference-signature entry points */

/* velnw.f95:5  subroutine velnw(km,jm,im,p,ro,dxs,u,dt,f,dys,v,g,dzs,w,h) */
int les_velnw(int km, int jm, int im, const float* p, float ro, const float* dxs, float* u, float dt,
              const float* f, const float* dys, float* v, const float* g, const float* dzs, float* w,
              const float* h);

/* bondv1.f95:5  subroutine bondv1(jm,u,z2,dzn,v,w,km,n,im,dt,dxs).
 * The inflow is the reference's 5 m/s (bondv1.f95:33-54) unless a profile
 * was set with les_bondv1_set_inflow for a domain of these extents. */
int les_bondv1(int jm, float* u, const float* z2, const float* dzn, float* v, float* w, int km, int n,
               int im, float dt, const float* dxs);

/* velFG.f95:6-8  subroutine velfg(km,jm,im,dx1,cov1,...,uspd,vspd) (calls vel2, vel2.f95:4) */
int les_velfg(int km, int jm, int im, const float* dx1, float* cov1, float* cov2, float* cov3,
              float* dfu1, float* diu1, float* diu2, const float* dy1, float* diu3, const float* dzn,
              float vn, float* f, float* cov4, float* cov5, float* cov6, float* dfv1, float* diu4,
              float* diu5, float* diu6, float* g, float* cov7, float* cov8, float* cov9, float* dfw1,
              float* diu7, float* diu8, float* diu9, const float* dzs, float* h, float* nou1,
              const float* u, float* nou5, const float* v, float* nou9, const float* w, float* nou2,
              float* nou3, float* nou4, float* nou6, float* nou7, float* nou8, float* uspd,
              float* vspd);

/* feedbf.f95:5-6  subroutine feedbf(km,jm,im,usum,u,bmask1,vsum,v,cmask1,wsum,w,dmask1,alpha,dt,beta,fx,fy,fz,f,g,h) */
int les_feedbf(int km, int jm, int im, float* usum, const float* u, const float* bmask1, float* vsum,
               const float* v, const float* cmask1, float* wsum, const float* w, const float* dmask1,
               float alpha, float dt, float beta, float* fx, float* fy, float* fz, float* f, float* g,
               float* h);

/* les.f95:6-7  subroutine les(km,delx1,dx1,dy1,dzn,jm,im,diu1,...,diu9,sm,f,g,h,uspd,vspd,dxs,dys)
 * (calls boundsm, boundsm.f95:5).  delx1 (intent out) is written on the host. */
int les_les(int km, float* delx1, const float* dx1, const float* dy1, const float* dzn, int jm, int im,
            const float* diu1, const float* diu2, const float* diu3, const float* diu4,
            const float* diu5, const float* diu6, const float* diu7, const float* diu8,
            const float* diu9, float* sm, float* f, float* g, float* h, const float* uspd,
            const float* vspd, const float* dxs, const float* dys);

/* adam.f95:5  subroutine adam(n,nmax,data21,fold,im,jm,km,gold,hold,fghold,f,g,h).
 * The character argument data21 names the file of the adam.f95:20-29 dump,
 * which stays on the Fortran side; it is not passed. */
int les_adam(int n, int nmax, float* fold, int im, int jm, int km, float* gold, float* hold,
             const float* fghold, float* f, float* g, float* h);

/* press.f95:9-10  subroutine press(km,jm,im,rhs,u,dx1,v,dy1,w,dzn,f,g,h,dt,cn1,cn2l,p,cn2s,
 * cn3l,cn3s,cn4l,cn4s,n,nmax,data20,usum,vsum,wsum) (calls bondfg, bondFG.f95:5, and
 * boundp1/boundp2, boundp.f95:7,28).  data20 (file name of the press.f95:205-217
 * dump) stays on the Fortran side. */
int les_press(int km, int jm, int im, float* rhs, const float* u, const float* dx1, const float* v,
              const float* dy1, const float* w, const float* dzn, float* f, float* g, float* h,
              float dt, const float* cn1, const float* cn2l, float* p, const float* cn2s,
              const float* cn3l, const float* cn3s, const float* cn4l, const float* cn4s, int n,
              int nmax, const float* usum, const float* vsum, const float* wsum);

/* press without allocation.  No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, serving
 * loops that forbid hipMalloc).  The argument list of les_press (press.f95:
 * 9-10) with every array a DEVICE pointer in the reference layout (extents
 * from im, jm, km), plus the stream to run on (a hipStream_t; NULL: the null
 * stream) and stats, a device int[2] that receives the SOR iteration count
 * and a non-finite flag (NULL: not written).  Asynchronous; nothing is
 * allocated and nothing waits.  nmaxp, pjuge and omega are those of
 * les_press_config.  The reductions' partial sums and control block live in
 * a fixed device array of 270 592 bytes per GPU, held until the process ends and
 * shared by every call on that GPU: calls that overlap on one GPU race, so
 * order them.  At most ceil(im/64)*ceil(jm/4) <= 8192 (im, jm up to 512 x
 * 4096, 1024 x 2048, ...); larger domains are refused.  Results are those of
 * les_press (p, rhs, f, g, h bit for bit where the FP64 volume sum of
 * dx1*dy1*dzn is exact, as for power-of-two dx1*dy1 and FP32 dzn); the
 * first call on a device must not be under graph capture (it looks up the
 * array's device address). */
int les_press_noalloc(int km, int jm, int im, float* rhs, const float* u, const float* dx1, const float* v,
                      const float* dy1, const float* w, const float* dzn, float* f, float* g, float* h,
                      float dt, const float* cn1, const float* cn2l, float* p, const float* cn2s,
                      const float* cn3l, const float* cn3s, const float* cn4l, const float* cn4s, int n,
                      int nmax, const float* usum, const float* vsum, const float* wsum, void* stream,
                      int* stats);

/* Replaces the values of the hard-coded inflow (bondv1.f95:33-54) with
 * u_in(j,k), j=1..jm fastest then k=1..km, for the domain of extents
 * (im,jm,km); prof == NULL restores 5 m/s.  No reference counterpart: the
 * reference's inflow is a literal. */
int les_bondv1_set_inflow(int im, int jm, int km, const float* prof);

/* The inflow planes of bondv1_data24 (bondv1_data24.f95:6, which reads
 * them from its data24_out<n+48000>.dat records, :43-68, and sets
 * u,v,w(i=0..1,j,k) from them, :79-87; the rest of that subroutine is
 * bondv1's).  u_in, v_in, w_in are (j,k), j=1..jm fastest then k=1..km;
 * v_in/w_in NULL mean 0, u_in NULL restores the 5 m/s of bondv1.  The next
 * les_bondv1 call on a domain of these extents uses them: setting the
 * planes each step and calling les_bondv1 is bondv1_data24. */
int les_bondv1_set_inflow_uvw(int im, int jm, int km, const float* u_in, const float* v_in,
                              const float* w_in);

/* aveflow.f95:5-6  subroutine aveflow(n,n1,km,jm,im,aveu,avev,avew,avep,avel,aveuu,avevv,
 * aveww,avesm,avesmsm,uwfx,avesu,...,amask1): the time statistics.
 * Arguments the statistics never touch (avesu..avesww, uwfxs, time,
 * amask1) and the file names data10..data14 are not passed.  The sums live
 * on the device; at n == nmax the final means, deviations and u-w flux are
 * copied into the host arrays, from which the caller writes data10/data11
 * (aveflow.f95:207-345).  Before that the host arrays are stale (les_sync_host
 * fetches them).  Layouts: aveu, avev (1:im,1:jm,0:km); avew
 * (1:im+1,1:jm,0:km+2); the rest (1:im,1:jm,1:km). */
int les_aveflow(int n, int n1, int km, int jm, int im, float* aveu, float* avev, float* avew, float* avep,
                float* avel, float* aveuu, float* avevv, float* aveww, float* avesm, float* avesmsm,
                float* uwfx, const float* u, const float* v, const float* w, const float* p,
                const float* sm, int nmax);

/* SOR controls fixed by parameters in the reference (press.f95:43-45):
 * nmaxp = 50, pjuge = 1e-4, omega = 1.  Applies to domains created after
 * the call and to reference-signature domains. */
int les_press_config(int nmaxp, float pjuge, float omega);
/* The values les_press_config set (the reference's parameters by default). */
int les_press_config_get(int* nmaxp, float* pjuge, float* omega);

/* Statistics of the last press call on the reference-signature domain of
 * extents (im,jm,km): SOR iterations, sum(reltmp^2) per iteration, the
 * volume-weighted means of rhs and p.  Synchronises. */
int les_press_stats(int im, int jm, int km, int* iters, double* sor_hist, int hist_len, double* rhsav,
                    double* pav, int* nonfinite);

/* Copies the device state of a host array handed to an earlier call back
 * into it.  Synchronises.  Returns 2 (and leaves the array alone) when no
 * call has received this address: its host contents are then current. */
int les_sync_host(const void* host_array);
/* The next call that receives this host array uploads it again. */
int les_mark_host_dirty(const void* host_array);
/* Waits for all work of every reference-signature domain. */
int les_synchronize(void);
/* Frees every reference-signature domain, forgets all host arrays, and
 * returns to the HIP runtime the device memory kept for reuse (below). */
int les_release_all(void);

/* --------------------------------------------------------- domain API */
typedef struct les_domain les_domain;

typedef struct {
  int im, jm, km;
  /* 1-D metrics in reference layout (grid.f95:5-172, init.f95:82-107):
   * dx1(-1:im+1) dy1(0:jm+1) dzn(-1:km+2) dzs(-1:km+2) dxs(0:im) dys(0:jm)
   * cn2l,cn2s(1:im) cn3l,cn3s(1:jm) cn4l,cn4s(1:km) */
  const float *dx1, *dy1, *dzn, *dzs, *dxs, *dys;
  const float *cn2l, *cn2s, *cn3l, *cn3s, *cn4l, *cn4s;
  /* cn1(1:im,1:jm,1:km); NULL selects the separable definition of
   * init.f95:100-107, 1/(2/(dxs dxs) + 2/(dys dys) + 2/(dz1 dz2)), which the
   * SOR evaluates on the fly (nothing is uploaded). */
  const float* cn1;
  /* Masks in reference layout, or all three NULL together with zbm and z2:
   * the masks are then built on the device by the rule of
   * feedbfm.f95:87-116 (1 inside buildings, zbm(i,j) > z2(k)+0.5*dzn(k),
   * 0 elsewhere and on every ghost cell), and the domain keeps zbm and the
   * level thresholds (im*jm + km floats), from which the whole step's
   * feedbf evaluates them instead of reading the mask fields.  Given masks
   * that are equal, 0 or 1, and solid from the ground up to a level in every
   * column are kept the same way (that level per column).  Uploading any
   * mask field afterwards drops that form: the step reads the fields from
   * then on. */
  const float* bmask1;   /* (-1:im+1,0:jm+1,0:km+1) */
  const float* cmask1;   /* (0:im+1,-1:jm+1,0:km+1) */
  const float* dmask1;   /* (0:im+1,0:jm+1,0:km+1) */
  const float* inflow;   /* u_in(j,k), NULL = 5 m/s */
  float ro, vn, alpha, beta, dt;   /* set.f95:89-95 */
  int nmaxp;
  float pjuge, omega;
  const float* zbm;      /* building heights zbm(1:im,1:jm), i fastest (feedbfm.f95:52-77) */
  const float* z2;       /* z2(0:km+2) (grid.f95:154-160) */
} les_domain_desc;

int les_domain_create(const les_domain_desc* desc, les_domain** out);
/* Waits for the domain's stream and releases it.  Its device memory is kept
 * for the next domains and batches (best fit, at most an eighth to spare)
 * instead of being freed: hipFree would wait for every stream of the device.
 * Kept memory goes back to the runtime when an allocation finds the device
 * full, and in les_release_all.  A domain must outlive the batches holding
 * it.  Creation, destruction and stepping of different domains may run on
 * different threads. */
int les_domain_destroy(les_domain* d);
/* One time step n (main.f95:260-296): the seven stages in order. */
int les_domain_step(les_domain* d, int n, int nmax);
/* Picks the launch shape of the fused SOR iteration (kernel: rows = 4 runs
 * two iterations per launch on 32 x 64 tiles with a four-cell ring (even jm,
 * separable cn1), rows = 2 or 1 one iteration with two or one tile rows per
 * thread; tile rows x pairs with the ring of the one-iteration kernel: 16 x
 * 64 or 32 x 64; k levels per block) for this grid shape: times
 * `reps` iterations of each candidate on the domain's SOR scratch arrays and
 * keeps the fastest.  With fused stages it also picks the k levels per
 * thread of velfg+feedbf, les_sm (when sm is not computed inside
 * velfg+feedbf) and les viscous+adam, timing `reps`
 * launches of each candidate on the domain's own fields, which it copies to
 * a temporary device allocation first and restores (les_domain_stencil_chunks
 * reports them).  The choices are remembered per (im, jm, km) for the
 * process; later calls with reps <= 0 reuse them without timing.  Results
 * do not depend on them.  Leaves the fields and the last press statistics
 * unchanged; synchronises.  Reports the SOR choice and the microseconds per
 * iteration measured, when the shape was timed, for it and for the untuned
 * default.  No reference counterpart: the per-shape autotuner of the serving
 * path. */
int les_domain_autotune(les_domain* d, int reps, int* rows, int* tile_rows, int* tile_width, int* kchunk,
                        double* us_best, double* us_default);
/* The k levels per thread of the fused stencil kernels (velfg+feedbf,
 * les_sm, les viscous+adam; 0 = automatic, and always 0 for les_sm when
 * velfg+feedbf computes sm) in use by the domain, as
 * les_domain_autotune picks them with the SOR shape (timing each candidate
 * on the domain's fields, which it restores), and the microseconds of one
 * velfg+feedbf plus one les+adam launch sequence with the tuned and with
 * the automatic chunks (0 when not timed).  Results do not depend on the
 * chunks.  No reference counterpart: part of the per-shape autotuner. */
int les_domain_stencil_chunks(les_domain* d, int* velfg_kchunk, int* sm_kchunk, int* visc_kchunk, double* us_best,
                              double* us_default);
/* Microseconds per fused SOR iteration with the domain's current launch
 * shape (a launch of two iterations counts as two; an odd iters is rounded
 * up to whole launches): `iters` back-to-back iterations (numbered 2.., never stopping) on
 * the domain's SOR scratch arrays between one pair of HIP events on its
 * stream.  Leaves the fields and the last press statistics unchanged;
 * synchronises.  No reference counterpart: the measurement of the dominant
 * kernel. */
int les_domain_time_sor(les_domain* d, int iters, double* us_per_iter);
/* With on != 0, steps n > 1 replay one captured HIP graph of the whole
 * step (recaptured when the domain's constants change); results are the
 * same as eager launches.  Off by default; ignored while timing is on. */
int les_domain_set_graph(les_domain* d, int on);
/* Advance steps n0..n1 inclusive. */
int les_domain_run(les_domain* d, int n0, int n1, int nmax);
/* Stage s of step n: 0 velnw, 1 bondv1, 2 velfg, 3 feedbf, 4 les, 5 adam, 6 press. */
int les_domain_stage(les_domain* d, int stage, int n, int nmax);
/* Field by reference name ("u","v","w","p","f","g","h","usum","vsum","wsum",
 * "fold","gold","hold","rhs","sm","uspd","vspd","diu1".."diu9","cn1",
 * "bmask1","cmask1","dmask1", and the aveflow statistics "aveu","avev",
 * "avew","avep","avel","aveuu","avevv","aveww","avesm","avesmsm","uwfx"),
 * host array in reference layout. */
int les_domain_upload(les_domain* d, const char* field, const float* host);
int les_domain_download(les_domain* d, const char* field, float* host);
int les_domain_synchronize(les_domain* d);
/* aveflow (aveflow.f95:80-203) on the domain's device state after step n:
 * zero at n == n1, accumulate for n >= n1, finalise at n == nmax.  Results
 * by les_domain_download("aveu", ...).  Does not synchronise. */
int les_domain_aveflow(les_domain* d, int n, int n1, int nmax);
/* Inflow planes for the domain's next steps, as les_bondv1_set_inflow_uvw.
 * The copy is ordered on the domain's stream (a step already enqueued keeps
 * the previous planes); the host arrays may be reused on return. */
int les_domain_set_inflow(les_domain* d, const float* u_in, const float* v_in, const float* w_in);
/* Pressure solver of the domain's press stage.  mode 0 (default): the
 * reference's SOR (press.f95:89-122).  mode 1: FP64 defect correction --
 * `outer` times, the defect d = (rhs - rhsav) - B(p) of the SOR's equation
 * is formed in FP64 and `inner` FP32 SOR iterations (from zero, no stop
 * test) solve B(e) = d, then p += e in FP64; p = p - pav in FP64 at the
 * end.  In mode 1 les_domain_press_stats reports the number of defect norms
 * recorded as the iteration count and sum((cn1*d)^2) before each outer
 * step's correction as the history.  No reference
 * counterpart: a precision variant (1 <= outer <= 64, 1 <= inner <= 1024). */
int les_domain_set_press_mode(les_domain* d, int mode, int outer, int inner);
/* Press mode 1: the relaxation of the inner FP32 SOR iterations (default 1,
 * the reference's; 0 < omega < 2).  Over-relaxation cuts the defect in far
 * fewer sweeps; the FP64 outer loop keeps the result that of the exact
 * equation.  The outer loop stops once sum((cn1*d)^2), the reference's
 * stop-test measure for the FP64 iterate, is below the domain's pjuge; the
 * statistics then report the norms up to that one.  No reference
 * counterpart: a precision variant. */
int les_domain_set_press_omega(les_domain* d, float omega);
/* 1 when every call issued on the domain has completed, 0 while work is
 * pending, -1 on error (does not wait).  No reference counterpart: used by
 * hosts that serve several domains at once. */
int les_domain_query(les_domain* d);
/* Self-test of the device reciprocal the kernels use in place of the
 * correctly rounded division 1.0f/x: counts the x with bit patterns in
 * [lo_bits, hi_bits) where the two differ, and the first such pattern
 * (0xffffffff when none).  Synchronises.  No reference counterpart. */
int les_selftest_reciprocal(unsigned lo_bits, unsigned hi_bits, unsigned long long* misses,
                            unsigned* first_bits);
/* Bytes of device memory a domain of these extents would hold. */
long long les_domain_bytes_needed(int im, int jm, int km);
int les_domain_press_stats(les_domain* d, int* iters, double* sor_hist, int hist_len, double* rhsav,
                           double* pav, int* nonfinite);
/* Bytes of device memory the domain holds. */
long long les_domain_bytes(const les_domain* d);

/* ------------------------------------------------------------ batches
 * Several domains advanced together, one step of each per les_batch_step.
 * Domains with a batchable whole step (fused reference SOR, separable cn1,
 * even jm, velfg+feedbf and les+adam fused) run every stage as one launch
 * for all of them: each block runs its own domain's stage on that domain's
 * grid, and the SOR iterations 2..nmaxp run as one launch per iteration
 * with one k chunk for the batch (LES_BATCH_SOR_KCHUNK and
 * LES_BATCH_SOR_TILE_ROWS override it).  A step in which such a domain has
 * n == 1 runs its stages on its own stream and only the SOR iterations
 * batched.  Other domains step whole on their own streams, joined by
 * events.  With the graph on (default), steps where every n > 1 replay one
 * graph (recaptured when a domain's state changes).  Fields and SOR
 * iteration counts are those of stepping each domain alone, bit for bit;
 * the FP64 residual history agrees to rounding (the batch's k chunk groups
 * the sum differently).  While in a batch a domain must not be stepped on
 * its own; the domain API may read it (les_domain_download waits for its
 * stream).  No reference counterpart: the batched step of the serving
 * path. */
typedef struct les_batch les_batch;
int les_batch_create(les_domain* const* domains, int n, les_batch** out);
/* Waits for the batch's work and frees the batch (not the domains). */
int les_batch_destroy(les_batch* b);
int les_batch_set_graph(les_batch* b, int on);
/* One step of every domain: n[q], nmax[q] for domain q (as les_domain_step). */
int les_batch_step(les_batch* b, const int* n, const int* nmax);
/* 1 when every step issued has completed, 0 while work is pending, -1 on error. */
int les_batch_query(les_batch* b);
int les_batch_synchronize(les_batch* b);

/* Per-kernel timing: while on (a nonzero bit mask of kernel classes, -1 for
 * all), every launch of a selected class is bracketed by a pair of HIP
 * events on the domain's stream.  les_domain_timing fills
 * ms[c] and count[c] for each kernel class c < les_kernel_classes() with
 * the launches recorded since the previous read, then clears them.
 * Synchronises.  No reference counterpart beyond the -DTIMINGS per-stage
 * clock of main.f95:243-303. */
int les_domain_set_timing(les_domain* d, int on);
int les_domain_timing(les_domain* d, double* ms, long long* count);
/* With on != 0, back-to-back launches of one selected class (no other
 * launch between them) share one event pair, so the time of a run of n
 * launches is n kernel durations plus the n-1 gaps between them, without
 * an event per launch.  Needs timing on.  No reference counterpart. */
int les_domain_set_timing_runs(les_domain* d, int on);
int les_kernel_classes(void);
const char* les_kernel_class_name(int c);

#ifdef __cplusplus
}
#endif
#endif
