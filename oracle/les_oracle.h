/*
 * CPU restatement of the MPI-LES per-time-step hot path -- TEST
 * INFRASTRUCTURE ONLY.  It is the checker for the HIP library: only tests/,
 * the smoke check and the CPU-baseline leg of the benchmark may load it.
 * The product never links or calls it.
 *
 * Every function takes exactly the argument list of the reference
 * subroutine it restates, with the reference's array layouts (Fortran
 * order, i fastest, explicit lower bounds, compile-time extents equal to
 * the run-time im,jm,km as the reference enforces at ifdata.f95:238-241).
 * Arrays are the caller's; nothing is zeroed here, exactly as the
 * reference relies on caller-owned static storage.
 *
 * Two instantiations from one source: orc32_* (real = float, the shipped
 * build) and orc64_* (real = double, the FP64 accuracy oracle).
 *
 * Reductions in press (rhsav/area, sor, pav/pco) follow the mode set by
 * orcNN_press_config: 0 = sequential in the working precision (the shipped
 * FP32 build), 1 = sequential FP64 accumulation of FP32 terms with the
 * mean subtraction done in FP64 (the GPU path's arithmetic).
 */
#ifndef LES_ORACLE_H
#define LES_ORACLE_H

#define ORC_DECLARE(P, real)                                                                      \
  void P##velnw(int km, int jm, int im, const real* p, real ro, const real* dxs, real* u, real dt,  \
                const real* f, const real* dys, real* v, const real* g, const real* dzs, real* w,    \
                const real* h);                                                                     \
  void P##bondv1(int jm, real* u, const real* z2, const real* dzn, real* v, real* w, int km, int n,  \
                 int im, real dt, const real* dxs);                                                 \
  void P##bondv1_set_inflow(const real* prof, int jm, int km);                                      \
  void P##bondv1_data24(int jm, real* u, const real* z2, const real* dzn, real* v, real* w, int km,  \
                        int n, int im, real dt, const real* dxs, const real* ua, const real* va,     \
                        const real* wa);                                                            \
  void P##aveflow(int n, int n1, int km, int jm, int im, real* aveu, real* avev, real* avew,        \
                  real* avep, real* avel, real* aveuu, real* avevv, real* aveww, real* avesm,       \
                  real* avesmsm, real* uwfx, const real* u, const real* v, const real* w,           \
                  const real* p, const real* sm, int nmax);                                         \
  void P##velfg(int km, int jm, int im, const real* dx1, real* cov1, real* cov2, real* cov3,        \
                real* dfu1, real* diu1, real* diu2, const real* dy1, real* diu3, const real* dzn,    \
                real vn, real* f, real* cov4, real* cov5, real* cov6, real* dfv1, real* diu4,        \
                real* diu5, real* diu6, real* g, real* cov7, real* cov8, real* cov9, real* dfw1,     \
                real* diu7, real* diu8, real* diu9, const real* dzs, real* h, real* nou1,            \
                const real* u, real* nou5, const real* v, real* nou9, const real* w, real* nou2,     \
                real* nou3, real* nou4, real* nou6, real* nou7, real* nou8, real* uspd, real* vspd); \
  void P##feedbf(int km, int jm, int im, real* usum, const real* u, const real* bmask1, real* vsum,  \
                 const real* v, const real* cmask1, real* wsum, const real* w, const real* dmask1,   \
                 real alpha, real dt, real beta, real* fx, real* fy, real* fz, real* f, real* g,     \
                 real* h);                                                                          \
  void P##les(int km, real* delx1, const real* dx1, const real* dy1, const real* dzn, int jm, int im, \
              const real* diu1, const real* diu2, const real* diu3, const real* diu4,                \
              const real* diu5, const real* diu6, const real* diu7, const real* diu8,                \
              const real* diu9, real* sm, real* f, real* g, real* h, const real* uspd,               \
              const real* vspd, const real* dxs, const real* dys);                                   \
  void P##boundsm(int km, int jm, real* sm, int im);                                                \
  void P##adam(int n, int nmax, real* fold, int im, int jm, int km, real* gold, real* hold,         \
               const real* fghold, real* f, real* g, real* h);                                      \
  void P##press(int km, int jm, int im, real* rhs, const real* u, const real* dx1, const real* v,   \
                const real* dy1, const real* w, const real* dzn, real* f, real* g, real* h, real dt, \
                const real* cn1, const real* cn2l, real* p, const real* cn2s, const real* cn3l,      \
                const real* cn3s, const real* cn4l, const real* cn4s, int n, int nmax,               \
                const real* usum, const real* vsum, const real* wsum);                              \
  void P##bondfg(int km, int jm, real* f, int im, real* g, real* h);                                \
  void P##boundp1(int km, int jm, real* p, int im);                                                 \
  void P##boundp2(int jm, int im, real* p, int km);                                                 \
  void P##press_config(int nmaxp, double pjuge, double omega, int redmode);                         \
  void P##press_stats(int* iters, double* sor_hist, int hist_len, double* rhsav, double* area,      \
                      double* pav, double* pco);

ORC_DECLARE(orc32_, float)
ORC_DECLARE(orc64_, double)

#endif
