/*
 * CPU restatement of the MPI-LES per-time-step hot path -- TEST
 * INFRASTRUCTURE ONLY (the checker, never the product).  See les_oracle.h.
 *
 * Compiled twice: without ORC_DOUBLE -> orc32_* (float), with ORC_DOUBLE ->
 * orc64_* (double).  Build with -ffp-contract=off and no fast-math so every
 * expression is evaluated exactly in the order written, which is the order
 * of the Fortran source line cited next to it (Fortran evaluates a+b+c as
 * (a+b)+c and a*b/c as (a*b)/c; x**2 is x*x and x**0.5 is sqrt(x) in the
 * flang build of the reference).
 *
 * Parity status: every hot-path function here is checked against the
 * reference's own subroutines built with flang (oracle/_ref) on seeded
 * inputs, stage by stage and over whole steps, by tests/test_oracle_vs_ref.py
 * and the committed fixtures under tests/golden/.  aveflow and
 * bondv1_data24 cannot be built from the reference without MPI (they call
 * distributeaveu/distributebondu of communication_helper_real.f95); they
 * are pinned instead by closed forms (means, population standard
 * deviations and the u-w covariance of fields with known time series, and
 * the inflow planes reducing to bondv1 with the same profile) in
 * tests/test_oracle_outputs.py.
 */
#include <math.h>
#include <stddef.h>
#include "les_oracle.h"

#ifdef ORC_DOUBLE
typedef double real;
#define ORC(n) orc64_##n
#define SQRT sqrt
#define LOG log
#define POW pow
#else
typedef float real;
#define ORC(n) orc32_##n
#define SQRT sqrtf
#define LOG logf
#define POW powf
#endif
#define C(x) ((real)(x))

/* The timing build of the restatement (liblesoracle_omp.so, -fopenmp)
 * splits the k loops of the stencil stages, the rhs and mean loops of press
 * and the SOR colour sweeps over threads; the serial build (liblesoracle.so,
 * the checker) compiles without OpenMP and ignores them.  The cells of one
 * such loop are independent -- each is written from arrays the loop does not
 * write, and a colour sweep reads only the other colour -- so the fields are
 * those of the serial loops; the FP64 sums of press (redmode 1) are reduced
 * per thread, in another order. */
#define PRAGMA_(x) _Pragma(#x)
#ifdef _OPENMP
#define PAR_K PRAGMA_(omp parallel for schedule(static))
#define PAR_K_SUM(...) PRAGMA_(omp parallel for schedule(static) reduction(+ : __VA_ARGS__))
#else
#define PAR_K
#define PAR_K_SUM(...)
#endif

/* ---- reference array layouts (main.f95:63-189 and the callee dummies) ---- */
typedef struct {
  long li, lj, lk, ni, nj;
} lay;
#define AT(a, L, i, j, k) \
  (a)[((long)(i) - (L).li) + ((long)(j) - (L).lj) * (L).ni + ((long)(k) - (L).lk) * (L).ni * (L).nj]

static lay layU(int im, int jm) { lay L = {0, -1, 0, im + 2, jm + 3}; return L; }  /* u,v (0:ip+1,-1:jp+1,0:kp+1) */
static lay layW(int im, int jm) { lay L = {0, -1, -1, im + 2, jm + 3}; return L; } /* w (0:ip+1,-1:jp+1,-1:kp+1) */
static lay layP(int im, int jm) { lay L = {0, 0, 0, im + 3, jm + 3}; return L; }   /* p (0:ip+2,0:jp+2,0:kp+1) */
static lay layF(int im, int jm) { lay L = {0, 0, 0, im + 1, jm + 1}; return L; }   /* f,g,h,usum.. (0:ip,0:jp,0:kp) */
static lay layI(int im, int jm) { lay L = {1, 1, 1, im, jm}; return L; }           /* fold..,cn1 (ip,jp,kp) */
static lay layR(int im, int jm) { lay L = {0, 0, 0, im + 2, jm + 2}; return L; }   /* rhs,amask1,dmask1 (0:ip+1,0:jp+1,0:kp+1) */
static lay layS(int im, int jm) { lay L = {-1, -1, 0, im + 3, jm + 3}; return L; } /* sm (-1:ip+1,-1:jp+1,0:kp+1) */
static lay layD1(int im, int jm) { lay L = {-1, 0, 0, im + 4, jm + 3}; return L; } /* diu1,5.. (-1:ip+2,0:jp+2,0:kp+2) */
static lay layD(int im, int jm) { lay L = {0, 0, 0, im + 3, jm + 3}; return L; }   /* other diu.. (0:ip+2,0:jp+2,0:kp+2) */
static lay layB(int im, int jm) { lay L = {-1, 0, 0, im + 3, jm + 2}; return L; }  /* bmask1 (-1:ip+1,0:jp+1,0:kp+1) */
static lay lay2(int im, int jm) { lay L = {0, 0, 0, im + 2, jm + 2}; return L; }   /* uspd,vspd (0:ip+1,0:jp+1) */
static lay layDU(int im, int jm) { lay L = {0, 1, 1, im + 1, jm}; return L; }      /* dfu1 (0:ip,jp,kp) */
static lay layDV(int im, int jm) { lay L = {1, 0, 1, im, jm + 1}; return L; }      /* dfv1 (ip,0:jp,kp) */

/* ---------------------------------------------------------------- velnw */
/* velnw.f95:5-71 */
void ORC(velnw)(int km, int jm, int im, const real* p, real ro, const real* dxs, real* u, real dt,
                const real* f, const real* dys, real* v, const real* g, const real* dzs, real* w,
                const real* h) {
  lay LP = layP(im, jm), LU = layU(im, jm), LW = layW(im, jm), LF = layF(im, jm);
  const real* DZS = dzs + 1; /* dzs(-1:kp+2) */
  PAR_K
  for (int k = 1; k <= km; k++) /* velnw.f95:27-35 */
    for (int j = 1; j <= jm; j++)
      for (int i = 1; i <= im; i++) {
        real pz = (-AT(p, LP, i, j, k) + AT(p, LP, i + 1, j, k)) / ro / dxs[i];
        AT(u, LU, i, j, k) = AT(u, LU, i, j, k) + dt * (AT(f, LF, i, j, k) - pz);
      }
  PAR_K
  for (int k = 1; k <= km; k++) /* velnw.f95:38-48 */
    for (int j = 1; j <= jm; j++)
      for (int i = 1; i <= im; i++) {
        real pz = (-AT(p, LP, i, j, k) + AT(p, LP, i, j + 1, k)) / ro / dys[j];
        AT(v, LU, i, j, k) = AT(v, LU, i, j, k) + dt * (AT(g, LF, i, j, k) - pz);
      }
  PAR_K
  for (int k = 1; k <= km - 1; k++) /* velnw.f95:51-59 */
    for (int j = 1; j <= jm; j++)
      for (int i = 1; i <= im; i++) {
        real pz = (-AT(p, LP, i, j, k) + AT(p, LP, i, j, k + 1)) / ro / DZS[k];
        AT(w, LW, i, j, k) = AT(w, LW, i, j, k) + dt * (AT(h, LF, i, j, k) - pz);
      }
}

/* --------------------------------------------------------------- bondv1 */
static const real* g_inflow = NULL; /* u_in(j,k), j fastest; NULL = reference 5 m/s */
static int g_inflow_jm = 0, g_inflow_km = 0;

void ORC(bondv1_set_inflow)(const real* prof, int jm, int km) {
  g_inflow = prof;
  g_inflow_jm = jm;
  g_inflow_km = km;
}

/* bondv1.f95:5-200 (serial branch).  The shipped inflow (bondv1.f95:33-54)
 * is 5.0 for k=1..78 and u(i,j,77)=5.0 above, i.e. 5.0 at every level; a
 * profile set with bondv1_set_inflow replaces those values (the configured
 * log-law inflow), nothing else changes. */
static void bondv1_after_inflow(int jm, real* u, real* v, real* w, int km, int n, int im, real dt,
                                const real* dxs);

void ORC(bondv1)(int jm, real* u, const real* z2, const real* dzn, real* v, real* w, int km, int n,
                 int im, real dt, const real* dxs) {
  lay LU = layU(im, jm), LW = layW(im, jm);
  (void)z2;
  (void)dzn;
  for (int i = 0; i <= 1; i++) /* bondv1.f95:33-54 */
    for (int k = 1; k <= km; k++)
      for (int j = 1; j <= jm; j++) {
        real uin = C(5.0);
        if (g_inflow && j <= g_inflow_jm && k <= g_inflow_km)
          uin = g_inflow[(j - 1) + (long)(k - 1) * g_inflow_jm];
        AT(u, LU, i, j, k) = uin;
        AT(v, LU, i, j, k) = C(0.0);
        AT(w, LW, i, j, k) = C(0.0);
      }
  bondv1_after_inflow(jm, u, v, w, km, n, im, dt, dxs);
}

/* bondv1_data24.f95:6-260 (one process: isMaster and isTopRow hold,
 * jpmax = jp).  The file records (bondv1_data24.f95:43-59) are read by the
 * caller; ua/va/wa(j,k), j fastest, hold ub/vb/wb(400+j,k) for j=1..jm
 * (bondv1_data24.f95:62-68).  Inflow planes i=0,1 for k=1..km, j=1..jm
 * (:79-87), then the rest of bondv1 unchanged (:124-258 repeat
 * bondv1.f95:61-195). */
void ORC(bondv1_data24)(int jm, real* u, const real* z2, const real* dzn, real* v, real* w, int km,
                        int n, int im, real dt, const real* dxs, const real* ua, const real* va,
                        const real* wa) {
  lay LU = layU(im, jm), LW = layW(im, jm);
  (void)z2;
  (void)dzn;
  for (int k = 1; k <= km; k++)
    for (int j = 1; j <= jm; j++)
      for (int i = 0; i <= 1; i++) {
        const long q = (j - 1) + (long)(k - 1) * jm;
        AT(u, LU, i, j, k) = ua[q];
        AT(v, LU, i, j, k) = va[q];
        AT(w, LW, i, j, k) = wa[q];
      }
  bondv1_after_inflow(jm, u, v, w, km, n, im, dt, dxs);
}

static void bondv1_after_inflow(int jm, real* u, real* v, real* w, int km, int n, int im, real dt,
                                const real* dxs) {
  lay LU = layU(im, jm), LW = layW(im, jm);
  if (n == 1) { /* bondv1.f95:61-88, ICAL == 0 */
    for (int k = 1; k <= km; k++)
      for (int j = 1; j <= jm; j++)
        for (int i = 2; i <= im; i++) {
          AT(u, LU, i, j, k) = AT(u, LU, 1, j, k);
          AT(v, LU, i, j, k) = AT(v, LU, 1, j, k);
          AT(w, LW, i, j, k) = AT(w, LW, 1, j, k);
        }
  }
  real aaa = C(0.0), bbb = C(0.0); /* bondv1.f95:98-114 */
  for (int k = 1; k <= km; k++)
    for (int j = 1; j <= jm; j++) {
      real x = AT(u, LU, im, j, k);
      aaa = (x > aaa) ? x : aaa;
      bbb = (x < bbb) ? x : bbb;
    }
  real uout = (aaa + bbb) / C(2.);
  for (int k = 1; k <= km; k++) /* bondv1.f95:118-122 */
    for (int j = 1; j <= jm; j++)
      AT(u, LU, im, j, k) = AT(u, LU, im, j, k) -
                            dt * uout * (AT(u, LU, im, j, k) - AT(u, LU, im - 1, j, k)) / dxs[im];
  for (int k = 1; k <= km; k++) /* bondv1.f95:124-128 */
    for (int j = 1; j <= jm; j++)
      AT(v, LU, im + 1, j, k) = AT(v, LU, im + 1, j, k) -
                                dt * uout * (AT(v, LU, im + 1, j, k) - AT(v, LU, im, j, k)) / dxs[im];
  for (int k = 1; k <= km; k++) /* bondv1.f95:130-134 */
    for (int j = 1; j <= jm; j++)
      AT(w, LW, im + 1, j, k) = AT(w, LW, im + 1, j, k) -
                                dt * uout * (AT(w, LW, im + 1, j, k) - AT(w, LW, im, j, k)) / dxs[im];
  for (int k = 0; k <= km + 1; k++) /* bondv1.f95:137-142 */
    for (int i = 0; i <= im + 1; i++) {
      AT(u, LU, i, 0, k) = AT(u, LU, i, jm, k);
      AT(u, LU, i, jm + 1, k) = AT(u, LU, i, 1, k);
    }
  for (int k = 0; k <= km + 1; k++) /* bondv1.f95:144-149 */
    for (int i = 0; i <= im + 1; i++) {
      AT(v, LU, i, 0, k) = AT(v, LU, i, jm, k);
      AT(v, LU, i, jm + 1, k) = AT(v, LU, i, 1, k);
    }
  for (int k = 0; k <= km; k++) /* bondv1.f95:151-156 */
    for (int i = 0; i <= im + 1; i++) {
      AT(w, LW, i, 0, k) = AT(w, LW, i, jm, k);
      AT(w, LW, i, jm + 1, k) = AT(w, LW, i, 1, k);
    }
  for (int j = 0; j <= jm + 1; j++) /* bondv1.f95:176-181 */
    for (int i = 0; i <= im + 1; i++) {
      AT(u, LU, i, j, 0) = -AT(u, LU, i, j, 1);
      AT(u, LU, i, j, km + 1) = AT(u, LU, i, j, km);
    }
  for (int j = 0; j <= jm + 1; j++) /* bondv1.f95:183-188 */
    for (int i = 0; i <= im + 1; i++) {
      AT(v, LU, i, j, 0) = -AT(v, LU, i, j, 1);
      AT(v, LU, i, j, km + 1) = AT(v, LU, i, j, km);
    }
  for (int j = -1; j <= jm + 1; j++) /* bondv1.f95:190-195 */
    for (int i = 0; i <= im + 1; i++) {
      AT(w, LW, i, j, 0) = C(0.0);
      AT(w, LW, i, j, km) = C(0.0);
    }
}

/* ------------------------------------------------------------ vel2/velfg */
/* vel2.f95:4-364 (serial branch) */
static void vel2(int km, int jm, int im, real* nou1, const real* u, real* diu1, const real* dx1,
                 real* nou5, const real* v, real* diu5, const real* dy1, real* nou9, const real* w,
                 real* diu9, const real* dzn, real* cov1, real* cov5, real* cov9, real* nou2,
                 real* diu2, real* cov2, real* nou3, real* diu3, const real* dzs, real* cov3,
                 real* nou4, real* diu4, real* cov4, real* nou6, real* diu6, real* cov6, real* nou7,
                 real* diu7, real* cov7, real* nou8, real* diu8, real* cov8, real* uspd, real* vspd) {
  lay LU = layU(im, jm), LW = layW(im, jm), LD1 = layD1(im, jm), LD = layD(im, jm), L2 = lay2(im, jm);
  const real* DX1 = dx1 + 1; /* dx1(-1:ip+1) */
  const real* DZN = dzn + 1; /* dzn(-1:kp+2) */
  const real* DZS = dzs + 1;
#define U_(i, j, k) AT(u, LU, i, j, k)
#define V_(i, j, k) AT(v, LU, i, j, k)
#define W_(i, j, k) AT(w, LW, i, j, k)
#define S2(a, i, j) AT(a, L2, i, j, 0)
  for (int j = 1; j <= jm; j++) /* vel2.f95:54-59 */
    for (int i = 1; i <= im; i++) {
      real t = (C(0.5) * (V_(i, j - 1, 1) + V_(i, j, 1)) * DX1[i + 1] +
                C(0.5) * (V_(i + 1, j - 1, 1) + V_(i + 1, j, 1)) * DX1[i]) /
               (DX1[i] + DX1[i + 1]);
      S2(uspd, i, j) = SQRT(U_(i, j, 1) * U_(i, j, 1) + t * t);
    }
  for (int j = 1; j <= jm; j++) /* vel2.f95:60-65 */
    for (int i = 1; i <= im; i++) {
      real t = (C(0.5) * (U_(i - 1, j, 1) + U_(i, j, 1)) * dy1[j + 1] +
                C(0.5) * (U_(i - 1, j + 1, 1) + U_(i, j + 1, 1)) * dy1[j]) /
               (dy1[j] + dy1[j + 1]);
      S2(vspd, i, j) = SQRT(V_(i, j, 1) * V_(i, j, 1) + t * t);
    }
  PAR_K
  for (int k = 1; k <= km; k++) /* vel2.f95:74-89 */
    for (int j = 1; j <= jm; j++)
      for (int i = 1; i <= im; i++) {
        AT(nou1, LD1, i, j, k) = (U_(i - 1, j, k) + U_(i, j, k)) / C(2.);
        AT(diu1, LD1, i, j, k) = (-U_(i - 1, j, k) + U_(i, j, k)) / DX1[i];
        AT(nou5, LD1, i, j, k) = (V_(i, j - 1, k) + V_(i, j, k)) / C(2.);
        AT(diu5, LD1, i, j, k) = (-V_(i, j - 1, k) + V_(i, j, k)) / dy1[j];
        AT(nou9, LD, i, j, k) = (W_(i, j, k - 1) + W_(i, j, k)) / C(2.);
        AT(diu9, LD, i, j, k) = (-W_(i, j, k - 1) + W_(i, j, k)) / DZN[k];
        AT(cov1, LD1, i, j, k) = AT(nou1, LD1, i, j, k) * AT(diu1, LD1, i, j, k);
        AT(cov5, LD1, i, j, k) = AT(nou5, LD1, i, j, k) * AT(diu5, LD1, i, j, k);
        AT(cov9, LD, i, j, k) = AT(nou9, LD, i, j, k) * AT(diu9, LD, i, j, k);
      }
  PAR_K
  for (int k = 1; k <= km; k++) /* vel2.f95:103-111 */
    for (int j = 1; j <= jm; j++)
      for (int i = 1; i <= im; i++) {
        AT(nou2, LD, i, j, k) =
            (DX1[i + 1] * V_(i, j - 1, k) + DX1[i] * V_(i + 1, j - 1, k)) / (DX1[i] + DX1[i + 1]);
        AT(diu2, LD, i, j, k) = C(2.) * (-U_(i, j - 1, k) + U_(i, j, k)) / (dy1[j - 1] + dy1[j]);
        AT(cov2, LD, i, j, k) = AT(nou2, LD, i, j, k) * AT(diu2, LD, i, j, k);
      }
  PAR_K
  for (int k = 2; k <= km + 1; k++) /* vel2.f95:119-127 */
    for (int j = 1; j <= jm; j++)
      for (int i = 1; i <= im; i++) {
        AT(nou3, LD, i, j, k) =
            (DX1[i + 1] * W_(i, j, k - 1) + DX1[i] * W_(i + 1, j, k - 1)) / (DX1[i] + DX1[i + 1]);
        AT(diu3, LD, i, j, k) = (-U_(i, j, k - 1) + U_(i, j, k)) / DZS[k - 1];
        AT(cov3, LD, i, j, k) = AT(nou3, LD, i, j, k) * AT(diu3, LD, i, j, k);
      }
  for (int j = 1; j <= jm; j++) /* vel2.f95:135-141, wall layer */
    for (int i = 1; i <= im; i++) {
      AT(nou3, LD, i, j, 1) =
          C(0.5) * (DX1[i + 1] * W_(i, j, 1) + DX1[i] * W_(i + 1, j, 1)) / (DX1[i] + DX1[i + 1]);
      AT(diu3, LD, i, j, 1) = S2(uspd, i, j) * C(0.4) / LOG(C(0.5) * DZN[1] / C(0.1)) /
                              (C(0.5) * DZN[1]) / C(0.4) * U_(i, j, 1) / S2(uspd, i, j);
      AT(cov3, LD, i, j, 1) = AT(nou3, LD, i, j, 1) * AT(diu3, LD, i, j, 1);
    }
  PAR_K
  for (int k = 1; k <= km; k++) /* vel2.f95:144-152 (u0 = 0) */
    for (int j = 1; j <= jm; j++)
      for (int i = 1; i <= im; i++) {
        AT(nou4, LD, i, j, k) =
            (dy1[j + 1] * U_(i - 1, j, k) + dy1[j] * U_(i - 1, j + 1, k)) / (dy1[j] + dy1[j + 1]);
        AT(diu4, LD, i, j, k) = C(2.) * (-V_(i - 1, j, k) + V_(i, j, k)) / (DX1[i - 1] + DX1[i]);
        AT(cov4, LD, i, j, k) = (AT(nou4, LD, i, j, k) - C(0)) * AT(diu4, LD, i, j, k);
      }
  PAR_K
  for (int k = 2; k <= km + 1; k++) /* vel2.f95:154-162 */
    for (int j = 1; j <= jm; j++)
      for (int i = 1; i <= im; i++) {
        AT(nou6, LD, i, j, k) =
            (dy1[j + 1] * W_(i, j, k - 1) + dy1[j] * W_(i, j + 1, k - 1)) / (dy1[j] + dy1[j + 1]);
        AT(diu6, LD, i, j, k) = (-V_(i, j, k - 1) + V_(i, j, k)) / DZS[k - 1];
        AT(cov6, LD, i, j, k) = AT(nou6, LD, i, j, k) * AT(diu6, LD, i, j, k);
      }
  for (int j = 1; j <= jm; j++) /* vel2.f95:170-176, wall layer */
    for (int i = 1; i <= im; i++) {
      AT(nou6, LD, i, j, 1) =
          C(0.5) * (dy1[j + 1] * W_(i, j, 1) + dy1[j] * W_(i, j + 1, 1)) / (dy1[j] + dy1[j + 1]);
      AT(diu6, LD, i, j, 1) = S2(vspd, i, j) * C(0.4) / LOG(C(0.5) * DZN[1] / C(0.1)) /
                              (C(0.5) * DZN[1]) / C(0.4) * V_(i, j, 1) / S2(vspd, i, j);
      AT(cov6, LD, i, j, 1) = AT(nou6, LD, i, j, 1) * AT(diu6, LD, i, j, 1);
    }
  PAR_K
  for (int k = 1; k <= km - 1; k++) /* vel2.f95:179-187 */
    for (int j = 1; j <= jm; j++)
      for (int i = 1; i <= im; i++) {
        AT(nou7, LD, i, j, k) =
            (DZN[k + 1] * U_(i - 1, j, k) + DZN[k] * U_(i - 1, j, k + 1)) / (DZN[k] + DZN[k + 1]);
        AT(diu7, LD, i, j, k) = C(2.) * (-W_(i - 1, j, k) + W_(i, j, k)) / (DX1[i - 1] + DX1[i]);
        AT(cov7, LD, i, j, k) = (AT(nou7, LD, i, j, k) - C(0)) * AT(diu7, LD, i, j, k);
      }
  PAR_K
  for (int k = 1; k <= km - 1; k++) /* vel2.f95:189-197 */
    for (int j = 1; j <= jm; j++)
      for (int i = 1; i <= im; i++) {
        AT(nou8, LD, i, j, k) =
            (DZN[k + 1] * V_(i, j - 1, k) + DZN[k] * V_(i, j - 1, k + 1)) / (DZN[k] + DZN[k + 1]);
        AT(diu8, LD, i, j, k) = C(2.) * (-W_(i, j - 1, k) + W_(i, j, k)) / (dy1[j - 1] + dy1[j]);
        AT(cov8, LD, i, j, k) = AT(nou8, LD, i, j, k) * AT(diu8, LD, i, j, k);
      }
  PAR_K
  for (int k = 1; k <= km; k++) /* vel2.f95:202-208 */
    for (int j = 1; j <= jm; j++) {
      AT(nou1, LD1, im + 1, j, k) = AT(nou1, LD1, im, j, k);
      AT(diu1, LD1, im + 1, j, k) = AT(diu1, LD1, im, j, k);
      AT(cov1, LD1, im + 1, j, k) = AT(cov1, LD1, im, j, k);
    }
  PAR_K
  for (int k = 1; k <= km; k++) /* vel2.f95:213-222 */
    for (int i = 1; i <= im; i++) {
      AT(nou2, LD, i, 0, k) = AT(nou2, LD, i, jm, k);
      AT(diu2, LD, i, 0, k) = AT(diu2, LD, i, jm, k);
      AT(cov2, LD, i, 0, k) = AT(cov2, LD, i, jm, k);
      AT(nou2, LD, i, jm + 1, k) = AT(nou2, LD, i, 1, k);
      AT(diu2, LD, i, jm + 1, k) = AT(diu2, LD, i, 1, k);
      AT(cov2, LD, i, jm + 1, k) = AT(cov2, LD, i, 1, k);
    }
  PAR_K
  for (int k = 1; k <= km; k++) /* vel2.f95:234-240 */
    for (int j = 1; j <= jm; j++) {
      AT(nou4, LD, im + 1, j, k) = AT(nou4, LD, im, j, k);
      AT(diu4, LD, im + 1, j, k) = AT(diu4, LD, im, j, k);
      AT(cov4, LD, im + 1, j, k) = AT(cov4, LD, im, j, k);
    }
  PAR_K
  for (int k = 1; k <= km; k++) /* vel2.f95:245-254 */
    for (int i = 1; i <= im; i++) {
      AT(nou5, LD1, i, 0, k) = AT(nou5, LD1, i, jm, k);
      AT(diu5, LD1, i, 0, k) = AT(diu5, LD1, i, jm, k);
      AT(cov5, LD1, i, 0, k) = AT(cov5, LD1, i, jm, k);
      AT(nou5, LD1, i, jm + 1, k) = AT(nou5, LD1, i, 1, k);
      AT(diu5, LD1, i, jm + 1, k) = AT(diu5, LD1, i, 1, k);
      AT(cov5, LD1, i, jm + 1, k) = AT(cov5, LD1, i, 1, k);
    }
  PAR_K
  for (int k = 1; k <= km - 1; k++) /* vel2.f95:266-272 */
    for (int j = 1; j <= jm; j++) {
      AT(nou7, LD, im + 1, j, k) = AT(nou7, LD, im, j, k);
      AT(diu7, LD, im + 1, j, k) = AT(diu7, LD, im, j, k);
      AT(cov7, LD, im + 1, j, k) = AT(cov7, LD, im, j, k);
    }
  PAR_K
  for (int k = 1; k <= km - 1; k++) /* vel2.f95:277-286 */
    for (int i = 1; i <= im; i++) {
      AT(nou8, LD, i, 0, k) = AT(nou8, LD, i, jm, k);
      AT(diu8, LD, i, 0, k) = AT(diu8, LD, i, jm, k);
      AT(cov8, LD, i, 0, k) = AT(cov8, LD, i, jm, k);
      AT(nou8, LD, i, jm + 1, k) = AT(nou8, LD, i, 1, k);
      AT(diu8, LD, i, jm + 1, k) = AT(diu8, LD, i, 1, k);
      AT(cov8, LD, i, jm + 1, k) = AT(cov8, LD, i, 1, k);
    }
  PAR_K
  for (int k = 1; k <= km + 1; k++) /* vel2.f95:299-304 */
    for (int j = 1; j <= jm + 1; j++) {
      AT(diu2, LD, im + 1, j, k) = AT(diu2, LD, im, j, k);
      AT(diu3, LD, im + 1, j, k) = AT(diu3, LD, im, j, k);
    }
  PAR_K
  for (int k = 1; k <= km + 1; k++) /* vel2.f95:309-314 */
    for (int i = 1; i <= im + 1; i++) {
      AT(diu4, LD, i, 0, k) = AT(diu4, LD, i, jm, k);
      AT(diu6, LD, i, 0, k) = AT(diu6, LD, i, jm, k);
    }
  for (int j = 1; j <= jm; j++) /* vel2.f95:351-356 */
    for (int i = 1; i <= im; i++) {
      S2(uspd, i, j) = U_(i, j, 1) / S2(uspd, i, j);
      S2(vspd, i, j) = V_(i, j, 1) / S2(vspd, i, j);
    }
#undef U_
#undef V_
#undef W_
#undef S2
}

/* velFG.f95:6-124 */
void ORC(velfg)(int km, int jm, int im, const real* dx1, real* cov1, real* cov2, real* cov3,
                real* dfu1, real* diu1, real* diu2, const real* dy1, real* diu3, const real* dzn,
                real vn, real* f, real* cov4, real* cov5, real* cov6, real* dfv1, real* diu4,
                real* diu5, real* diu6, real* g, real* cov7, real* cov8, real* cov9, real* dfw1,
                real* diu7, real* diu8, real* diu9, const real* dzs, real* h, real* nou1,
                const real* u, real* nou5, const real* v, real* nou9, const real* w, real* nou2,
                real* nou3, real* nou4, real* nou6, real* nou7, real* nou8, real* uspd, real* vspd) {
  lay LD1 = layD1(im, jm), LD = layD(im, jm), LF = layF(im, jm), LI = layI(im, jm);
  lay LDU = layDU(im, jm), LDV = layDV(im, jm);
  const real* DX1 = dx1 + 1;
  const real* DZN = dzn + 1;
  const real* DZS = dzs + 1;
  vel2(km, jm, im, nou1, u, diu1, dx1, nou5, v, diu5, dy1, nou9, w, diu9, dzn, cov1, cov5, cov9,
       nou2, diu2, cov2, nou3, diu3, dzs, cov3, nou4, diu4, cov4, nou6, diu6, cov6, nou7, diu7,
       cov7, nou8, diu8, cov8, uspd, vspd); /* velFG.f95:61-62 */
  PAR_K
  for (int k = 1; k <= km; k++) /* velFG.f95:64-77 */
    for (int j = 1; j <= jm; j++)
      for (int i = 1; i <= im; i++) {
        real covx1 = (DX1[i + 1] * AT(cov1, LD1, i, j, k) + DX1[i] * AT(cov1, LD1, i + 1, j, k)) /
                     (DX1[i] + DX1[i + 1]);
        real covy1 = (AT(cov2, LD, i, j, k) + AT(cov2, LD, i, j + 1, k)) / C(2.);
        real covz1 = (AT(cov3, LD, i, j, k) + AT(cov3, LD, i, j, k + 1)) / C(2.);
        real covc = covx1 + covy1 + covz1;
        real d = C(2.) * (-AT(diu1, LD1, i, j, k) + AT(diu1, LD1, i + 1, j, k)) / (DX1[i] + DX1[i + 1]) +
                 (-AT(diu2, LD, i, j, k) + AT(diu2, LD, i, j + 1, k)) / dy1[j] +
                 (-AT(diu3, LD, i, j, k) + AT(diu3, LD, i, j, k + 1)) / DZN[k];
        AT(dfu1, LDU, i, j, k) = d;
        real df = vn * d;
        AT(f, LF, i, j, k) = (-covc + df);
      }
  PAR_K
  for (int k = 1; k <= km; k++) /* velFG.f95:80-93 */
    for (int j = 1; j <= jm; j++)
      for (int i = 1; i <= im; i++) {
        real covx1 = (AT(cov4, LD, i, j, k) + AT(cov4, LD, i + 1, j, k)) / C(2.);
        real covy1 = (dy1[j + 1] * AT(cov5, LD1, i, j, k) + dy1[j] * AT(cov5, LD1, i, j + 1, k)) /
                     (dy1[j] + dy1[j + 1]);
        real covz1 = (AT(cov6, LD, i, j, k) + AT(cov6, LD, i, j, k + 1)) / C(2.);
        real covc = covx1 + covy1 + covz1;
        real d = (-AT(diu4, LD, i, j, k) + AT(diu4, LD, i + 1, j, k)) / DX1[i] +
                 C(2.) * (-AT(diu5, LD1, i, j, k) + AT(diu5, LD1, i, j + 1, k)) / (dy1[j] + dy1[j + 1]) +
                 (-AT(diu6, LD, i, j, k) + AT(diu6, LD, i, j, k + 1)) / DZN[k];
        AT(dfv1, LDV, i, j, k) = d;
        real df = vn * d;
        AT(g, LF, i, j, k) = (-covc + df);
      }
  PAR_K
  for (int k = 1; k <= km - 1; k++) /* velFG.f95:97-110 */
    for (int j = 1; j <= jm; j++)
      for (int i = 1; i <= im; i++) {
        real covx1 = (AT(cov7, LD, i, j, k) + AT(cov7, LD, i + 1, j, k)) / C(2.);
        real covy1 = (AT(cov8, LD, i, j, k) + AT(cov8, LD, i, j + 1, k)) / C(2.);
        real covz1 = (DZN[k + 1] * AT(cov9, LD, i, j, k) + DZN[k] * AT(cov9, LD, i, j, k + 1)) /
                     (DZN[k] + DZN[k + 1]);
        real covc = covx1 + covy1 + covz1;
        real d = (-AT(diu7, LD, i, j, k) + AT(diu7, LD, i + 1, j, k)) / DX1[i] +
                 (-AT(diu8, LD, i, j, k) + AT(diu8, LD, i, j + 1, k)) / dy1[j] +
                 (-AT(diu9, LD, i, j, k) + AT(diu9, LD, i, j, k + 1)) / DZS[k];
        AT(dfw1, LI, i, j, k) = d;
        real df = vn * d;
        AT(h, LF, i, j, k) = (-covc + df);
      }
}

/* --------------------------------------------------------------- feedbf */
/* feedbf.f95:5-76 */
void ORC(feedbf)(int km, int jm, int im, real* usum, const real* u, const real* bmask1, real* vsum,
                 const real* v, const real* cmask1, real* wsum, const real* w, const real* dmask1,
                 real alpha, real dt, real beta, real* fx, real* fy, real* fz, real* f, real* g,
                 real* h) {
  lay LF = layF(im, jm), LU = layU(im, jm), LW = layW(im, jm), LB = layB(im, jm), LR = layR(im, jm);
  PAR_K
  for (int k = 1; k <= km; k++) /* feedbf.f95:35-50 */
    for (int j = 1; j <= jm; j++)
      for (int i = 1; i <= im; i++) {
        AT(usum, LF, i, j, k) = (AT(usum, LF, i, j, k) + AT(u, LU, i, j, k)) * AT(bmask1, LB, i, j, k);
        AT(vsum, LF, i, j, k) = (AT(vsum, LF, i, j, k) + AT(v, LU, i, j, k)) * AT(cmask1, LU, i, j, k);
        AT(wsum, LF, i, j, k) = (AT(wsum, LF, i, j, k) + AT(w, LW, i, j, k)) * AT(dmask1, LR, i, j, k);
        real f1x = alpha * AT(usum, LF, i, j, k) * dt;
        real f1y = alpha * AT(vsum, LF, i, j, k) * dt;
        real f1z = alpha * AT(wsum, LF, i, j, k) * dt;
        real f2x = beta * AT(u, LU, i, j, k) * AT(bmask1, LB, i, j, k);
        real f2y = beta * AT(v, LU, i, j, k) * AT(cmask1, LU, i, j, k);
        real f2z = beta * AT(w, LW, i, j, k) * AT(dmask1, LR, i, j, k);
        AT(fx, LF, i, j, k) = f1x + f2x;
        AT(fy, LF, i, j, k) = f1y + f2y;
        AT(fz, LF, i, j, k) = f1z + f2z;
      }
  PAR_K
  for (int k = 1; k <= km; k++) /* feedbf.f95:53-61 */
    for (int j = 1; j <= jm; j++)
      for (int i = 1; i <= im; i++) {
        AT(f, LF, i, j, k) = AT(f, LF, i, j, k) + AT(fx, LF, i, j, k);
        AT(g, LF, i, j, k) = AT(g, LF, i, j, k) + AT(fy, LF, i, j, k);
        AT(h, LF, i, j, k) = AT(h, LF, i, j, k) + AT(fz, LF, i, j, k);
      }
}

/* -------------------------------------------------------------- boundsm */
/* boundsm.f95:5-69 (serial: both Neumann branches execute) */
void ORC(boundsm)(int km, int jm, real* sm, int im) {
  lay LS = layS(im, jm);
  for (int k = 0; k <= km + 1; k++) /* boundsm.f95:18-32 */
    for (int j = -1; j <= jm + 1; j++) {
      AT(sm, LS, 0, j, k) = AT(sm, LS, 1, j, k);
      AT(sm, LS, im + 1, j, k) = AT(sm, LS, im, j, k);
    }
  for (int k = 0; k <= km + 1; k++) /* boundsm.f95:40-54 */
    for (int i = 0; i <= im + 1; i++) {
      AT(sm, LS, i, jm + 1, k) = AT(sm, LS, i, jm, k);
      AT(sm, LS, i, 0, k) = AT(sm, LS, i, 1, k);
    }
  for (int j = -1; j <= jm + 1; j++) /* boundsm.f95:59-64 */
    for (int i = 0; i <= im + 1; i++) {
      AT(sm, LS, i, j, 0) = -AT(sm, LS, i, j, 1);
      AT(sm, LS, i, j, km + 1) = AT(sm, LS, i, j, km);
    }
}

/* ------------------------------------------------------------------ les */
/* les.f95:6-246 */
void ORC(les)(int km, real* delx1, const real* dx1, const real* dy1, const real* dzn, int jm, int im,
              const real* diu1, const real* diu2, const real* diu3, const real* diu4,
              const real* diu5, const real* diu6, const real* diu7, const real* diu8,
              const real* diu9, real* sm, real* f, real* g, real* h, const real* uspd,
              const real* vspd, const real* dxs, const real* dys) {
  lay LD1 = layD1(im, jm), LD = layD(im, jm), LS = layS(im, jm), LF = layF(im, jm), L2 = lay2(im, jm);
  const real* DX1 = dx1 + 1;
  const real* DZN = dzn + 1;
  real* DELX1 = delx1 - 1; /* delx1(kp) */
  const real cs0 = C(.1);  /* les.f95:36 */
#define SM(i, j, k) AT(sm, LS, i, j, k)
#define D1(a, i, j, k) AT(a, LD1, i, j, k)
#define DD(a, i, j, k) AT(a, LD, i, j, k)
  PAR_K
  for (int k = 1; k <= km; k++) /* les.f95:38-42 */
    DELX1[k] = POW(DX1[0] * dy1[0] * DZN[k], C(1.) / C(3.));
  PAR_K
  for (int k = 1; k <= km; k++) /* les.f95:45-65 */
    for (int j = 1; j <= jm; j++)
      for (int i = 1; i <= im; i++) {
        real dudxx1 = D1(diu1, i, j, k);
        real dudyx1 = (DD(diu2, i - 1, j, k) + DD(diu2, i - 1, j + 1, k) + DD(diu2, i, j, k) +
                       DD(diu2, i, j + 1, k)) * C(.25);
        real dudzx1 = (DD(diu3, i - 1, j, k) + DD(diu3, i - 1, j, k + 1) + DD(diu3, i, j, k) +
                       DD(diu3, i, j, k + 1)) * C(.25);
        real dvdxx1 = (DD(diu4, i, j, k) + DD(diu4, i, j - 1, k) + DD(diu4, i + 1, j, k) +
                       DD(diu4, i + 1, j - 1, k)) * C(.25);
        real dvdyx1 = D1(diu5, i, j, k);
        real dvdzx1 = (DD(diu6, i, j - 1, k) + DD(diu6, i, j - 1, k + 1) + DD(diu6, i, j, k) +
                       DD(diu6, i, j, k + 1)) * C(.25);
        real dwdxx1 = (DD(diu7, i, j, k) + DD(diu7, i, j, k - 1) + DD(diu7, i + 1, j, k) +
                       DD(diu7, i + 1, j, k - 1)) * C(.25);
        real dwdyx1 = (DD(diu8, i, j, k) + DD(diu8, i, j, k - 1) + DD(diu8, i, j + 1, k) +
                       DD(diu8, i, j + 1, k - 1)) * C(.25);
        real dwdzx1 = DD(diu9, i, j, k);
        real csx1 = cs0;
        real a = csx1 * DELX1[k];
        real s = C(2.) * (dudxx1 * dudxx1 + dvdyx1 * dvdyx1 + dwdzx1 * dwdzx1) +
                 (dudyx1 + dvdxx1) * (dudyx1 + dvdxx1) + (dwdyx1 + dvdzx1) * (dwdyx1 + dvdzx1) +
                 (dudzx1 + dwdxx1) * (dudzx1 + dwdxx1);
        SM(i, j, k) = a * a * SQRT(s);
      }
  ORC(boundsm)(km, jm, sm, im); /* les.f95:74 */
  const real wl = LOG(C(0.5) * DZN[1] / C(0.1));
  PAR_K
  for (int k = 2; k <= km; k++) /* les.f95:83-114 */
    for (int j = 1; j <= jm; j++)
      for (int i = 1; i <= im; i++) {
        real evsx2 = SM(i + 1, j, k);
        real evsx1 = SM(i, j, k);
        real evsy2 = (dy1[j + 1] * ((DX1[i + 1] * SM(i, j, k) + DX1[i] * SM(i + 1, j, k)) / (DX1[i] + DX1[i + 1])) +
                      dy1[j] * ((DX1[i + 1] * SM(i, j + 1, k) + DX1[i] * SM(i + 1, j + 1, k)) / (DX1[i] + DX1[i + 1]))) /
                     (dy1[j] + dy1[j + 1]);
        real evsy1 = (dy1[j + 1] * ((DX1[i + 1] * SM(i, j - 1, k) + DX1[i] * SM(i + 1, j - 1, k)) / (DX1[i] + DX1[i + 1])) +
                      dy1[j] * ((DX1[i + 1] * SM(i, j, k) + DX1[i] * SM(i + 1, j, k)) / (DX1[i] + DX1[i + 1]))) /
                     (dy1[j] + dy1[j + 1]);
        real evsz2 = (DZN[k + 1] * ((DX1[i + 1] * SM(i, j, k) + DX1[i] * SM(i + 1, j, k)) / (DX1[i] + DX1[i + 1])) +
                      DZN[k] * ((DX1[i + 1] * SM(i, j, k + 1) + DX1[i] * SM(i + 1, j, k + 1)) / (DX1[i] + DX1[i + 1]))) /
                     (DZN[k] + DZN[k + 1]);
        real evsz1 = (DZN[k] * ((DX1[i + 1] * SM(i, j, k - 1) + DX1[i] * SM(i + 1, j, k - 1)) / (DX1[i] + DX1[i + 1])) +
                      DZN[k - 1] * ((DX1[i + 1] * SM(i, j, k) + DX1[i] * SM(i + 1, j, k)) / (DX1[i] + DX1[i + 1]))) /
                     (DZN[k - 1] + DZN[k]);
        real visux2 = (evsx2) * C(2.) * D1(diu1, i + 1, j, k);
        real visux1 = (evsx1) * C(2.) * D1(diu1, i, j, k);
        real visuy2 = (evsy2) * (DD(diu2, i, j + 1, k) + DD(diu4, i + 1, j, k));
        real visuy1 = (evsy1) * (DD(diu2, i, j, k) + DD(diu4, i + 1, j - 1, k));
        real visuz2 = (evsz2) * (DD(diu3, i, j, k + 1) + DD(diu7, i + 1, j, k));
        real visuz1 = (evsz1) * (DD(diu3, i, j, k) + DD(diu7, i + 1, j, k - 1));
        real vfu = (visux2 - visux1) / dxs[i] + (visuy2 - visuy1) / dy1[j] + (visuz2 - visuz1) / DZN[k];
        AT(f, LF, i, j, k) = (AT(f, LF, i, j, k) + vfu);
      }
  for (int j = 1; j <= jm; j++) /* les.f95:118-139, wall function (Q1: double count) */
    for (int i = 1; i <= im; i++) {
      real evsx2 = SM(i + 1, j, 1);
      real evsx1 = SM(i, j, 1);
      real evsy2 = (dy1[j + 1] * ((DX1[i + 1] * SM(i, j, 1) + DX1[i] * SM(i + 1, j, 1)) / (DX1[i] + DX1[i + 1])) +
                    dy1[j] * ((DX1[i + 1] * SM(i, j + 1, 1) + DX1[i] * SM(i + 1, j + 1, 1)) / (DX1[i] + DX1[i + 1]))) /
                   (dy1[j] + dy1[j + 1]);
      real evsy1 = (dy1[j + 1] * ((DX1[i + 1] * SM(i, j - 1, 1) + DX1[i] * SM(i + 1, j - 1, 1)) / (DX1[i] + DX1[i + 1])) +
                    dy1[j] * ((DX1[i + 1] * SM(i, j, 1) + DX1[i] * SM(i + 1, j, 1)) / (DX1[i] + DX1[i + 1]))) /
                   (dy1[j] + dy1[j + 1]);
      real evsz2 = (DZN[2] * ((DX1[i + 1] * SM(i, j, 1) + DX1[i] * SM(i + 1, j, 1)) / (DX1[i] + DX1[i + 1])) +
                    DZN[1] * ((DX1[i + 1] * SM(i, j, 2) + DX1[i] * SM(i + 1, j, 2)) / (DX1[i] + DX1[i + 1]))) /
                   (DZN[1] + DZN[2]);
      real visux2 = (evsx2) * C(2.) * D1(diu1, i + 1, j, 1);
      real visux1 = (evsx1) * C(2.) * D1(diu1, i, j, 1);
      real visuy2 = (evsy2) * (DD(diu2, i, j + 1, 1) + DD(diu4, i + 1, j, 1));
      real visuy1 = (evsy1) * (DD(diu2, i, j, 1) + DD(diu4, i + 1, j - 1, 1));
      real visuz2 = (evsz2) * (DD(diu3, i, j, 2) + DD(diu7, i + 1, j, 1));
      real t = C(0.4) * AT(uspd, L2, i, j, 0) / wl;
      real visuz1 = t * t * AT(uspd, L2, i, j, 0);
      real vfu = (visux2 - visux1) / dxs[i] + (visuy2 - visuy1) / dy1[j] + (visuz2 - visuz1) / DZN[1] +
                 (visuy2 - visuy1) / dy1[j] + (visuz2 - visuz1) / DZN[1];
      AT(f, LF, i, j, 1) = (AT(f, LF, i, j, 1) + vfu);
    }
  PAR_K
  for (int k = 2; k <= km; k++) /* les.f95:143-174 */
    for (int j = 1; j <= jm; j++)
      for (int i = 1; i <= im; i++) {
        real evsy2 = SM(i, j + 1, k);
        real evsy1 = SM(i, j, k);
        real evsx2 = (dy1[j + 1] * ((DX1[i + 1] * SM(i, j, k) + DX1[i] * SM(i + 1, j, k)) / (DX1[i] + DX1[i + 1])) +
                      dy1[j] * ((DX1[i + 1] * SM(i, j + 1, k) + DX1[i] * SM(i + 1, j + 1, k)) / (DX1[i] + DX1[i + 1]))) /
                     (dy1[j] + dy1[j + 1]);
        real evsx1 = (dy1[j + 1] * ((DX1[i] * SM(i - 1, j, k) + DX1[i - 1] * SM(i, j, k)) / (DX1[i - 1] + DX1[i])) +
                      dy1[j] * ((DX1[i] * SM(i - 1, j + 1, k) + DX1[i - 1] * SM(i, j + 1, k)) / (DX1[i - 1] + DX1[i]))) /
                     (dy1[j] + dy1[j + 1]);
        real evsz2 = (DZN[k + 1] * ((DX1[i + 1] * SM(i, j, k) + DX1[i] * SM(i + 1, j, k)) / (DX1[i] + DX1[i + 1])) +
                      DZN[k] * ((DX1[i + 1] * SM(i, j, k + 1) + DX1[i] * SM(i + 1, j, k + 1)) / (DX1[i] + DX1[i + 1]))) /
                     (DZN[k] + DZN[k + 1]);
        real evsz1 = (DZN[k] * ((DX1[i + 1] * SM(i, j, k - 1) + DX1[i] * SM(i + 1, j, k - 1)) / (DX1[i] + DX1[i + 1])) +
                      DZN[k - 1] * ((DX1[i + 1] * SM(i, j, k) + DX1[i] * SM(i + 1, j, k)) / (DX1[i] + DX1[i + 1]))) /
                     (DZN[k - 1] + DZN[k]);
        real visvx2 = (evsx2) * (DD(diu2, i, j + 1, k) + DD(diu4, i + 1, j, k));
        real visvx1 = (evsx1) * (DD(diu2, i - 1, j + 1, k) + DD(diu4, i, j, k));
        real visvy2 = (evsy2) * C(2.) * D1(diu5, i, j + 1, k);
        real visvy1 = (evsy1) * C(2.) * D1(diu5, i, j, k);
        real visvz2 = (evsz2) * (DD(diu6, i, j, k + 1) + DD(diu8, i, j + 1, k));
        real visvz1 = (evsz1) * (DD(diu6, i, j, k) + DD(diu8, i, j + 1, k - 1));
        real vfv = (visvx2 - visvx1) / DX1[i] + (visvy2 - visvy1) / dys[j] + (visvz2 - visvz1) / DZN[k];
        AT(g, LF, i, j, k) = (AT(g, LF, i, j, k) + vfv);
      }
  for (int j = 1; j <= jm; j++) /* les.f95:178-201, wall function */
    for (int i = 1; i <= im; i++) {
      real evsy2 = SM(i, j + 1, 1);
      real evsy1 = SM(i, j, 1);
      real evsx2 = (dy1[j + 1] * ((DX1[i + 1] * SM(i, j, 1) + DX1[i] * SM(i + 1, j, 1)) / (DX1[i] + DX1[i + 1])) +
                    dy1[j] * ((DX1[i + 1] * SM(i, j + 1, 1) + DX1[i] * SM(i + 1, j + 1, 1)) / (DX1[i] + DX1[i + 1]))) /
                   (dy1[j] + dy1[j + 1]);
      real evsx1 = (dy1[j + 1] * ((DX1[i] * SM(i - 1, j, 1) + DX1[i - 1] * SM(i, j, 1)) / (DX1[i - 1] + DX1[i])) +
                    dy1[j] * ((DX1[i] * SM(i - 1, j + 1, 1) + DX1[i - 1] * SM(i, j + 1, 1)) / (DX1[i - 1] + DX1[i]))) /
                   (dy1[j] + dy1[j + 1]);
      real evsz2 = (DZN[2] * ((DX1[i + 1] * SM(i, j, 1) + DX1[i] * SM(i + 1, j, 1)) / (DX1[i] + DX1[i + 1])) +
                    DZN[1] * ((DX1[i + 1] * SM(i, j, 2) + DX1[i] * SM(i + 1, j, 2)) / (DX1[i] + DX1[i + 1]))) /
                   (DZN[1] + DZN[2]);
      real visvx2 = (evsx2) * (DD(diu2, i, j + 1, 1) + DD(diu4, i + 1, j, 1));
      real visvx1 = (evsx1) * (DD(diu2, i - 1, j + 1, 1) + DD(diu4, i, j, 1));
      real visvy2 = (evsy2) * C(2.) * D1(diu5, i, j + 1, 1);
      real visvy1 = (evsy1) * C(2.) * D1(diu5, i, j, 1);
      real visvz2 = (evsz2) * (DD(diu6, i, j, 2) + DD(diu8, i, j + 1, 1));
      real t = C(0.4) * AT(vspd, L2, i, j, 0) / wl;
      real visvz1 = t * t * AT(vspd, L2, i, j, 0);
      real vfv = (visvx2 - visvx1) / DX1[i] + (visvy2 - visvy1) / dys[j] + (visvz2 - visvz1) / DZN[1];
      AT(g, LF, i, j, 1) = (AT(g, LF, i, j, 1) + vfv);
    }
  PAR_K
  for (int k = 1; k <= km; k++) /* les.f95:206-242 */
    for (int j = 1; j <= jm; j++)
      for (int i = 1; i <= im; i++) {
        real evsz2 = SM(i, j, k + 1);
        real evsz1 = SM(i, j, k);
        real evsx2 = (DZN[k + 1] * ((DX1[i + 1] * SM(i, j, k) + DX1[i] * SM(i + 1, j, k)) / (DX1[i] + DX1[i + 1])) +
                      DZN[k] * ((DX1[i + 1] * SM(i, j, k + 1) + DX1[i] * SM(i + 1, j, k + 1)) / (DX1[i] + DX1[i + 1]))) /
                     (DZN[k] + DZN[k + 1]);
        real evsx1 = (DZN[k + 1] * ((DX1[i] * SM(i - 1, j, k) + DX1[i - 1] * SM(i, j, k)) / (DX1[i - 1] + DX1[i])) +
                      DZN[k] * ((DX1[i] * SM(i - 1, j, k + 1) + DX1[i - 1] * SM(i, j, k + 1)) / (DX1[i - 1] + DX1[i]))) /
                     (DZN[k] + DZN[k + 1]);
        real evsy2 = (DZN[k + 1] * ((dy1[j + 1] * SM(i, j, k) + dy1[j] * SM(i, j + 1, k)) / (dy1[j] + dy1[j + 1])) +
                      DZN[k] * ((dy1[j + 1] * SM(i, j, k + 1) + dy1[j] * SM(i, j + 1, k + 1)) / (dy1[j] + dy1[j + 1]))) /
                     (DZN[k] + DZN[k + 1]);
        real evsy1 = (DZN[k + 1] * ((dy1[j] * SM(i, j - 1, k) + dy1[j - 1] * SM(i, j, k)) / (dy1[j - 1] + dy1[j])) +
                      DZN[k] * ((dy1[j] * SM(i, j - 1, k + 1) + dy1[j - 1] * SM(i, j, k + 1)) / (dy1[j - 1] + dy1[j]))) /
                     (DZN[k] + DZN[k + 1]);
        real viswx2 = (evsx2) * (DD(diu3, i, j, k + 1) + DD(diu7, i + 1, j, k));
        real viswx1 = (evsx1) * (DD(diu3, i - 1, j, k + 1) + DD(diu7, i, j, k));
        real viswy2 = (evsy2) * (DD(diu6, i, j, k + 1) + DD(diu8, i, j + 1, k));
        real viswy1 = (evsy1) * (DD(diu6, i, j - 1, k + 1) + DD(diu8, i, j, k));
        real viswz2 = (evsz2) * C(2.) * DD(diu9, i, j, k + 1);
        real viswz1 = (evsz1) * C(2.) * DD(diu9, i, j, k);
        real vfw = (viswx2 - viswx1) / DX1[i] + (viswy2 - viswy1) / dy1[j] + (viswz2 - viswz1) / DZN[k];
        AT(h, LF, i, j, k) = (AT(h, LF, i, j, k) + vfw);
      }
#undef SM
#undef D1
#undef DD
}

/* ----------------------------------------------------------------- adam */
/* adam.f95:5-51 (the data21 dump at adam.f95:20-29 is file output, not restated) */
void ORC(adam)(int n, int nmax, real* fold, int im, int jm, int km, real* gold, real* hold,
               const real* fghold, real* f, real* g, real* h) {
  lay LF = layF(im, jm), LI = layI(im, jm);
  (void)n;
  (void)nmax;
  (void)fghold;
  PAR_K
  for (int k = 1; k <= km; k++) /* adam.f95:30-44 */
    for (int j = 1; j <= jm; j++)
      for (int i = 1; i <= im; i++) {
        real fd = AT(f, LF, i, j, k), gd = AT(g, LF, i, j, k), hd = AT(h, LF, i, j, k);
        AT(f, LF, i, j, k) = C(1.5) * AT(f, LF, i, j, k) - C(0.5) * AT(fold, LI, i, j, k);
        AT(g, LF, i, j, k) = C(1.5) * AT(g, LF, i, j, k) - C(0.5) * AT(gold, LI, i, j, k);
        AT(h, LF, i, j, k) = C(1.5) * AT(h, LF, i, j, k) - C(0.5) * AT(hold, LI, i, j, k);
        AT(fold, LI, i, j, k) = fd;
        AT(gold, LI, i, j, k) = gd;
        AT(hold, LI, i, j, k) = hd;
      }
}

/* --------------------------------------------------------- bondfg/boundp */
/* bondFG.f95:5-49 (serial branch) */
void ORC(bondfg)(int km, int jm, real* f, int im, real* g, real* h) {
  lay LF = layF(im, jm);
  for (int k = 1; k <= km; k++) /* bondFG.f95:18-22 */
    for (int j = 1; j <= jm; j++) AT(f, LF, 0, j, k) = AT(f, LF, 1, j, k);
  for (int k = 1; k <= km; k++) /* bondFG.f95:28-32 (one-sided, Q13) */
    for (int i = 1; i <= im; i++) AT(g, LF, i, 0, k) = AT(g, LF, i, jm, k);
  for (int j = 1; j <= jm; j++) /* bondFG.f95:37-42 */
    for (int i = 1; i <= im; i++) {
      AT(h, LF, i, j, 0) = C(0.0);
      AT(h, LF, i, j, km) = C(0.0);
    }
}

/* boundp.f95:7-26 */
void ORC(boundp2)(int jm, int im, real* p, int km) {
  lay LP = layP(im, jm);
  for (int j = 0; j <= jm + 1; j++)
    for (int i = 0; i <= im + 1; i++) {
      AT(p, LP, i, j, 0) = AT(p, LP, i, j, 1);
      AT(p, LP, i, j, km + 1) = AT(p, LP, i, j, km);
    }
}

/* boundp.f95:28-78 (serial branch: both i faces, then periodic j) */
void ORC(boundp1)(int km, int jm, real* p, int im) {
  lay LP = layP(im, jm);
  for (int k = 0; k <= km + 1; k++)
    for (int j = 0; j <= jm + 1; j++) {
      AT(p, LP, 0, j, k) = AT(p, LP, 1, j, k);
      AT(p, LP, im + 1, j, k) = AT(p, LP, im, j, k);
    }
  for (int k = 0; k <= km + 1; k++)
    for (int i = 0; i <= im + 1; i++) {
      AT(p, LP, i, 0, k) = AT(p, LP, i, jm, k);
      AT(p, LP, i, jm + 1, k) = AT(p, LP, i, 1, k);
    }
}

/* ---------------------------------------------------------------- press */
#define SOR_HIST_MAX 4096
static int g_nmaxp = 50;        /* press.f95:44 */
static double g_pjuge = 0.0001; /* press.f95:43 */
static double g_omega = 1.;     /* press.f95:45 */
static int g_redmode = 0;
static int g_iters = 0;
static double g_sor_hist[SOR_HIST_MAX];
static double g_rhsav = 0, g_area = 0, g_pav = 0, g_pco = 0;

void ORC(press_config)(int nmaxp, double pjuge, double omega, int redmode) {
  g_nmaxp = nmaxp < SOR_HIST_MAX ? nmaxp : SOR_HIST_MAX;
  g_pjuge = pjuge;
  g_omega = omega;
  g_redmode = redmode;
}

void ORC(press_stats)(int* iters, double* sor_hist, int hist_len, double* rhsav, double* area,
                      double* pav, double* pco) {
  if (iters) *iters = g_iters;
  for (int l = 0; l < hist_len && l < g_iters; l++) sor_hist[l] = g_sor_hist[l];
  if (rhsav) *rhsav = g_rhsav;
  if (area) *area = g_area;
  if (pav) *pav = g_pav;
  if (pco) *pco = g_pco;
}

/* press.f95:9-220 (VERBOSE prints and the data20 dump are not restated) */
void ORC(press)(int km, int jm, int im, real* rhs, const real* u, const real* dx1, const real* v,
                const real* dy1, const real* w, const real* dzn, real* f, real* g, real* h, real dt,
                const real* cn1, const real* cn2l, real* p, const real* cn2s, const real* cn3l,
                const real* cn3s, const real* cn4l, const real* cn4s, int n, int nmax,
                const real* usum, const real* vsum, const real* wsum) {
  lay LR = layR(im, jm), LU = layU(im, jm), LW = layW(im, jm), LF = layF(im, jm), LP = layP(im, jm);
  lay LI = layI(im, jm);
  const real* DX1 = dx1 + 1;
  const real* DZN = dzn + 1;
  const real* CN2L = cn2l - 1; /* cn2l(ip) */
  const real* CN2S = cn2s - 1;
  const real* CN3L = cn3l - 1;
  const real* CN3S = cn3s - 1;
  const real* CN4L = cn4l - 1;
  const real* CN4S = cn4s - 1;
  const real pjuge = (real)g_pjuge, omega = (real)g_omega;
  /* FP64 accumulation mode only exists for the float instantiation; the
   * double instantiation accumulates in double anyway. */
  const int dmode = (sizeof(real) == sizeof(double)) || g_redmode == 1;
  (void)n;
  (void)nmax;
  (void)usum;
  (void)vsum;
  (void)wsum;

  ORC(bondfg)(km, jm, f, im, g, h); /* press.f95:47 */
  PAR_K
  for (int k = 1; k <= km; k++) /* press.f95:49-60 */
    for (int j = 1; j <= jm; j++)
      for (int i = 1; i <= im; i++) {
        real r = (-AT(u, LU, i - 1, j, k) + AT(u, LU, i, j, k)) / DX1[i] +
                 (-AT(v, LU, i, j - 1, k) + AT(v, LU, i, j, k)) / dy1[j] +
                 (-AT(w, LW, i, j, k - 1) + AT(w, LW, i, j, k)) / DZN[k];
        AT(rhs, LR, i, j, k) = (AT(f, LF, i, j, k) - AT(f, LF, i - 1, j, k)) / DX1[i] +
                               (AT(g, LF, i, j, k) - AT(g, LF, i, j - 1, k)) / dy1[j] +
                               (AT(h, LF, i, j, k) - AT(h, LF, i, j, k - 1)) / DZN[k] + r / dt;
      }
  /* press.f95:63-87: volume-weighted mean of rhs, subtracted */
  if (dmode) {
    double rs = 0.0, ar = 0.0;
    PAR_K_SUM(rs, ar)
    for (int k = 1; k <= km; k++)
      for (int j = 1; j <= jm; j++)
        for (int i = 1; i <= im; i++) {
          rs = rs + (double)(DX1[i] * dy1[j] * DZN[k] * AT(rhs, LR, i, j, k));
          ar = ar + (double)(DX1[i] * dy1[j] * DZN[k]);
        }
    rs = rs / ar;
    g_rhsav = rs;
    g_area = ar;
    PAR_K
    for (int k = 1; k <= km; k++)
      for (int j = 1; j <= jm; j++)
        for (int i = 1; i <= im; i++) AT(rhs, LR, i, j, k) = (real)((double)AT(rhs, LR, i, j, k) - rs);
  } else {
    real rs = C(0.0), ar = C(0.0);
    for (int k = 1; k <= km; k++)
      for (int j = 1; j <= jm; j++)
        for (int i = 1; i <= im; i++) {
          rs = rs + DX1[i] * dy1[j] * DZN[k] * AT(rhs, LR, i, j, k);
          ar = ar + DX1[i] * dy1[j] * DZN[k];
        }
    rs = rs / ar;
    g_rhsav = rs;
    g_area = ar;
    for (int k = 1; k <= km; k++)
      for (int j = 1; j <= jm; j++)
        for (int i = 1; i <= im; i++) AT(rhs, LR, i, j, k) = AT(rhs, LR, i, j, k) - rs;
  }
  /* press.f95:89-122: red-black SOR; nrd=0 updates cells with i+j+k odd */
  int l;
  g_iters = 0;
  for (l = 1; l <= g_nmaxp; l++) {
    double sord = 0.0;
    real sorf = C(0.0);
    for (int nrd = 0; nrd <= 1; nrd++) {
      PAR_K_SUM(sord, sorf)
      for (int k = 1; k <= km; k++)
        for (int j = 1; j <= jm; j++)
          for (int i = 1 + (k + j + nrd) % 2; i <= im; i += 2) {
            real reltmp = omega * (AT(cn1, LI, i, j, k) *
                                       (CN2L[i] * AT(p, LP, i + 1, j, k) + CN2S[i] * AT(p, LP, i - 1, j, k) +
                                        CN3L[j] * AT(p, LP, i, j + 1, k) + CN3S[j] * AT(p, LP, i, j - 1, k) +
                                        CN4L[k] * AT(p, LP, i, j, k + 1) + CN4S[k] * AT(p, LP, i, j, k - 1) -
                                        AT(rhs, LR, i, j, k)) -
                                   AT(p, LP, i, j, k));
            AT(p, LP, i, j, k) = AT(p, LP, i, j, k) + reltmp;
            if (dmode)
              sord = sord + (double)reltmp * (double)reltmp;
            else
              sorf = sorf + reltmp * reltmp;
          }
      ORC(boundp1)(km, jm, p, im); /* press.f95:105 */
    }
    ORC(boundp2)(jm, im, p, km); /* press.f95:107 */
    double sor = dmode ? sord : (double)sorf;
    g_sor_hist[l - 1] = sor;
    g_iters = l;
    if (dmode ? (sor < (double)pjuge) : (sorf < pjuge)) break; /* press.f95:119-121 */
  }
  /* press.f95:124-150: volume-weighted mean of p, subtracted */
  if (dmode) {
    double pa = 0.0, pc = 0.0;
    PAR_K_SUM(pa, pc)
    for (int k = 1; k <= km; k++)
      for (int j = 1; j <= jm; j++)
        for (int i = 1; i <= im; i++) {
          pa = pa + (double)(AT(p, LP, i, j, k) * DX1[i] * dy1[j] * DZN[k]);
          pc = pc + (double)(DX1[i] * dy1[j] * DZN[k]);
        }
    pa = pa / pc;
    g_pav = pa;
    g_pco = pc;
    PAR_K
    for (int k = 1; k <= km; k++)
      for (int j = 1; j <= jm; j++)
        for (int i = 1; i <= im; i++) AT(p, LP, i, j, k) = (real)((double)AT(p, LP, i, j, k) - pa);
  } else {
    real pa = C(0.0), pc = C(0.0);
    for (int k = 1; k <= km; k++)
      for (int j = 1; j <= jm; j++)
        for (int i = 1; i <= im; i++) {
          pa = pa + AT(p, LP, i, j, k) * DX1[i] * dy1[j] * DZN[k];
          pc = pc + DX1[i] * dy1[j] * DZN[k];
        }
    pa = pa / pc;
    g_pav = pa;
    g_pco = pc;
    for (int k = 1; k <= km; k++)
      for (int j = 1; j <= jm; j++)
        for (int i = 1; i <= im; i++) AT(p, LP, i, j, k) = AT(p, LP, i, j, k) - pa;
  }
  ORC(boundp1)(km, jm, p, im); /* press.f95:155 */
  ORC(boundp2)(jm, im, p, km); /* press.f95:159 */
}

/* --------------------------------------------------------------- aveflow */
/* aveflow.f95:5-354, the statistics part (one process, ipmax=ip, jpmax=jp;
 * the data10/data11 writes at :207-345 are the caller's).  Layouts:
 * aveu, avev (ip,jp,0:kp); avew (ip+1,jp,0:kp+2); every other statistic
 * (ip,jp,kp).  avel is zeroed at n==n1 and never accumulated (:86). */
void ORC(aveflow)(int n, int n1, int km, int jm, int im, real* aveu, real* avev, real* avew,
                  real* avep, real* avel, real* aveuu, real* avevv, real* aveww, real* avesm,
                  real* avesmsm, real* uwfx, const real* u, const real* v, const real* w,
                  const real* p, const real* sm, int nmax) {
  lay LU = layU(im, jm), LW = layW(im, jm), LP = layP(im, jm), LS = layS(im, jm), LI = layI(im, jm);
  lay LA = {1, 1, 0, im, jm};      /* aveu, avev (ip,jp,0:kp) */
  lay LAW = {1, 1, 0, im + 1, jm}; /* avew (ip+1,jp,0:kp+2) */
  if (n == n1) { /* :80-113 */
    for (int k = 1; k <= km; k++)
      for (int j = 1; j <= jm; j++)
        for (int i = 1; i <= im; i++) {
          AT(avev, LA, i, j, k) = C(0.0);
          AT(avep, LI, i, j, k) = C(0.0);
          AT(avel, LI, i, j, k) = C(0.0);
          AT(aveuu, LI, i, j, k) = C(0.0);
          AT(avevv, LI, i, j, k) = C(0.0);
          AT(aveww, LI, i, j, k) = C(0.0);
          AT(avesm, LI, i, j, k) = C(0.0);
          AT(avesmsm, LI, i, j, k) = C(0.0);
          AT(uwfx, LI, i, j, k) = C(0.0);
        }
    for (int k = 0; k <= km; k++)
      for (int j = 1; j <= jm; j++)
        for (int i = 1; i <= im; i++) AT(aveu, LA, i, j, k) = C(0.0);
    for (int k = 0; k <= km; k++)
      for (int j = 1; j <= jm; j++)
        for (int i = 1; i <= im + 1; i++) AT(avew, LAW, i, j, k) = C(0.0);
  }
  if (n >= n1) { /* :116-149 */
    for (int k = 1; k <= km; k++)
      for (int j = 1; j <= jm; j++)
        for (int i = 1; i <= im; i++) {
          const real uc = AT(u, LU, i, j, k), vc = AT(v, LU, i, j, k), wc = AT(w, LW, i, j, k);
          const real s = AT(sm, LS, i, j, k);
          AT(avev, LA, i, j, k) = AT(avev, LA, i, j, k) + vc;
          AT(avep, LI, i, j, k) = AT(avep, LI, i, j, k) + AT(p, LP, i, j, k);
          AT(aveuu, LI, i, j, k) = AT(aveuu, LI, i, j, k) + uc * uc;
          AT(avevv, LI, i, j, k) = AT(avevv, LI, i, j, k) + vc * vc;
          AT(aveww, LI, i, j, k) = AT(aveww, LI, i, j, k) + wc * wc;
          AT(avesm, LI, i, j, k) = AT(avesm, LI, i, j, k) + s;
          AT(avesmsm, LI, i, j, k) = AT(avesmsm, LI, i, j, k) + s * s;
          AT(uwfx, LI, i, j, k) = AT(uwfx, LI, i, j, k) +
                                  C(0.5) * (AT(u, LU, i, j, k - 1) + uc) * C(0.5) *
                                      (AT(w, LW, i, j, k - 1) + AT(w, LW, i + 1, j, k - 1));
        }
    for (int k = 0; k <= km; k++)
      for (int j = 1; j <= jm; j++)
        for (int i = 1; i <= im; i++) AT(aveu, LA, i, j, k) = AT(aveu, LA, i, j, k) + AT(u, LU, i, j, k);
    for (int k = 0; k <= km; k++)
      for (int j = 1; j <= jm; j++)
        for (int i = 1; i <= im + 1; i++)
          AT(avew, LAW, i, j, k) = AT(avew, LAW, i, j, k) + AT(w, LW, i, j, k);
  }
  if (n == nmax) { /* :151-203 */
    const real c = (real)(nmax - n1 + 1); /* float(nmax-n1+1) */
    for (int k = 1; k <= km; k++)
      for (int j = 1; j <= jm; j++)
        for (int i = 1; i <= im; i++) {
          AT(avev, LA, i, j, k) = AT(avev, LA, i, j, k) / c;
          AT(avep, LI, i, j, k) = AT(avep, LI, i, j, k) / c;
          AT(avel, LI, i, j, k) = AT(avel, LI, i, j, k) / c;
          AT(aveuu, LI, i, j, k) = AT(aveuu, LI, i, j, k) / c;
          AT(avevv, LI, i, j, k) = AT(avevv, LI, i, j, k) / c;
          AT(aveww, LI, i, j, k) = AT(aveww, LI, i, j, k) / c;
          AT(avesm, LI, i, j, k) = AT(avesm, LI, i, j, k) / c;
          AT(avesmsm, LI, i, j, k) = AT(avesmsm, LI, i, j, k) / c;
        }
    for (int k = 0; k <= km; k++)
      for (int j = 1; j <= jm; j++)
        for (int i = 1; i <= im; i++) AT(aveu, LA, i, j, k) = AT(aveu, LA, i, j, k) / c;
    for (int k = 0; k <= km; k++)
      for (int j = 1; j <= jm; j++)
        for (int i = 1; i <= im + 1; i++) AT(avew, LAW, i, j, k) = AT(avew, LAW, i, j, k) / c;
    for (int k = 1; k <= km; k++)
      for (int j = 1; j <= jm; j++)
        for (int i = 1; i <= im; i++)
          AT(uwfx, LI, i, j, k) = AT(uwfx, LI, i, j, k) / c -
                                  C(0.5) * (AT(aveu, LA, i, j, k - 1) + AT(aveu, LA, i, j, k)) * C(0.5) *
                                      (AT(avew, LAW, i, j, k - 1) + AT(avew, LAW, i + 1, j, k - 1));
    for (int k = 1; k <= km; k++)
      for (int j = 1; j <= jm; j++)
        for (int i = 1; i <= im; i++) {
          const real au = AT(aveu, LA, i, j, k), av = AT(avev, LA, i, j, k), aw = AT(avew, LAW, i, j, k);
          AT(aveuu, LI, i, j, k) = SQRT(fabs(AT(aveuu, LI, i, j, k) - au * au));
          AT(avevv, LI, i, j, k) = SQRT(fabs(AT(avevv, LI, i, j, k) - av * av));
          AT(aveww, LI, i, j, k) = SQRT(fabs(AT(aveww, LI, i, j, k) - aw * aw));
        }
  }
}
