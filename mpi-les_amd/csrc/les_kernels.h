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

struct DevDomain {
  Grid g;
  Metrics m;
  DevFields F;
  Ctrl* ctrl;
  double* partials;
  int npartials;
  hipStream_t stream;
  float wl;          // logf(0.5*dzn(1)/0.1), the wall-function log (vel2.f95:138, les.f95:133)
  double area;       // sum of dx1*dy1*dzn over the interior (press.f95:63-72)
  int nmaxp;         // SOR iteration cap (press.f95:44)
  float pjuge;       // SOR stop threshold on sum(reltmp^2) (press.f95:43)
  float omega;       // SOR relaxation (press.f95:45)
};

void launch_velnw(const DevDomain& d, float ro, float dt);
void launch_bondv1(const DevDomain& d, int n, float dt);
void launch_velfg(const DevDomain& d, float vn);
void launch_feedbf(const DevDomain& d, float alpha, float dt, float beta);
void launch_les(const DevDomain& d);
void launch_adam(const DevDomain& d);
void launch_press(const DevDomain& d, float dt);

// layout conversion between a packed reference-layout staging buffer and
// the device layout: the reference array covers [lo, lo+n) per axis
void launch_unpack(const DevDomain& d, const float* src, float* dst, int li, int lj, int lk, int ni,
                   int nj, int nk);
void launch_pack(const DevDomain& d, const float* src, float* dst, int li, int lj, int lk, int ni,
                 int nj, int nk);

}  // namespace les
