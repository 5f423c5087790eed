// Host runtime of the LES step library: device domains, the mirrors of
// host arrays handed to the reference-signature entry points, layout
// conversion, and the C ABI declared in include/les_amd.h.
#include <math.h>
#include <cmath>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/les_amd.h"
#include "les_kernels.h"

using namespace les;

namespace {

thread_local std::string g_err;

int fail(const std::string& msg) {
  g_err = msg;
  return 1;
}

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t _e = (expr);                                                        \
    if (_e != hipSuccess)                                                          \
      return fail(std::string(#expr) + ": " + hipGetErrorString(_e));              \
  } while (0)

// ------------------------------------------------------------ device memory
// Blocks released by destroyed domains and batches are kept for reuse rather
// than returned with hipFree: hipFree waits for every stream of the device,
// so a server creating and destroying domains while others step would drain
// the GPU at each release, and from another thread that wait breaks a graph
// capture under way.  A request takes the smallest kept block of its device
// that holds it with at most an eighth to spare, else new memory; when the
// device is out of memory the kept blocks are returned and the allocation
// is tried once more.  les_release_all returns them as well.  g_capture_mu
// orders the calls that may wait for the whole device (hipMalloc, hipFree,
// stream creation and destruction) against graph captures.
std::mutex g_pool_mu;
std::multimap<std::pair<int, size_t>, void*> g_pool_kept;  // (device, bytes) -> block
std::map<void*, std::pair<int, size_t>> g_pool_live;
std::recursive_mutex g_capture_mu;

void pool_trim() {
  std::lock_guard<std::recursive_mutex> ck(g_capture_mu);
  std::lock_guard<std::mutex> lk(g_pool_mu);
  for (auto& kv : g_pool_kept) hipFree(kv.second);
  g_pool_kept.clear();
}

hipError_t dev_alloc(void** p, size_t nb) {
  *p = nullptr;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  nb = std::max<size_t>(nb, 256);
  {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    auto it = g_pool_kept.lower_bound({dev, nb});
    if (it != g_pool_kept.end() && it->first.first == dev && it->first.second - nb <= nb / 8) {
      *p = it->second;
      g_pool_live[*p] = it->first;
      g_pool_kept.erase(it);
      return hipSuccess;
    }
  }
  {
    std::lock_guard<std::recursive_mutex> ck(g_capture_mu);
    e = hipMalloc(p, nb);
  }
  if (e == hipErrorOutOfMemory) {
    (void)hipGetLastError();
    pool_trim();
    std::lock_guard<std::recursive_mutex> ck(g_capture_mu);
    e = hipMalloc(p, nb);
  }
  if (e == hipSuccess) {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    g_pool_live[*p] = {dev, nb};
  } else {
    *p = nullptr;
  }
  return e;
}

template <class T>
hipError_t dev_alloc(T** p, size_t nb) {
  void* v = nullptr;
  const hipError_t e = dev_alloc(&v, nb);
  *p = (T*)v;
  return e;
}

// The block must no longer be used by any stream: its next owner may write
// it at once.
void dev_release(void* p) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(g_pool_mu);
  auto it = g_pool_live.find(p);
  if (it == g_pool_live.end()) return;
  g_pool_kept.emplace(it->second, p);
  g_pool_live.erase(it);
}

hipError_t stream_create(hipStream_t* s) {
  std::lock_guard<std::recursive_mutex> ck(g_capture_mu);
  return hipStreamCreateWithFlags(s, hipStreamNonBlocking);
}

void stream_destroy(hipStream_t s) {
  if (!s) return;
  std::lock_guard<std::recursive_mutex> ck(g_capture_mu);
  hipStreamDestroy(s);
}

// A blocking copy ordered on stream s: the null stream's implicit
// synchronisation would reach graph captures under way on other threads.
hipError_t copy_sync(void* dst, const void* src, size_t nb, hipMemcpyKind kind, hipStream_t s) {
  const hipError_t e = hipMemcpyAsync(dst, src, nb, kind, s);
  return e == hipSuccess ? hipStreamSynchronize(s) : e;
}

// ------------------------------------------------------------ field table
enum Kind { KU, KW, KP, KF, KI, KR, KS, KD1, KD, KB, K2D, KAU, KAW };

struct KindInfo {
  int li, lj, lk;    // reference lower bounds
  int ei, ej, ek;    // extents minus (im, jm, km)
};
// main.f95:63-189 and the callee dummy declarations
const KindInfo kKinds[] = {
    {0, -1, 0, 2, 3, 2},    // KU  u, v, cmask1   (0:ip+1,-1:jp+1,0:kp+1)
    {0, -1, -1, 2, 3, 3},   // KW  w              (0:ip+1,-1:jp+1,-1:kp+1)
    {0, 0, 0, 3, 3, 2},     // KP  p              (0:ip+2,0:jp+2,0:kp+1)
    {0, 0, 0, 1, 1, 1},     // KF  f,g,h,usum..   (0:ip,0:jp,0:kp)
    {1, 1, 1, 0, 0, 0},     // KI  fold..,cn1     (ip,jp,kp)
    {0, 0, 0, 2, 2, 2},     // KR  rhs, dmask1    (0:ip+1,0:jp+1,0:kp+1)
    {-1, -1, 0, 3, 3, 2},   // KS  sm             (-1:ip+1,-1:jp+1,0:kp+1)
    {-1, 0, 0, 4, 3, 3},    // KD1 diu1, diu5     (-1:ip+2,0:jp+2,0:kp+2)
    {0, 0, 0, 3, 3, 3},     // KD  other diu      (0:ip+2,0:jp+2,0:kp+2)
    {-1, 0, 0, 3, 2, 2},    // KB  bmask1         (-1:ip+1,0:jp+1,0:kp+1)
    {0, 0, -1, 2, 2, -1},   // K2D uspd, vspd     (0:ip+1,0:jp+1), stored in plane k=-1
    {1, 1, 0, 0, 0, 1},     // KAU aveu, avev     (ip,jp,0:kp)        aveflow.f95:18,20
    {1, 1, 0, 1, 0, 3},     // KAW avew           (ip+1,jp,0:kp+2)    aveflow.f95:22
};

enum Field {
  FU, FV, FW, FP, FF, FG, FH, FUSUM, FVSUM, FWSUM, FFOLD, FGOLD, FHOLD, FRHS, FSM, FUSPD, FVSPD,
  FDIU1, FDIU2, FDIU3, FDIU4, FDIU5, FDIU6, FDIU7, FDIU8, FDIU9, FBMASK, FCMASK, FDMASK, FCN1,
  // the time statistics of aveflow (aveflow.f95:8-23), allocated on first use
  FAVEU, FAVEV, FAVEW, FAVEP, FAVEL, FAVEUU, FAVEVV, FAVEWW, FAVESM, FAVESMSM, FUWFX,
  NFIELDS
};
constexpr int kLazyFirst = FAVEU;  // fields from here on are allocated when first used
struct FieldInfo {
  const char* name;
  Kind kind;
};
const FieldInfo kFields[NFIELDS] = {
    {"u", KU},       {"v", KU},       {"w", KW},       {"p", KP},       {"f", KF},
    {"g", KF},       {"h", KF},       {"usum", KF},    {"vsum", KF},    {"wsum", KF},
    {"fold", KI},    {"gold", KI},    {"hold", KI},    {"rhs", KR},     {"sm", KS},
    {"uspd", K2D},   {"vspd", K2D},   {"diu1", KD1},   {"diu2", KD},    {"diu3", KD},
    {"diu4", KD},    {"diu5", KD1},   {"diu6", KD},    {"diu7", KD},    {"diu8", KD},
    {"diu9", KD},    {"bmask1", KB},  {"cmask1", KU},  {"dmask1", KR},  {"cn1", KI},
    {"aveu", KAU},   {"avev", KAU},   {"avew", KAW},   {"avep", KI},    {"avel", KI},
    {"aveuu", KI},   {"avevv", KI},   {"aveww", KI},   {"avesm", KI},   {"avesmsm", KI},
    {"uwfx", KI},
};

int field_by_name(const char* n) {
  for (int f = 0; f < NFIELDS; f++)
    if (strcmp(kFields[f].name, n) == 0) return f;
  return -1;
}

// 1-D metrics: reference lower bound, length minus n, which n (0 i, 1 j, 2 k)
enum Metric {
  MDX1, MDY1, MDZN, MDZS, MDXS, MDYS, MCN2L, MCN2S, MCN3L, MCN3S, MCN4L, MCN4S, MDELX1,
  MCNAX, MCNBY, MCNCZ, NMETRICS
};
struct MetricInfo {
  int lb, elen, axis;
};
const MetricInfo kMetrics[NMETRICS] = {
    {-1, 3, 0},  // dx1(-1:ip+1)
    {0, 2, 1},   // dy1(0:jp+1)
    {-1, 4, 2},  // dzn(-1:kp+2)
    {-1, 4, 2},  // dzs(-1:kp+2)
    {0, 1, 0},   // dxs(0:ip)
    {0, 1, 1},   // dys(0:jp)
    {1, 0, 0},   // cn2l(ip)
    {1, 0, 0},   // cn2s(ip)
    {1, 0, 1},   // cn3l(jp)
    {1, 0, 1},   // cn3s(jp)
    {1, 0, 2},   // cn4l(kp)
    {1, 0, 2},   // cn4s(kp)
    {1, 0, 2},   // delx1(kp)
    {1, 0, 0},   // cnax(ip)   derived
    {1, 0, 1},   // cnby(jp)   derived
    {1, 0, 2},   // cncz(kp)   derived
};

// Event pairs recorded around launches while timing is on.
struct Timer {
  std::vector<hipEvent_t> ev;  // 2 per pair
  std::vector<int> cls;        // class of each pair
  std::vector<int> n;          // launches each pair brackets
  size_t used = 0;             // pairs recorded since the last read
  unsigned int mask = ~0u;     // classes to record
  bool open = false;           // a start event is waiting for its end
  bool runs = false;           // one pair per run of back-to-back launches of one class
  bool chain = false;          // the last launch was recorded and ended a pair
  bool extend = false;         // the open launch extends the last pair
};

const char* kClassNames[KC_N] = {"velnw", "bondv1", "vel2", "vel2_ghost", "velfg", "feedbf", "les_sm",
                                 "boundsm", "les_visc", "adam", "rhs", "sor", "boundp", "pav", "padj",
                                 "copy", "sor1"};

}  // namespace

void les::timer_begin(const DevDomain& d, int c) {
  Timer* T = (Timer*)d.timer;
  if (!T) return;
  if (!(T->mask & (1u << c))) {  // an unrecorded launch ends a run
    T->chain = false;
    return;
  }
  if (T->runs && T->chain && T->used > 0 && T->cls[T->used - 1] == c) {
    T->extend = true;  // the previous launch was the same class: move its end event
    T->open = true;
    return;
  }
  if (2 * T->used + 2 > T->ev.size()) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
    T->ev.push_back(a);
    T->ev.push_back(b);
    T->cls.push_back(0);
    T->n.push_back(0);
  }
  T->cls[T->used] = c;
  hipEventRecord(T->ev[2 * T->used], d.stream);
  T->open = true;
}

void les::timer_end(const DevDomain& d, int c) {
  (void)c;
  Timer* T = (Timer*)d.timer;
  if (!T || !T->open) return;
  if (T->extend) {
    hipEventRecord(T->ev[2 * (T->used - 1) + 1], d.stream);
    T->n[T->used - 1]++;
    T->extend = false;
  } else {
    hipEventRecord(T->ev[2 * T->used + 1], d.stream);
    T->n[T->used] = 1;
    T->used++;
  }
  T->open = false;
  T->chain = true;
}

struct les_domain {
  DevDomain d{};
  int n3 = 0;                       // floats per 3-D field
  float* fields[NFIELDS] = {};
  float* cov[10] = {};
  float* mbuf = nullptr;            // metric pack on device
  float* inflow_dev = nullptr;
  std::vector<float> mhost;         // metric pack on host
  int moff[NMETRICS] = {};          // offset of each metric in the pack
  bool mdirty = true;
  bool have_inflow = false;
  float* split = nullptr;           // colour-split SOR arrays
  float* mask_src = nullptr;        // column values and level thresholds of the masks (DevDomain::mask_src)
  int use_graph = 0;                // replay steps n > 1 from a captured graph
  hipGraphExec_t gexec = nullptr;   // captured step
  DevDomain gkey{};                 // domain state the graph was captured with
  float gscal[5] = {};              // ro, vn, alpha, beta, dt at capture
  float* staging = nullptr;         // reference-layout staging buffer
  size_t staging_floats = 0;
  const void* bound[NFIELDS] = {};  // host address each field mirrors (reference-signature layer)
  float ro = 1.1763f, vn = 0.0f, alpha = -10.0f, beta = -1.0f, dt = 0.05f;
  long long bytes = 0;
  // the reference-layout rhs is behind the split rhs (a press of mode 0 ran
  // since it was last written); a download gathers it first
  bool rhs_lazy = false;
  double st_us_best = 0.0, st_us_default = 0.0;  // the stencil tuner's timings for this shape
};

namespace {

int axis_n(const les_domain* D, int axis) { return axis == 0 ? D->d.g.im : (axis == 1 ? D->d.g.jm : D->d.g.km); }

void ref_extent(const les_domain* D, Kind k, int* ni, int* nj, int* nk) {
  const KindInfo& K = kKinds[k];
  *ni = D->d.g.im + K.ei;
  *nj = D->d.g.jm + K.ej;
  *nk = (k == K2D) ? 1 : D->d.g.km + K.ek;
}

size_t ref_floats(const les_domain* D, Kind k) {
  int a, b, c;
  ref_extent(D, k, &a, &b, &c);
  return (size_t)a * b * c;
}

int ensure_staging(les_domain* D, size_t n) {
  if (D->staging_floats >= n) return 0;
  dev_release(D->staging);
  D->staging = nullptr;
  HIP_TRY(dev_alloc(&D->staging, n * sizeof(float)));
  D->staging_floats = n;
  return 0;
}

void derive(les_domain* D);

// cn1 of init.f95:100-107 is 1/(cnax(i)+cnby(j)+cncz(k)); when the array
// handed over equals that bit for bit the SOR recomputes it instead of
// reading 4 bytes per cell per sweep.
int env_int(const char* name, int dflt);

// The kernels recompute cn1 with rcp_rn, which equals the correctly rounded
// 1.0f/x for x in [2^-125, 2^125) (checked exhaustively by
// les_selftest_reciprocal).  Rounded FP32 addition is monotone, so bounding
// the per-axis minima and maxima bounds every (cnax+cnby)+cncz.
bool cn_rcp_range_ok(const les_domain* D) {
  const Grid& g = D->d.g;
  const float* ax = D->mhost.data() + D->moff[MCNAX] + 1;
  const float* by = D->mhost.data() + D->moff[MCNBY] + 1;
  const float* cz = D->mhost.data() + D->moff[MCNCZ] + 1;
  auto lim = [](const float* a, int n, float* lo, float* hi) {
    *lo = a[1];
    *hi = a[1];
    for (int x = 1; x <= n; x++) {
      if (!(a[x] >= 0.0f) || !std::isfinite(a[x])) return false;  // NaN, negative or infinite
      *lo = std::min(*lo, a[x]);
      *hi = std::max(*hi, a[x]);
    }
    return true;
  };
  float l0, h0, l1, h1, l2, h2;
  if (!lim(ax, g.im, &l0, &h0) || !lim(by, g.jm, &l1, &h1) || !lim(cz, g.km, &l2, &h2)) return false;
  const float lo = (l0 + l1) + l2, hi = (h0 + h1) + h2;
  return lo >= 0x1p-125f && hi < 0x1p125f;
}

void check_cn1(les_domain* D, const float* host) {
  derive(D);
  if (env_int("LES_SOR_CN1_LOAD", 0) || !cn_rcp_range_ok(D)) {
    // read cn1 in the SOR: asked for, or its recomputation would not be exact
    D->d.cn1_separable = 0;
    return;
  }
  const Grid& g = D->d.g;
  const float* ax = D->mhost.data() + D->moff[MCNAX] + 1;
  const float* by = D->mhost.data() + D->moff[MCNBY] + 1;
  const float* cz = D->mhost.data() + D->moff[MCNCZ] + 1;
  size_t q = 0;
  for (int k = 1; k <= g.km; k++)
    for (int j = 1; j <= g.jm; j++)
      for (int i = 1; i <= g.im; i++, q++) {
        const float c = 1.0f / ((ax[i] + by[j]) + cz[k]);
        if (memcmp(&c, host + q, sizeof(float)) != 0) {
          D->d.cn1_separable = 0;
          return;
        }
      }
  D->d.cn1_separable = 1;
}

// Fields allocated on first use start at zero, as the reference's static
// arrays do.
int ensure_field(les_domain* D, int f) {
  if (D->fields[f]) return 0;
  const size_t nb = (size_t)D->n3 * sizeof(float);
  float* a = nullptr;
  HIP_TRY(dev_alloc(&a, nb));
  hipError_t e = hipMemsetAsync(a, 0, nb, D->d.stream);
  if (e != hipSuccess) {
    dev_release(a);
    return fail(std::string("field allocation: ") + hipGetErrorString(e));
  }
  D->fields[f] = a;
  D->bytes += nb;
  return 0;
}

int upload_field(les_domain* D, int f, const float* host) {
  if (ensure_field(D, f)) return 1;
  if (f == FCN1) check_cn1(D, host);
  if (f == FRHS) D->rhs_lazy = false;
  const Kind k = kFields[f].kind;
  int ni, nj, nk;
  ref_extent(D, k, &ni, &nj, &nk);
  const size_t n = (size_t)ni * nj * nk;
  if (ensure_staging(D, n)) return 1;
  HIP_TRY(hipMemcpyAsync(D->staging, host, n * sizeof(float), hipMemcpyHostToDevice, D->d.stream));
  const KindInfo& K = kKinds[k];
  launch_unpack(D->d, D->staging, D->fields[f], K.li, K.lj, K.lk, ni, nj, nk);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(D->d.stream));  // the host array may change after return
  if ((f == FBMASK || f == FCMASK || f == FDMASK) && D->mask_src) {
    // an uploaded mask replaces the masks' column form: k_velfg_feedbf reads
    // the mask fields again (the changed DevDomain also retires a captured step)
    D->d.mask_src = nullptr;
    dev_release(D->mask_src);
    D->mask_src = nullptr;
    D->bytes -= ((long long)D->d.g.im * D->d.g.jm + D->d.g.km) * (long long)sizeof(float);
  }
  return 0;
}

int download_field(les_domain* D, int f, float* host) {
  if (ensure_field(D, f)) return 1;
  if (f == FRHS && D->rhs_lazy) {
    launch_rhs_gather(D->d);
    HIP_TRY(hipGetLastError());
    D->rhs_lazy = false;
  }
  const Kind k = kFields[f].kind;
  int ni, nj, nk;
  ref_extent(D, k, &ni, &nj, &nk);
  const size_t n = (size_t)ni * nj * nk;
  if (ensure_staging(D, n)) return 1;
  const KindInfo& K = kKinds[k];
  launch_pack(D->d, D->fields[f], D->staging, K.li, K.lj, K.lk, ni, nj, nk);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(host, D->staging, n * sizeof(float), hipMemcpyDeviceToHost, D->d.stream));
  HIP_TRY(hipStreamSynchronize(D->d.stream));
  return 0;
}

// Copy a reference-layout 1-D metric into the host pack; marks the pack
// dirty when a value changed.
void set_metric(les_domain* D, int mi, const float* host) {
  if (!host) return;
  const MetricInfo& M = kMetrics[mi];
  const int len = axis_n(D, M.axis) + M.elen;
  float* dst = D->mhost.data() + D->moff[mi] + (M.lb + 1);  // slot x+1 holds index x
  if (memcmp(dst, host, len * sizeof(float)) != 0) {
    memcpy(dst, host, len * sizeof(float));
    D->mdirty = true;
    if (mi == MDXS || mi == MDYS || mi == MDZS) D->d.cn1_separable = 0;
  }
}

float metric_at(const les_domain* D, int mi, int x) { return D->mhost[D->moff[mi] + x + 1]; }

// Host-side constants derived from the metrics: delx1 (les.f95:38-42),
// the wall-function log (vel2.f95:138) and sum(dx1*dy1*dzn) (press.f95:63-72,
// accumulated in FP64 in i-fastest order).
void derive(les_domain* D) {
  const Grid& g = D->d.g;
  const float dx0 = metric_at(D, MDX1, 0), dy0 = metric_at(D, MDY1, 0);
  for (int k = 1; k <= g.km; k++) {
    const float v = powf(dx0 * dy0 * metric_at(D, MDZN, k), 1.0f / 3.0f);
    D->mhost[D->moff[MDELX1] + k + 1] = v;
  }
  D->d.wl = logf(0.5f * metric_at(D, MDZN, 1) / 0.1f);
  // the three terms of cn1 (init.f95:100-107), FP32 as written there
  for (int i = 1; i <= g.im; i++)
    D->mhost[D->moff[MCNAX] + i + 1] = 2.0f / (metric_at(D, MDXS, i - 1) * metric_at(D, MDXS, i));
  for (int j = 1; j <= g.jm; j++)
    D->mhost[D->moff[MCNBY] + j + 1] = 2.0f / (metric_at(D, MDYS, j - 1) * metric_at(D, MDYS, j));
  for (int k = 1; k <= g.km; k++)
    D->mhost[D->moff[MCNCZ] + k + 1] = 2.0f / (metric_at(D, MDZS, k - 1) * metric_at(D, MDZS, k));
  double area = 0.0;
  for (int k = 1; k <= g.km; k++) {
    const float dz = metric_at(D, MDZN, k);
    for (int j = 1; j <= g.jm; j++) {
      const float dy = metric_at(D, MDY1, j);
      for (int i = 1; i <= g.im; i++) area = area + (double)(metric_at(D, MDX1, i) * dy * dz);
    }
  }
  D->d.area = area;
}

int flush_metrics(les_domain* D) {
  if (!D->mdirty) return 0;
  derive(D);
  HIP_TRY(hipMemcpyAsync(D->mbuf, D->mhost.data(), D->mhost.size() * sizeof(float),
                         hipMemcpyHostToDevice, D->d.stream));
  HIP_TRY(hipStreamSynchronize(D->d.stream));
  D->mdirty = false;
  return 0;
}

// Inflow planes u_in, v_in, w_in (j,k), j fastest: u_in alone replaces
// bondv1's literal (bondv1.f95:33-54, v = w = 0); all three are the planes
// bondv1_data24 reads from its files (bondv1_data24.f95:62-87).  The copy is
// ordered on the domain's stream, so a step already enqueued still sees the
// previous planes.
int set_inflow(les_domain* D, const float* u_in, const float* v_in = nullptr, const float* w_in = nullptr) {
  const Grid& g = D->d.g;
  Metrics& m = D->d.m;
  if (!u_in) {
    if (v_in || w_in) return fail("inflow: v and w planes need a u plane");
    D->have_inflow = false;
    m.inflow = m.inflow_v = m.inflow_w = nullptr;
    return 0;
  }
  const size_t plane = (size_t)g.jm * g.km;
  if (!D->inflow_dev) {
    HIP_TRY(dev_alloc(&D->inflow_dev, 3 * plane * sizeof(float)));
    D->bytes += 3 * plane * sizeof(float);
  }
  const float* src[3] = {u_in, v_in, w_in};
  for (int c = 0; c < 3; c++)
    if (src[c])
      HIP_TRY(hipMemcpyAsync(D->inflow_dev + c * plane, src[c], plane * sizeof(float), hipMemcpyHostToDevice,
                             D->d.stream));
  D->have_inflow = true;
  m.inflow = D->inflow_dev;
  m.inflow_v = v_in ? D->inflow_dev + plane : nullptr;
  m.inflow_w = w_in ? D->inflow_dev + 2 * plane : nullptr;
  return 0;
}

int env_int(const char* name, int dflt) {
  const char* v = getenv(name);
  return (v && *v) ? atoi(v) : dflt;
}

int nmaxp_default = 50;
float pjuge_default = 1.0e-4f;
float omega_default = 1.0f;

void timer_free(les_domain* D) {
  Timer* T = (Timer*)D->d.timer;
  if (!T) return;
  for (hipEvent_t e : T->ev) hipEventDestroy(e);
  delete T;
  D->d.timer = nullptr;
}

void domain_free(les_domain* D) {
  if (!D) return;
  if (D->d.stream) hipStreamSynchronize(D->d.stream);  // nothing of this domain may still run
  timer_free(D);
  if (D->gexec) hipGraphExecDestroy(D->gexec);
  for (int f = 0; f < NFIELDS; f++) dev_release(D->fields[f]);
  for (int q = 1; q <= 9; q++) dev_release(D->cov[q]);
  dev_release(D->split);
  dev_release(D->mask_src);
  dev_release(D->mbuf);
  dev_release(D->inflow_dev);
  dev_release(D->staging);
  dev_release(D->d.ctrl);
  dev_release(D->d.partials);
  dev_release(D->d.dc64);
  stream_destroy(D->d.stream);
  delete D;
}

int domain_alloc(int im, int jm, int km, les_domain** out) {
  if (im < 2 || jm < 2 || km < 2) return fail("domain extents must be at least 2 in every direction");
  const long sj = row_pitch(im);
  const long sk = sj * (jm + 4);
  const long n3 = sk * (km + 4);
  if (n3 >= (1L << 31)) return fail("domain too large for 32-bit field indexing");
  les_domain* D = new les_domain();
  Grid& g = D->d.g;
  g.im = im;
  g.jm = jm;
  g.km = km;
  g.sj = (int)sj;
  g.sk = (int)sk;
  D->n3 = (int)n3;
  D->d.n3 = (int)n3;
  auto bail = [&](const std::string& m) {
    domain_free(D);
    return fail(m);
  };
  hipError_t e = stream_create(&D->d.stream);
  if (e != hipSuccess) return bail(std::string("hipStreamCreate: ") + hipGetErrorString(e));
  for (int f = 0; f < kLazyFirst; f++) {
    const size_t nb = (kFields[f].kind == K2D ? (size_t)sk : (size_t)n3) * sizeof(float);
    e = dev_alloc(&D->fields[f], nb);
    if (e == hipSuccess) e = hipMemsetAsync(D->fields[f], 0, nb, D->d.stream);
    if (e != hipSuccess) return bail(std::string("field allocation: ") + hipGetErrorString(e));
    D->bytes += nb;
  }
  // cov1..cov9 are scratch of the stage-by-stage velfg only; they are
  // allocated on its first use (ensure_cov), the whole-step path never
  // stores them
  {
    // colour-split SOR arrays: 4 iterate buffers, 2 rhs, 2 cn1, 2 rhs before
    // the mean removal, after one scratch plane and followed by another.  Rows are padded to at least
    // one pair beyond the last (cp > nt): the SOR kernel sends the stores of
    // lanes that own no cell there, and the planes at either end take the
    // loads and stores of levels outside 1..km.  Nothing reads either.
    SplitGrid& sg = D->d.sg;
    sg.nt = (im + 1) / 2;
    sg.cp = (sg.nt + 1 + 31) / 32 * 32;
    sg.cplane = sg.cp * jm;
    const size_t cs = (size_t)sg.cplane * km;
    const size_t nplanes = 10 * (size_t)km + 2;
    const size_t nb = nplanes * sg.cplane * sizeof(float);
    e = dev_alloc(&D->split, nb);
    if (e == hipSuccess) e = hipMemsetAsync(D->split, 0, nb, D->d.stream);
    if (e != hipSuccess) return bail(std::string("SOR buffer allocation: ") + hipGetErrorString(e));
    D->bytes += nb;
    SplitBufs& sb = D->d.sb;
    float* arr0 = D->split + sg.cplane;  // level 1 of the first array
    for (int q = 0; q < 4; q++) sb.p[q >> 1][q & 1] = arr0 + q * cs;
    for (int c = 0; c < 2; c++) {
      sb.rhs[c] = arr0 + (4 + c) * cs;
      sb.cn1[c] = arr0 + (6 + c) * cs;
      sb.p[2][c] = sb.cn1[c];  // the third iterate set of two-iteration launches (separable cn1 only)
      sb.rhsr[c] = arr0 + (8 + c) * cs;  // rhs[0..1] and rhsr[0..1] each one after the other (k_rhs_sub)
    }
    sb.base = D->split;
    sb.nplanes = (int)nplanes;
  }
  // metric pack: each array gets n+4 slots (indices -1..n+2)
  int off = 0;
  for (int mi = 0; mi < NMETRICS; mi++) {
    D->moff[mi] = off;
    off += axis_n(D, kMetrics[mi].axis) + 4;
  }
  D->mhost.assign(off, 0.0f);
  e = dev_alloc(&D->mbuf, off * sizeof(float));
  if (e != hipSuccess) return bail(std::string("metric allocation: ") + hipGetErrorString(e));
  Metrics& m = D->d.m;
  m.dx1 = D->mbuf + D->moff[MDX1];
  m.dy1 = D->mbuf + D->moff[MDY1];
  m.dzn = D->mbuf + D->moff[MDZN];
  m.dzs = D->mbuf + D->moff[MDZS];
  m.dxs = D->mbuf + D->moff[MDXS];
  m.dys = D->mbuf + D->moff[MDYS];
  m.cn2l = D->mbuf + D->moff[MCN2L];
  m.cn2s = D->mbuf + D->moff[MCN2S];
  m.cn3l = D->mbuf + D->moff[MCN3L];
  m.cn3s = D->mbuf + D->moff[MCN3S];
  m.cn4l = D->mbuf + D->moff[MCN4L];
  m.cn4s = D->mbuf + D->moff[MCN4S];
  m.delx1 = D->mbuf + D->moff[MDELX1];
  m.cnax = D->mbuf + D->moff[MCNAX];
  m.cnby = D->mbuf + D->moff[MCNBY];
  m.cncz = D->mbuf + D->moff[MCNCZ];
  m.inflow = nullptr;
  // reduction scratch: reduction partials, then the four SOR partial slots
  // at least the 2 x 256 slots of the bondv1 outflow's block max/min
  D->d.npartials = std::max(2 * reduction_parts_needed(D->d.g), 512);
  D->d.sor_blocks = env_int("LES_SOR_BLOCKS", 2048);
  D->d.xcd_remap = env_int("LES_XCD_REMAP", 1);
  D->d.sor_fused = env_int("LES_SOR_FUSED", 1);
  D->d.fuse_velfg = env_int("LES_FUSE_VELFG", 1);
  D->d.velfg_kchunk = env_int("LES_VELFG_KCHUNK", 0);
  D->d.sm_kchunk = std::max(0, env_int("LES_SM_KCHUNK", 0));
  D->d.visc_kchunk = std::max(0, env_int("LES_VISC_KCHUNK", 0));
  D->d.fuse_les = env_int("LES_FUSE_LES", 1);
  D->d.fuse_sm = env_int("LES_FUSE_SM", 1) != 0 && D->d.fuse_velfg && D->d.fuse_les;
  D->d.sor_fused_blocks = env_int("LES_SOR_FUSED_BLOCKS", 0);
  {
    // fused SOR kernel: 4 = two iterations per launch (default; even jm and
    // separable cn1, else 2), 2 = one iteration, two rows per thread, 1 = one row
    const int r = env_int("LES_SOR_ROWS", 4);
    D->d.sor_rows = (r == 1 || r == 2) ? r : 4;
  }
  D->d.sor_kchunk = std::max(0, env_int("LES_SOR_KCHUNK", 0));
  D->d.sor_tile_rows = env_int("LES_SOR_TILE_ROWS", 16) == 32 ? 32 : 16;
  D->d.sor_tile_width = (D->d.sor_tile_rows == 32 && env_int("LES_SOR_TILE_WIDTH", 64) == 32) ? 32 : 64;
  D->d.sor_entry_fast = env_int("LES_SOR_ENTRY_FAST", 1) != 0;
  // the fused SOR kernels address the split arrays through one buffer
  // descriptor with 32-bit byte offsets; beyond 4 GiB the colour sweeps run
  if ((size_t)D->d.sb.nplanes * D->d.sg.cplane * sizeof(float) > 0xffffffffull) D->d.sor_fused = 0;
  D->d.dc_omega = 1.0f;
  {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess)
      D->d.ncu = ncu;
  }
  // SOR partial slots for any k chunk of the fused iteration (down to one
  // level per block) and for the colour sweeps
  const int nsor = std::max({sor_parts_needed(D->d.g, D->d.sg, D->d.sor_blocks),
                             sor_fused_tiles(D->d.g, D->d.sg) * D->d.g.km,
                             sor_pair_tiles(D->d.g, D->d.sg) * D->d.g.km});
  e = dev_alloc(&D->d.partials, ((size_t)D->d.npartials + 4 * (size_t)nsor) * sizeof(double));
  if (e == hipSuccess)
    for (int q = 0; q < 4; q++) D->d.sb.part[q >> 1][q & 1] = D->d.partials + D->d.npartials + (size_t)q * nsor;
  if (e == hipSuccess) e = dev_alloc(&D->d.ctrl, sizeof(Ctrl));
  if (e == hipSuccess) e = hipMemsetAsync(D->d.ctrl, 0, sizeof(Ctrl), D->d.stream);
  if (e != hipSuccess) return bail(std::string("control allocation: ") + hipGetErrorString(e));
  D->d.nmaxp = nmaxp_default;
  D->d.pjuge = pjuge_default;
  D->d.omega = omega_default;
  DevFields& F = D->d.F;
  F.u = D->fields[FU];
  F.v = D->fields[FV];
  F.w = D->fields[FW];
  F.p = D->fields[FP];
  F.f = D->fields[FF];
  F.g = D->fields[FG];
  F.h = D->fields[FH];
  F.usum = D->fields[FUSUM];
  F.vsum = D->fields[FVSUM];
  F.wsum = D->fields[FWSUM];
  F.fold = D->fields[FFOLD];
  F.gold = D->fields[FGOLD];
  F.hold = D->fields[FHOLD];
  F.rhs = D->fields[FRHS];
  F.sm = D->fields[FSM];
  F.uspd = D->fields[FUSPD];
  F.vspd = D->fields[FVSPD];
  F.diu[0] = nullptr;
  F.cov[0] = nullptr;
  for (int q = 1; q <= 9; q++) {
    F.diu[q] = D->fields[FDIU1 + q - 1];
    F.cov[q] = D->cov[q];
  }
  F.bmask = D->fields[FBMASK];
  F.cmask = D->fields[FCMASK];
  F.dmask = D->fields[FDMASK];
  F.cn1 = D->fields[FCN1];
  D->bytes += (long long)off * 4 + ((long long)D->d.npartials + 4LL * nsor) * 8 + (long long)sizeof(Ctrl);
  *out = D;
  return 0;
}

int run_stage(les_domain* D, int s, int n, int nmax);

// The products cov1..cov9 of the stage kernels of velfg, allocated once.
int ensure_cov(les_domain* D) {
  if (D->cov[1]) return 0;
  const size_t nb = (size_t)D->n3 * sizeof(float);
  for (int q = 1; q <= 9; q++) {
    hipError_t e = dev_alloc(&D->cov[q], nb);
    if (e == hipSuccess) e = hipMemsetAsync(D->cov[q], 0, nb, D->d.stream);
    if (e != hipSuccess) {
      for (int r = 1; r <= q; r++) {
        dev_release(D->cov[r]);
        D->cov[r] = nullptr;
      }
      return fail(std::string("scratch allocation: ") + hipGetErrorString(e));
    }
    D->d.F.cov[q] = D->cov[q];
    D->bytes += nb;
  }
  return 0;
}

// One whole step n (main.f95:260-296); with fuse_velfg the velfg and feedbf
// stages run as one kernel, and with fuse_les adam is applied inside les.
int run_step(les_domain* D, int n, int nmax) {
  for (int s = 0; s < 7; s++) {
    if (s == 2 && D->d.fuse_velfg) {
      if (flush_metrics(D)) return 1;
      launch_velfg_feedbf(D->d, D->vn, D->alpha, D->dt, D->beta);
      HIP_TRY(hipGetLastError());
      s = 3;
      continue;
    }
    if (s == 4 && D->d.fuse_les) {
      if (flush_metrics(D)) return 1;
      launch_les_adam(D->d);
      HIP_TRY(hipGetLastError());
      s = 5;
      continue;
    }
    if (run_stage(D, s, n, nmax)) return 1;
  }
  return 0;
}

// Stages velnw..adam of a step (as run_step), then press: whole when the
// domain's SOR cannot be batched, else its first part (the right-hand side
// and SOR iteration 1, with the batch's SOR k chunk kc and tile rows ty).
// Returns whether the press was split.
int run_step_head(les_domain* D, int n, int nmax, int kc, int ty, int tw, bool* split) {
  for (int s = 0; s < 6; s++) {
    if (s == 2 && D->d.fuse_velfg) {
      if (flush_metrics(D)) return 1;
      launch_velfg_feedbf(D->d, D->vn, D->alpha, D->dt, D->beta);
      HIP_TRY(hipGetLastError());
      s = 3;
      continue;
    }
    if (s == 4 && D->d.fuse_les) {
      if (flush_metrics(D)) return 1;
      launch_les_adam(D->d);
      HIP_TRY(hipGetLastError());
      s = 5;
      continue;
    }
    if (run_stage(D, s, n, nmax)) return 1;
  }
  *split = sor_batchable(D->d);
  if (!*split) return run_stage(D, 6, n, nmax);
  launch_press_part(D->d, D->dt, 0, kc, ty, tw);
  HIP_TRY(hipGetLastError());
  return 0;
}

int run_stage(les_domain* D, int s, int n, int nmax) {
  (void)nmax;
  if (flush_metrics(D)) return 1;
  switch (s) {
    case 0: launch_velnw(D->d, D->ro, D->dt); break;
    case 1: launch_bondv1(D->d, n, D->dt); break;
    case 2:
      if (ensure_cov(D)) return 1;
      launch_velfg(D->d, D->vn);
      break;
    case 3: launch_feedbf(D->d, D->alpha, D->dt, D->beta); break;
    case 4: launch_les(D->d); break;
    case 5: launch_adam(D->d); break;
    case 6:
      launch_press(D->d, D->dt);
      // press mode 0 leaves the split rhs equal to rhs on the interior (two
      // iterations per launch write only the split rhs); mode 1 overwrites
      // the split rhs with its defects and writes rhs itself
      D->rhs_lazy = D->d.press_mode == 0;
      break;
    default: return fail("unknown stage");
  }
  HIP_TRY(hipGetLastError());
  return 0;
}

// One aveflow call on the device statistics (allocated, zero, on first use).
int run_aveflow(les_domain* D, int n, int n1, int nmax) {
  for (int f = FAVEU; f <= FUWFX; f++)
    if (ensure_field(D, f)) return 1;
  AveBufs a;
  a.aveu = D->fields[FAVEU];
  a.avev = D->fields[FAVEV];
  a.avew = D->fields[FAVEW];
  a.avep = D->fields[FAVEP];
  a.avel = D->fields[FAVEL];
  a.aveuu = D->fields[FAVEUU];
  a.avevv = D->fields[FAVEVV];
  a.aveww = D->fields[FAVEWW];
  a.avesm = D->fields[FAVESM];
  a.avesmsm = D->fields[FAVESMSM];
  a.uwfx = D->fields[FUWFX];
  launch_aveflow(D->d, a, n, n1, nmax);
  HIP_TRY(hipGetLastError());
  return 0;
}

// ------------------------------------------------------------ autotuner
// The fused SOR iteration's launch shape (tile rows per thread, k levels per
// block) changes only the speed, never the iterates (every tiling gives the
// colour sweeps' values bit for bit).  The tuner times short runs of each
// candidate on the domain's own split arrays, which hold only scratch
// between press calls, keeps the fastest and remembers it per grid shape.
struct TuneKey {
  int im, jm, km, sep, ncu;
  bool operator<(const TuneKey& o) const {
    return std::tie(im, jm, km, sep, ncu) < std::tie(o.im, o.jm, o.km, o.sep, o.ncu);
  }
};
struct TuneResult {
  int rows = 2, tile_rows = 16, tile_width = 64, kchunk = 0;
  double us_best = 0.0, us_default = 0.0;
  // k levels per thread of k_velfg_feedbf, k_les_sm and k_les_visc<true> (0:
  // automatic) and the microseconds of one velfg+feedbf and one les+adam
  // launch sequence with them and with the automatic ones
  int velfg_kc = 0, sm_kc = 0, visc_kc = 0;
  double st_us_best = 0.0, st_us_default = 0.0;
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneResult> g_tuned;

// Microseconds per iteration of `reps` back-to-back fused SOR iterations
// (numbered 2.., never stopping) with the given shape, between one pair of
// events on the domain's stream, after two untimed ones.  The caller saves
// and restores the control block.
int time_sor_trial(les_domain* D, const Ctrl& base, int rows, int ty, int tw, int kc, int reps, hipEvent_t e0,
                   hipEvent_t e1, double* us) {
  Ctrl c = base;
  c.sor_done = 0;
  c.nonfinite = 0;
  hipError_t e = hipMemcpyAsync(D->d.ctrl, &c, sizeof(Ctrl), hipMemcpyHostToDevice, D->d.stream);
  if (e == hipSuccess) {
    launch_sor_trial(D->d, rows, ty, tw, kc, 2);  // warm-up
    e = hipEventRecord(e0, D->d.stream);
  }
  if (e == hipSuccess) {
    launch_sor_trial(D->d, rows, ty, tw, kc, reps);
    e = hipEventRecord(e1, D->d.stream);
  }
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (e == hipSuccess) e = hipGetLastError();
  if (e != hipSuccess) return fail(std::string("SOR trial: ") + hipGetErrorString(e));
  *us = 1e3 * ms / reps;
  return 0;
}

int tune_sor(les_domain* D, int reps, TuneResult* out) {
  const Grid& g = D->d.g;
  HIP_TRY(hipStreamSynchronize(D->d.stream));
  Ctrl saved;
  HIP_TRY(copy_sync(&saved, D->d.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, D->d.stream));
  const int ncu = D->d.ncu > 0 ? D->d.ncu : 256;
  const bool pair_ok = (g.jm & 1) == 0 && D->d.cn1_separable;
  // candidate k chunks of a kernel with `tiles` tiles: its automatic chunk
  // and every distinct split of km into nz chunks with between a quarter
  // and eight times as many blocks as compute units
  auto chunks = [&](int tiles, int kdef) {
    std::vector<int> kcs{kdef};
    for (int nz = 1; nz <= g.km; nz++) {
      const int kc = (g.km + nz - 1) / nz;
      const long blocks = (long)tiles * ((g.km + kc - 1) / kc);
      if (blocks * 4 < ncu || blocks > 8L * ncu) continue;
      if (std::find(kcs.begin(), kcs.end(), kc) == kcs.end()) kcs.push_back(kc);
    }
    return kcs;
  };
  const int kdef1 = sor_fused_kchunk(g, D->d.sg, D->d.sor_fused_blocks, D->d.ncu);
  const int kdefp = sor_pair_kchunk(g, D->d.sg, D->d.ncu);
  const std::vector<int> kcs1 = chunks(sor_fused_tiles(g, D->d.sg), kdef1);
  const std::vector<int> kcsp = chunks(sor_pair_tiles(g, D->d.sg), kdefp);
  // (kernel, tile rows, tile width): k_sor_pair (4; two iterations per
  // launch, 32 x 64 tiles), k_sor_iter2 (2) with 16 x 64 or 32 x 64 tiles
  // (even jm only), k_sor_iter (1).  The 32 x 32 tile of k_sor_iter2
  // (LES_SOR_TILE_WIDTH=32) is not a candidate: at 126 x 8 x 6 its stop-test
  // sums differ from the colour sweeps' though its iterates are the same.
  std::vector<std::tuple<int, int, int>> kernels{{1, 16, 64}};
  if ((g.jm & 1) == 0) kernels.insert(kernels.begin(), {{2, 16, 64}, {2, 32, 64}});
  if (pair_ok) kernels.insert(kernels.begin(), {4, 32, 64});
  // the shape used without tuning
  const int rdef = (D->d.sor_rows == 4 && !pair_ok) ? 2 : D->d.sor_rows;
  const std::tuple<int, int, int> kdef2 =
      (rdef == 4 && pair_ok) ? std::make_tuple(4, 32, 64)
      : ((g.jm & 1) == 0 && rdef >= 2)
          ? std::make_tuple(2, D->d.sor_tile_rows, D->d.sor_tile_rows == 32 ? D->d.sor_tile_width : 64)
          : std::make_tuple(1, 16, 64);
  const int kdef = std::get<0>(kdef2) == 4 ? kdefp : kdef1;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIP_TRY(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    hipEventDestroy(e0);
    return fail("hipEventCreate");
  }
  int rc = 0;
  double best = -1.0;
  TuneResult r;
  r.rows = D->d.sor_rows;
  r.tile_rows = D->d.sor_tile_rows;
  r.tile_width = D->d.sor_tile_width;
  r.kchunk = 0;
  for (const auto& kt : kernels) {
    const int rows = std::get<0>(kt), ty = std::get<1>(kt), tw = std::get<2>(kt);
    for (int kc : (rows == 4 ? kcsp : kcs1)) {
      double us = 0.0;
      if ((rc = time_sor_trial(D, saved, rows, ty, tw, kc, reps, e0, e1, &us)) != 0) break;
      if (kt == kdef2 && kc == kdef) r.us_default = us;
      if (best < 0.0 || us < best) {
        best = us;
        r.rows = rows;
        r.tile_rows = ty;
        r.tile_width = tw;
        r.kchunk = kc;
      }
    }
    if (rc) break;
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  // the last press call's statistics, as they were
  if (copy_sync(D->d.ctrl, &saved, sizeof(Ctrl), hipMemcpyHostToDevice, D->d.stream) != hipSuccess && !rc)
    rc = fail("autotune: restoring the control block");
  if (rc) return rc;
  r.us_best = best;
  *out = r;
  return 0;
}

// The stencil kernels' k levels per thread -- k_velfg_feedbf, k_les_sm and
// k_les_visc<true> -- change only the speed: a thread computes each of its
// cells with the same expressions whatever its run, and no partial sum
// depends on it.  Each candidate is timed (`reps` launches between one pair
// of events, after one untimed launch) on the domain's own fields, which
// are restored from a copy before each candidate and at the end: every
// field those kernels write (velfg+feedbf: diu1..9, f, g, h, usum, vsum,
// wsum, uspd, vspd; les+adam: sm, f, g, h, fold, gold, hold).  Velfg+feedbf
// is timed alone, les_sm and les_visc by the les+adam sequence (with
// boundsm), one chunk varied at a time.
int tune_stencils(les_domain* D, int reps, TuneResult* r) {
  const Grid& g = D->d.g;
  if (!D->d.fuse_velfg || !D->d.fuse_les) return 0;  // the unfused stages keep their shapes
  if (flush_metrics(D)) return 1;
  static const int kSaved[] = {FF,    FG,    FH,    FUSUM, FVSUM, FWSUM, FFOLD, FGOLD, FHOLD, FSM,  FUSPD,
                               FVSPD, FDIU1, FDIU2, FDIU3, FDIU4, FDIU5, FDIU6, FDIU7, FDIU8, FDIU9};
  constexpr int kNs = (int)(sizeof(kSaved) / sizeof(kSaved[0]));
  size_t off[kNs + 1];
  off[0] = 0;
  for (int q = 0; q < kNs; q++)
    off[q + 1] = off[q] + (kFields[kSaved[q]].kind == K2D ? (size_t)D->d.g.sk : (size_t)D->n3);
  float* save = nullptr;
  hipError_t e = dev_alloc(&save, off[kNs] * sizeof(float));
  if (e != hipSuccess) return fail(std::string("autotune: field copy allocation: ") + hipGetErrorString(e));
  auto copy_all = [&](bool back) {
    for (int q = 0; q < kNs; q++) {
      float* f = D->fields[kSaved[q]];
      const size_t nb = (off[q + 1] - off[q]) * sizeof(float);
      const hipError_t ce = back ? hipMemcpyAsync(f, save + off[q], nb, hipMemcpyDeviceToDevice, D->d.stream)
                                 : hipMemcpyAsync(save + off[q], f, nb, hipMemcpyDeviceToDevice, D->d.stream);
      if (ce != hipSuccess) return ce;
    }
    return hipSuccess;
  };
  int rc = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if ((e = copy_all(false)) != hipSuccess) rc = fail(std::string("autotune: field copy: ") + hipGetErrorString(e));
  if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = fail("hipEventCreate");
  // candidates: the automatic chunk (0) and every distinct chunk whose
  // grid has between one and 64 blocks per compute unit
  const int gx = (g.im + 63) / 64, gy = (g.jm + 3) / 4;
  const long ncu = D->d.ncu > 0 ? D->d.ncu : 256;
  std::vector<int> kcs{0};
  for (int nz = 1; nz <= g.km; nz++) {
    const int kc = (g.km + nz - 1) / nz;
    const long blocks = (long)gx * gy * ((g.km + kc - 1) / kc);
    if (blocks < ncu || blocks > 64 * ncu) continue;
    if (std::find(kcs.begin(), kcs.end(), kc) == kcs.end()) kcs.push_back(kc);
  }
  // which: 0 velfg+feedbf, 1 les+adam
  auto time_one = [&](int which, int vkc, int skc, int ckc, double* us) -> int {
    DevDomain d = D->d;
    d.timer = nullptr;
    d.velfg_kchunk = vkc;
    d.sm_kchunk = skc;
    d.visc_kchunk = ckc;
    hipError_t te = copy_all(true);
    for (int q = -1; q < reps && te == hipSuccess; q++) {
      if (q == 0) te = hipEventRecord(e0, d.stream);
      if (which == 0)
        launch_velfg_feedbf(d, D->vn, D->alpha, D->dt, D->beta);
      else
        launch_les_adam(d);
      if (te == hipSuccess) te = hipGetLastError();
    }
    if (te == hipSuccess) te = hipEventRecord(e1, d.stream);
    if (te == hipSuccess) te = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (te == hipSuccess) te = hipEventElapsedTime(&ms, e0, e1);
    if (te != hipSuccess) return fail(std::string("autotune: stencil timing: ") + hipGetErrorString(te));
    *us = 1e3 * (double)ms / reps;
    return 0;
  };
  double vdef = 0.0, vbest = -1.0, ldef = 0.0, lbest = -1.0;
  int vkc = 0, skc = 0, ckc = 0;
  for (int kc : kcs) {
    double us = 0.0;
    if (rc || (rc = time_one(0, kc, 0, 0, &us))) break;
    if (kc == 0) vdef = us;
    if (vbest < 0.0 || us < vbest) {
      vbest = us;
      vkc = kc;
    }
  }
  for (int kc : kcs) {  // les_sm's chunk, les_visc's automatic
    double us = 0.0;
    if (D->d.fuse_sm && kc != 0) continue;  // sm is computed in velfg+feedbf
    if (rc || (rc = time_one(1, vkc, kc, 0, &us))) break;
    if (kc == 0) ldef = us;
    if (lbest < 0.0 || us < lbest) {
      lbest = us;
      skc = kc;
    }
  }
  for (int kc : kcs) {  // then les_visc's
    double us = 0.0;
    if (kc == 0) continue;  // timed above with les_sm's best
    if (rc || (rc = time_one(1, vkc, skc, kc, &us))) break;
    if (us < lbest) {
      lbest = us;
      ckc = kc;
    }
  }
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  if ((e = copy_all(true)) != hipSuccess && !rc) rc = fail(std::string("autotune: field restore: ") + hipGetErrorString(e));
  if (hipStreamSynchronize(D->d.stream) != hipSuccess && !rc) rc = fail("autotune: synchronize");
  dev_release(save);
  if (rc) return rc;
  r->velfg_kc = vkc;
  r->sm_kc = skc;
  r->visc_kc = ckc;
  r->st_us_best = vbest + lbest;
  r->st_us_default = vdef + ldef;
  return 0;
}

// -------------------------------------------- reference-signature layer
std::mutex g_mu;
std::map<std::tuple<int, int, int>, les_domain*> g_domains;
struct InflowPlanes {
  std::vector<float> u, v, w;  // v, w empty: zero
};
std::map<std::tuple<int, int, int>, InflowPlanes> g_inflows;

const float* opt(const std::vector<float>& a) { return a.empty() ? nullptr : a.data(); }

int ref_domain(int im, int jm, int km, les_domain** out) {
  auto key = std::make_tuple(im, jm, km);
  auto it = g_domains.find(key);
  if (it != g_domains.end()) {
    *out = it->second;
    return 0;
  }
  les_domain* D = nullptr;
  if (domain_alloc(im, jm, km, &D)) return 1;
  auto pit = g_inflows.find(key);
  if (pit != g_inflows.end() && set_inflow(D, opt(pit->second.u), opt(pit->second.v), opt(pit->second.w))) {
    domain_free(D);
    return 1;
  }
  g_domains[key] = D;
  *out = D;
  return 0;
}

// Mirror a host array: upload when this field last mirrored another address.
int bind(les_domain* D, int f, const float* host) {
  if (!host) return fail(std::string("null array for ") + kFields[f].name);
  if (D->bound[f] == host) return 0;
  if (upload_field(D, f, host)) return 1;
  D->bound[f] = host;
  return 0;
}

#define BIND(f, ptr) \
  if (bind(D, f, ptr)) return 1

#define REF_DOMAIN(im, jm, km)                 \
  std::lock_guard<std::mutex> _lk(g_mu);       \
  les_domain* D = nullptr;                     \
  if (ref_domain(im, jm, km, &D)) return 1

}  // namespace

// Handles handed out by les_domain_create and les_batch_create and not yet
// destroyed.  Every domain and batch entry point checks its handle here, so
// a call on a destroyed domain (or a batch step after one of its domains
// was destroyed) is refused with an error instead of touching freed memory.
std::mutex g_live_mu;
std::set<const void*> g_live;
void live_add(const void* p) {
  std::lock_guard<std::mutex> lk(g_live_mu);
  g_live.insert(p);
}
bool live_remove(const void* p) {
  std::lock_guard<std::mutex> lk(g_live_mu);
  return g_live.erase(p) == 1;
}
bool live_has(const void* p) {
  if (!p) return false;
  std::lock_guard<std::mutex> lk(g_live_mu);
  return g_live.count(p) == 1;
}

namespace les {
// For the entry points defined in other files (k_noalloc.hip).
int report_error(const std::string& msg) { return fail(msg); }
int press_config_values(int* nmaxp, float* pjuge, float* omega) {
  *nmaxp = nmaxp_default;
  *pjuge = pjuge_default;
  *omega = omega_default;
  return 0;
}
}  // namespace les

extern "C" {

const char* les_last_error(void) { return g_err.c_str(); }

int les_velnw(int km, int jm, int im, const float* p, float ro, const float* dxs, float* u, float dt,
              const float* f, const float* dys, float* v, const float* g, const float* dzs, float* w,
              const float* h) {
  REF_DOMAIN(im, jm, km);
  set_metric(D, MDXS, dxs);
  set_metric(D, MDYS, dys);
  set_metric(D, MDZS, dzs);
  BIND(FP, p);
  BIND(FU, u);
  BIND(FF, f);
  BIND(FV, v);
  BIND(FG, g);
  BIND(FW, w);
  BIND(FH, h);
  D->ro = ro;
  D->dt = dt;
  return run_stage(D, 0, 0, 0);
}

int les_bondv1(int jm, float* u, const float* z2, const float* dzn, float* v, float* w, int km, int n,
               int im, float dt, const float* dxs) {
  (void)z2;
  REF_DOMAIN(im, jm, km);
  set_metric(D, MDZN, dzn);
  set_metric(D, MDXS, dxs);
  BIND(FU, u);
  BIND(FV, v);
  BIND(FW, w);
  D->dt = dt;
  return run_stage(D, 1, n, 0);
}

int les_velfg(int km, int jm, int im, const float* dx1, float* cov1, float* cov2, float* cov3,
              float* dfu1, float* diu1, float* diu2, const float* dy1, float* diu3, const float* dzn,
              float vn, float* f, float* cov4, float* cov5, float* cov6, float* dfv1, float* diu4,
              float* diu5, float* diu6, float* g, float* cov7, float* cov8, float* cov9, float* dfw1,
              float* diu7, float* diu8, float* diu9, const float* dzs, float* h, float* nou1,
              const float* u, float* nou5, const float* v, float* nou9, const float* w, float* nou2,
              float* nou3, float* nou4, float* nou6, float* nou7, float* nou8, float* uspd,
              float* vspd) {
  (void)cov1; (void)cov2; (void)cov3; (void)cov4; (void)cov5; (void)cov6; (void)cov7; (void)cov8;
  (void)cov9; (void)dfu1; (void)dfv1; (void)dfw1; (void)nou1; (void)nou2; (void)nou3; (void)nou4;
  (void)nou5; (void)nou6; (void)nou7; (void)nou8; (void)nou9;
  REF_DOMAIN(im, jm, km);
  set_metric(D, MDX1, dx1);
  set_metric(D, MDY1, dy1);
  set_metric(D, MDZN, dzn);
  set_metric(D, MDZS, dzs);
  float* diu[9] = {diu1, diu2, diu3, diu4, diu5, diu6, diu7, diu8, diu9};
  for (int q = 0; q < 9; q++) BIND(FDIU1 + q, diu[q]);
  BIND(FF, f);
  BIND(FG, g);
  BIND(FH, h);
  BIND(FU, u);
  BIND(FV, v);
  BIND(FW, w);
  BIND(FUSPD, uspd);
  BIND(FVSPD, vspd);
  D->vn = vn;
  return run_stage(D, 2, 0, 0);
}

int les_feedbf(int km, int jm, int im, float* usum, const float* u, const float* bmask1, float* vsum,
               const float* v, const float* cmask1, float* wsum, const float* w, const float* dmask1,
               float alpha, float dt, float beta, float* fx, float* fy, float* fz, float* f, float* g,
               float* h) {
  (void)fx; (void)fy; (void)fz;
  REF_DOMAIN(im, jm, km);
  BIND(FUSUM, usum);
  BIND(FU, u);
  BIND(FBMASK, bmask1);
  BIND(FVSUM, vsum);
  BIND(FV, v);
  BIND(FCMASK, cmask1);
  BIND(FWSUM, wsum);
  BIND(FW, w);
  BIND(FDMASK, dmask1);
  BIND(FF, f);
  BIND(FG, g);
  BIND(FH, h);
  D->alpha = alpha;
  D->dt = dt;
  D->beta = beta;
  return run_stage(D, 3, 0, 0);
}

int les_les(int km, float* delx1, const float* dx1, const float* dy1, const float* dzn, int jm, int im,
            const float* diu1, const float* diu2, const float* diu3, const float* diu4,
            const float* diu5, const float* diu6, const float* diu7, const float* diu8,
            const float* diu9, float* sm, float* f, float* g, float* h, const float* uspd,
            const float* vspd, const float* dxs, const float* dys) {
  REF_DOMAIN(im, jm, km);
  set_metric(D, MDX1, dx1);
  set_metric(D, MDY1, dy1);
  set_metric(D, MDZN, dzn);
  set_metric(D, MDXS, dxs);
  set_metric(D, MDYS, dys);
  const float* diu[9] = {diu1, diu2, diu3, diu4, diu5, diu6, diu7, diu8, diu9};
  for (int q = 0; q < 9; q++) BIND(FDIU1 + q, diu[q]);
  BIND(FSM, sm);
  BIND(FF, f);
  BIND(FG, g);
  BIND(FH, h);
  BIND(FUSPD, uspd);
  BIND(FVSPD, vspd);
  if (flush_metrics(D)) return 1;
  if (delx1)
    for (int k = 1; k <= km; k++) delx1[k - 1] = metric_at(D, MDELX1, k);
  return run_stage(D, 4, 0, 0);
}

int les_adam(int n, int nmax, float* fold, int im, int jm, int km, float* gold, float* hold,
             const float* fghold, float* f, float* g, float* h) {
  (void)fghold;
  REF_DOMAIN(im, jm, km);
  BIND(FFOLD, fold);
  BIND(FGOLD, gold);
  BIND(FHOLD, hold);
  BIND(FF, f);
  BIND(FG, g);
  BIND(FH, h);
  return run_stage(D, 5, n, nmax);
}

int les_press(int km, int jm, int im, float* rhs, const float* u, const float* dx1, const float* v,
              const float* dy1, const float* w, const float* dzn, float* f, float* g, float* h,
              float dt, const float* cn1, const float* cn2l, float* p, const float* cn2s,
              const float* cn3l, const float* cn3s, const float* cn4l, const float* cn4s, int n,
              int nmax, const float* usum, const float* vsum, const float* wsum) {
  (void)usum; (void)vsum; (void)wsum;
  REF_DOMAIN(im, jm, km);
  set_metric(D, MDX1, dx1);
  set_metric(D, MDY1, dy1);
  set_metric(D, MDZN, dzn);
  set_metric(D, MCN2L, cn2l);
  set_metric(D, MCN2S, cn2s);
  set_metric(D, MCN3L, cn3l);
  set_metric(D, MCN3S, cn3s);
  set_metric(D, MCN4L, cn4l);
  set_metric(D, MCN4S, cn4s);
  BIND(FRHS, rhs);
  BIND(FU, u);
  BIND(FV, v);
  BIND(FW, w);
  BIND(FF, f);
  BIND(FG, g);
  BIND(FH, h);
  BIND(FCN1, cn1);
  BIND(FP, p);
  D->dt = dt;
  D->d.nmaxp = nmaxp_default;
  D->d.pjuge = pjuge_default;
  D->d.omega = omega_default;
  return run_stage(D, 6, n, nmax);
}

int les_aveflow(int n, int n1, int km, int jm, int im, float* aveu, float* avev, float* avew, float* avep,
                float* avel, float* aveuu, float* avevv, float* aveww, float* avesm, float* avesmsm,
                float* uwfx, const float* u, const float* v, const float* w, const float* p,
                const float* sm, int nmax) {
  REF_DOMAIN(im, jm, km);
  float* st[11] = {aveu, avev, avew, avep, avel, aveuu, avevv, aveww, avesm, avesmsm, uwfx};
  for (int q = 0; q < 11; q++) BIND(FAVEU + q, st[q]);
  BIND(FU, u);
  BIND(FV, v);
  BIND(FW, w);
  BIND(FP, p);
  BIND(FSM, sm);
  if (run_aveflow(D, n, n1, nmax)) return 1;
  if (n == nmax)  // the caller writes data10/data11 from these (aveflow.f95:207-345)
    for (int q = 0; q < 11; q++)
      if (download_field(D, FAVEU + q, st[q])) return 1;
  return 0;
}

int les_bondv1_set_inflow_uvw(int im, int jm, int km, const float* u_in, const float* v_in,
                              const float* w_in) {
  if (!u_in && (v_in || w_in)) return fail("inflow: v and w planes need a u plane");
  std::lock_guard<std::mutex> lk(g_mu);
  auto key = std::make_tuple(im, jm, km);
  const size_t n = (size_t)jm * km;
  auto copy = [n](const float* a) { return a ? std::vector<float>(a, a + n) : std::vector<float>(); };
  if (u_in)
    g_inflows[key] = InflowPlanes{copy(u_in), copy(v_in), copy(w_in)};
  else
    g_inflows.erase(key);
  auto it = g_domains.find(key);
  if (it != g_domains.end()) return set_inflow(it->second, u_in, v_in, w_in);
  return 0;
}

int les_bondv1_set_inflow(int im, int jm, int km, const float* prof) {
  return les_bondv1_set_inflow_uvw(im, jm, km, prof, nullptr, nullptr);
}

int les_press_config_get(int* nmaxp, float* pjuge, float* omega) {
  if (!nmaxp || !pjuge || !omega) return fail("null argument");
  *nmaxp = nmaxp_default;
  *pjuge = pjuge_default;
  *omega = omega_default;
  return 0;
}

int les_press_config(int nmaxp, float pjuge, float omega) {
  if (nmaxp < 1 || nmaxp > kSorHist) return fail("nmaxp out of range (1..1024)");
  nmaxp_default = nmaxp;
  pjuge_default = pjuge;
  omega_default = omega;
  return 0;
}

static int read_stats(les_domain* D, int* iters, double* sor_hist, int hist_len, double* rhsav,
                      double* pav, int* nonfinite) {
  Ctrl c;
  HIP_TRY(hipStreamSynchronize(D->d.stream));
  HIP_TRY(copy_sync(&c, D->d.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, D->d.stream));
  if (iters) *iters = c.sor_iters;
  for (int l = 0; sor_hist && l < hist_len && l < c.sor_iters && l < kSorHist; l++) sor_hist[l] = c.sor_hist[l];
  if (rhsav) *rhsav = c.rhsav;
  if (pav) *pav = c.pav;
  if (nonfinite) *nonfinite = c.nonfinite;
  return 0;
}

int les_press_stats(int im, int jm, int km, int* iters, double* sor_hist, int hist_len, double* rhsav,
                    double* pav, int* nonfinite) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_domains.find(std::make_tuple(im, jm, km));
  if (it == g_domains.end()) return fail("no domain of these extents");
  return read_stats(it->second, iters, sor_hist, hist_len, rhsav, pav, nonfinite);
}

int les_sync_host(const void* host_array) {
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto& kv : g_domains)
    for (int f = 0; f < NFIELDS; f++)
      if (kv.second->bound[f] == host_array) return download_field(kv.second, f, (float*)host_array);
  g_err = "array was never handed to the library";
  return 2;
}

int les_mark_host_dirty(const void* host_array) {
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto& kv : g_domains)
    for (int f = 0; f < NFIELDS; f++)
      if (kv.second->bound[f] == host_array) kv.second->bound[f] = nullptr;
  return 0;
}

int les_synchronize(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto& kv : g_domains) HIP_TRY(hipStreamSynchronize(kv.second->d.stream));
  return 0;
}

int les_release_all(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto& kv : g_domains) {
    hipStreamSynchronize(kv.second->d.stream);
    domain_free(kv.second);
  }
  g_domains.clear();
  g_inflows.clear();
  pool_trim();
  return 0;
}

// ------------------------------------------------------------ domain API
// Masks from the height map on the device (feedbfm.f95:87-116); the level
// tops z2(k)+0.5*dzn(k) are summed here in FP32 as the reference does.
int build_masks(les_domain* D, const float* zbm, const float* z2, const float* dzn) {
  const Grid& g = D->d.g;
  std::vector<float> top(g.km);
  for (int k = 1; k <= g.km; k++) top[k - 1] = z2[k] + 0.5f * dzn[k + 1];  // z2(0:), dzn(-1:)
  float* buf = nullptr;
  const size_t nz = (size_t)g.im * g.jm;
  HIP_TRY(dev_alloc(&buf, (nz + g.km) * sizeof(float)));
  D->bytes += (long long)((nz + g.km) * sizeof(float));
  hipError_t e = hipMemcpyAsync(buf, zbm, nz * sizeof(float), hipMemcpyHostToDevice, D->d.stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(buf + nz, top.data(), g.km * sizeof(float), hipMemcpyHostToDevice, D->d.stream);
  if (e == hipSuccess) {
    launch_masks(D->d, buf, buf + nz);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(D->d.stream);
  if (env_int("LES_MASK_FIELDS", 0)) {  // the fused feedbf reads the mask fields
    dev_release(buf);
    D->bytes -= (long long)((nz + g.km) * sizeof(float));
  } else {  // kept: k_velfg_feedbf evaluates the masks from it instead of reading them
    D->mask_src = buf;
    D->d.mask_src = buf;
  }
  if (e != hipSuccess) return fail(std::string("mask generation: ") + hipGetErrorString(e));
  return 0;
}

// Masks given as fields whose three arrays are equal on the interior, hold
// only 0 and 1, and in every column (i,j) are 1 up to some level K(i,j) and
// 0 above it (buildings standing on the ground, as feedbfm.f95:87-116 makes
// them from a height map): the domain then keeps K(i,j) and the thresholds
// k - 1/2 in the height map's layout, so that k_velfg_feedbf's comparison
// K(i,j) > k - 1/2, true exactly for k <= K(i,j), gives each field's value
// without reading the fields.  Any other masks (or LES_MASK_FIELDS=1) leave
// the fields to be read.  Returns nonzero only on a device error.
int mask_columns(les_domain* D, const float* bm, const float* cm, const float* dm) {
  if (env_int("LES_MASK_FIELDS", 0)) return 0;
  const Grid& g = D->d.g;
  const size_t nz = (size_t)g.im * g.jm;
  const auto base = [&](Kind k, int& ni, int& nj) {
    int nk;
    ref_extent(D, k, &ni, &nj, &nk);
    const KindInfo& K = kKinds[k];
    return (size_t)(-K.li) + (size_t)ni * ((size_t)(-K.lj) + (size_t)nj * (size_t)(-K.lk));
  };
  int bi, bj, ci, cj, di, dj;
  const size_t b0 = base(KB, bi, bj), c0 = base(KU, ci, cj), d0 = base(KR, di, dj);
  const unsigned one = 0x3f800000u, zero = 0u;
  std::vector<float> col(nz + g.km, 0.0f);  // K(i,j), then the thresholds
  std::vector<unsigned char> top_seen(nz, 0);  // a 0 was met below in the column
  for (int k = 1; k <= g.km; k++)
    for (int j = 1; j <= g.jm; j++)
      for (int i = 1; i <= g.im; i++) {
        unsigned b, c, d;
        memcpy(&b, bm + b0 + i + (size_t)bi * (j + (size_t)bj * k), 4);
        memcpy(&c, cm + c0 + i + (size_t)ci * (j + (size_t)cj * k), 4);
        memcpy(&d, dm + d0 + i + (size_t)di * (j + (size_t)dj * k), 4);
        const size_t q = (size_t)(i - 1) + (size_t)(j - 1) * g.im;
        if (b != c || b != d) return 0;
        if (b == one) {
          if (top_seen[q]) return 0;
          col[q] = (float)k;
        } else if (b == zero) {
          top_seen[q] = 1;
        } else {
          return 0;
        }
      }
  for (int k = 1; k <= g.km; k++) col[nz + k - 1] = (float)k - 0.5f;
  float* buf = nullptr;
  HIP_TRY(dev_alloc(&buf, col.size() * sizeof(float)));
  D->bytes += (long long)(col.size() * sizeof(float));
  D->mask_src = buf;
  const hipError_t e = copy_sync(buf, col.data(), col.size() * sizeof(float), hipMemcpyHostToDevice, D->d.stream);
  if (e != hipSuccess) return fail(std::string("mask columns: ") + hipGetErrorString(e));
  D->d.mask_src = buf;
  return 0;
}

int les_domain_create(const les_domain_desc* s, les_domain** out) {
  if (!s || !out) return fail("null argument");
  les_domain* D = nullptr;
  if (domain_alloc(s->im, s->jm, s->km, &D)) return 1;
  set_metric(D, MDX1, s->dx1);
  set_metric(D, MDY1, s->dy1);
  set_metric(D, MDZN, s->dzn);
  set_metric(D, MDZS, s->dzs);
  set_metric(D, MDXS, s->dxs);
  set_metric(D, MDYS, s->dys);
  set_metric(D, MCN2L, s->cn2l);
  set_metric(D, MCN2S, s->cn2s);
  set_metric(D, MCN3L, s->cn3l);
  set_metric(D, MCN3S, s->cn3s);
  set_metric(D, MCN4L, s->cn4l);
  set_metric(D, MCN4S, s->cn4s);
  int rc = 0;
  if (s->cn1) {
    rc |= upload_field(D, FCN1, s->cn1);
  } else if (env_int("LES_SOR_CN1_LOAD", 0)) {
    rc |= fail("LES_SOR_CN1_LOAD needs cn1 in the domain description");
  } else {
    derive(D);
    D->d.cn1_separable = 1;  // init.f95:100-107 by definition
    if (!cn_rcp_range_ok(D))
      rc |= fail("cn1 terms out of the exact-reciprocal range [2^-125, 2^125): pass cn1 explicitly");
  }
  const bool gen_masks = !s->bmask1 && !s->cmask1 && !s->dmask1;
  if (gen_masks && !(s->zbm && s->z2)) rc |= fail("masks: give bmask1/cmask1/dmask1, or zbm and z2");
  if (s->bmask1) rc |= upload_field(D, FBMASK, s->bmask1);
  if (s->cmask1) rc |= upload_field(D, FCMASK, s->cmask1);
  if (s->dmask1) rc |= upload_field(D, FDMASK, s->dmask1);
  rc |= set_inflow(D, s->inflow);
  rc |= flush_metrics(D);
  if (!rc && gen_masks) rc |= build_masks(D, s->zbm, s->z2, s->dzn);
  if (!rc && s->bmask1 && s->cmask1 && s->dmask1) rc |= mask_columns(D, s->bmask1, s->cmask1, s->dmask1);
  if (rc) {
    std::string e = g_err;
    domain_free(D);
    return fail(e);
  }
  D->ro = s->ro;
  D->vn = s->vn;
  D->alpha = s->alpha;
  D->beta = s->beta;
  D->dt = s->dt;
  if (s->nmaxp < 1 || s->nmaxp > kSorHist) {
    domain_free(D);
    return fail("nmaxp out of range (1..1024)");
  }
  D->d.nmaxp = s->nmaxp;
  D->d.pjuge = s->pjuge;
  D->d.omega = s->omega;
  live_add(D);
  *out = D;
  return 0;
}

int les_domain_set_press_mode(les_domain* D, int mode, int outer, int inner) {
  if (!live_has(D)) return fail("null, unknown or destroyed domain");
  if (mode == 0) {
    D->d.press_mode = 0;
    return 0;
  }
  if (mode != 1) return fail("press mode must be 0 (reference SOR) or 1 (FP64 defect correction)");
  if (outer < 1 || outer > kDcMaxOuter) return fail("defect-correction steps out of range (1..64)");
  if (inner < 1 || inner > kSorHist) return fail("inner SOR iterations out of range (1..1024)");
  if (!D->d.dc64) {
    const size_t nb = (2 * (size_t)D->n3 + kDcMaxOuter) * sizeof(double);
    hipError_t e = dev_alloc(&D->d.dc64, nb);
    if (e == hipSuccess) e = hipMemsetAsync(D->d.dc64, 0, nb, D->d.stream);
    if (e != hipSuccess) {
      dev_release(D->d.dc64);
      D->d.dc64 = nullptr;
      return fail(std::string("defect-correction buffers: ") + hipGetErrorString(e));
    }
    D->bytes += (long long)nb;
  }
  D->d.press_mode = 1;
  D->d.dc_outer = outer;
  D->d.dc_inner = inner;
  return 0;
}

int les_domain_set_press_omega(les_domain* D, float omega) {
  if (!live_has(D)) return fail("null, unknown or destroyed domain");
  if (!(omega > 0.0f && omega < 2.0f)) return fail("inner relaxation must lie in (0, 2)");
  D->d.dc_omega = omega;
  return 0;
}

int les_selftest_reciprocal(unsigned lo_bits, unsigned hi_bits, unsigned long long* misses,
                            unsigned* first_bits) {
  if (!misses || !first_bits) return fail("null argument");
  if (lo_bits >= hi_bits) return fail("empty bit range");
  if (selftest_reciprocal(lo_bits, hi_bits, misses, first_bits)) return fail("reciprocal self-test failed to run");
  return 0;
}

int les_domain_query(les_domain* D) {
  if (!D) {
    fail("null domain");
    return -1;
  }
  const hipError_t e = hipStreamQuery(D->d.stream);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  fail(std::string("hipStreamQuery: ") + hipGetErrorString(e));
  return -1;
}

long long les_domain_bytes_needed(int im, int jm, int km) {
  if (im < 2 || jm < 2 || km < 2) return -1;
  const long long sj = row_pitch(im), sk = sj * (jm + 4), n3 = sk * (km + 4);
  const long long nt = (im + 1) / 2, cp = (nt + 1 + 31) / 32 * 32, cs = cp * jm * km;
  long long b = 0;
  for (int f = 0; f < kLazyFirst; f++) b += (kFields[f].kind == K2D ? sk : n3) * 4;
  b += (10 * cs + 2 * cp * jm) * 4;  // colour-split SOR arrays and their two scratch planes
  b += ((long long)im * jm + km) * 4;  // the masks' column values and level thresholds
  return b + (1LL << 20);  // metrics, partial sums, control block (bounded by 1 MiB here)
}

int les_domain_destroy(les_domain* D) {
  if (!D) return 0;
  if (!live_remove(D)) return fail("les_domain_destroy: unknown or already destroyed domain");
  hipStreamSynchronize(D->d.stream);
  domain_free(D);
  return 0;
}

int les_domain_stage(les_domain* D, int stage, int n, int nmax) {
  if (!live_has(D)) return fail("null, unknown or destroyed domain");
  return run_stage(D, stage, n, nmax);
}

// One step from a graph captured on the domain's stream.  Every launch
// argument of a step n > 1 is fixed by the domain state (pointers, sizes,
// physical constants, SOR controls), so one capture serves every such step;
// a change of that state triggers a new capture.  Step 1 differs (initial
// copy in bondv1) and runs eagerly.
static int step_graph(les_domain* D, int n, int nmax) {
  if (flush_metrics(D)) return 1;
  const float sc[5] = {D->ro, D->vn, D->alpha, D->beta, D->dt};
  if (D->gexec && (memcmp(&D->gkey, &D->d, sizeof(DevDomain)) != 0 || memcmp(D->gscal, sc, sizeof(sc)) != 0)) {
    HIP_TRY(hipStreamSynchronize(D->d.stream));  // the previous replay may still be running
    hipGraphExecDestroy(D->gexec);
    D->gexec = nullptr;
  }
  if (!D->gexec) {
    // allocations cannot be captured: make the stage scratch exist first
    if (!D->d.fuse_velfg && ensure_cov(D)) return 1;
    std::lock_guard<std::recursive_mutex> ck(g_capture_mu);
    HIP_TRY(hipStreamBeginCapture(D->d.stream, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    rc = run_step(D, n, nmax);
    hipGraph_t gr = nullptr;
    hipError_t e = hipStreamEndCapture(D->d.stream, &gr);
    if (rc) {
      if (gr) hipGraphDestroy(gr);
      return 1;
    }
    if (e != hipSuccess) return fail(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&D->gexec, gr, nullptr, nullptr, 0);
    hipGraphDestroy(gr);
    if (e != hipSuccess) {
      D->gexec = nullptr;
      return fail(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    memcpy(&D->gkey, &D->d, sizeof(DevDomain));
    memcpy(D->gscal, sc, sizeof(sc));
  }
  HIP_TRY(hipGraphLaunch(D->gexec, D->d.stream));
  D->rhs_lazy = D->d.press_mode == 0;
  return 0;
}

int les_domain_step(les_domain* D, int n, int nmax) {
  if (!live_has(D)) return fail("null, unknown or destroyed domain");
  if (D->use_graph && n != 1 && !D->d.timer) return step_graph(D, n, nmax);
  return run_step(D, n, nmax);
}

int les_domain_aveflow(les_domain* D, int n, int n1, int nmax) {
  if (!live_has(D)) return fail("null, unknown or destroyed domain");
  return run_aveflow(D, n, n1, nmax);
}

int les_domain_set_inflow(les_domain* D, const float* u_in, const float* v_in, const float* w_in) {
  if (!live_has(D)) return fail("null, unknown or destroyed domain");
  return set_inflow(D, u_in, v_in, w_in);
}

int les_domain_autotune(les_domain* D, int reps, int* rows, int* tile_rows, int* tile_width, int* kchunk,
                        double* us_best, double* us_default) {
  if (!live_has(D)) return fail("null, unknown or destroyed domain");
  std::lock_guard<std::mutex> lk(g_tune_mu);
  const Grid& g = D->d.g;
  const TuneKey key{g.im, g.jm, g.km, D->d.cn1_separable, D->d.ncu};
  auto it = g_tuned.find(key);
  if (it == g_tuned.end() || reps > 0) {
    TuneResult t;
    if (tune_sor(D, reps > 0 ? reps : 10, &t)) return 1;
    if (tune_stencils(D, reps > 0 ? reps : 10, &t)) return 1;
    it = g_tuned.insert_or_assign(key, t).first;
  }
  const TuneResult& r = it->second;
  D->d.sor_rows = r.rows;
  D->d.sor_tile_rows = r.tile_rows;
  D->d.sor_tile_width = r.tile_width;
  D->d.sor_kchunk = r.kchunk;
  D->d.velfg_kchunk = r.velfg_kc;
  D->d.sm_kchunk = r.sm_kc;
  D->d.visc_kchunk = r.visc_kc;
  D->st_us_best = r.st_us_best;
  D->st_us_default = r.st_us_default;
  if (rows) *rows = r.rows;
  if (tile_rows) *tile_rows = r.rows == 2 ? r.tile_rows : (r.rows == 4 ? 32 : 16);
  if (tile_width) *tile_width = r.rows == 2 ? r.tile_width : 64;
  if (kchunk) *kchunk = r.kchunk;
  if (us_best) *us_best = r.us_best;
  if (us_default) *us_default = r.us_default;
  return 0;
}

int les_domain_stencil_chunks(les_domain* D, int* velfg_kchunk, int* sm_kchunk, int* visc_kchunk, double* us_best,
                              double* us_default) {
  if (!live_has(D)) return fail("null, unknown or destroyed domain");
  if (velfg_kchunk) *velfg_kchunk = D->d.velfg_kchunk;
  if (sm_kchunk) *sm_kchunk = D->d.sm_kchunk;
  if (visc_kchunk) *visc_kchunk = D->d.visc_kchunk;
  if (us_best) *us_best = D->st_us_best;
  if (us_default) *us_default = D->st_us_default;
  return 0;
}

int les_domain_time_sor(les_domain* D, int iters, double* us_per_iter) {
  if (!D || !us_per_iter) return fail("null argument");
  if (iters < 1) return fail("les_domain_time_sor: iters < 1");
  if (!D->d.sor_fused) return fail("les_domain_time_sor: the fused SOR iteration is off");
  HIP_TRY(hipStreamSynchronize(D->d.stream));
  Ctrl saved;
  HIP_TRY(copy_sync(&saved, D->d.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, D->d.stream));
  const bool pair = D->d.sor_rows == 4 && D->d.cn1_separable && (D->d.g.jm & 1) == 0;
  const int kc = D->d.sor_kchunk > 0 ? D->d.sor_kchunk
                 : pair             ? sor_pair_kchunk(D->d.g, D->d.sg, D->d.ncu)
                                    : sor_fused_kchunk(D->d.g, D->d.sg, D->d.sor_fused_blocks, D->d.ncu);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIP_TRY(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    hipEventDestroy(e0);
    return fail("hipEventCreate");
  }
  int rc = time_sor_trial(D, saved, pair ? 4 : std::min(D->d.sor_rows, 2), D->d.sor_tile_rows,
                          D->d.sor_tile_width, kc, iters + (pair ? iters % 2 : 0), e0, e1, us_per_iter);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (copy_sync(D->d.ctrl, &saved, sizeof(Ctrl), hipMemcpyHostToDevice, D->d.stream) != hipSuccess && !rc)
    rc = fail("les_domain_time_sor: restoring the control block");
  return rc;
}

// ------------------------------------------------------------ batches
// Several domains advanced together.  Domains whose whole step can be
// batched run every stage as one launch for all of them and the SOR
// iterations 2..nmaxp as one launch per iteration (k_sor_iter2_batch), on
// the batch's stream; the others step whole on their own streams, joined by
// events.  Steps with every n > 1 replay one graph.

int les_domain_set_graph(les_domain* D, int on) {
  if (!live_has(D)) return fail("null, unknown or destroyed domain");
  D->use_graph = on ? 1 : 0;
  return 0;
}

int les_domain_run(les_domain* D, int n0, int n1, int nmax) {
  for (int n = n0; n <= n1; n++)
    if (les_domain_step(D, n, nmax)) return 1;
  return 0;
}

int les_domain_upload(les_domain* D, const char* field, const float* host) {
  if (!D || !field || !host) return fail("null argument");
  const int f = field_by_name(field);
  if (f < 0) return fail(std::string("unknown field ") + field);
  return upload_field(D, f, host);
}

int les_domain_download(les_domain* D, const char* field, float* host) {
  if (!D || !field || !host) return fail("null argument");
  const int f = field_by_name(field);
  if (f < 0) return fail(std::string("unknown field ") + field);
  return download_field(D, f, host);
}

int les_domain_synchronize(les_domain* D) {
  if (!live_has(D)) return fail("null, unknown or destroyed domain");
  HIP_TRY(hipStreamSynchronize(D->d.stream));
  return 0;
}

int les_domain_press_stats(les_domain* D, int* iters, double* sor_hist, int hist_len, double* rhsav,
                           double* pav, int* nonfinite) {
  if (!live_has(D)) return fail("null, unknown or destroyed domain");
  return read_stats(D, iters, sor_hist, hist_len, rhsav, pav, nonfinite);
}

long long les_domain_bytes(const les_domain* D) { return live_has(D) ? D->bytes : 0; }

int les_domain_set_timing(les_domain* D, int on) {
  if (!live_has(D)) return fail("null, unknown or destroyed domain");
  HIP_TRY(hipStreamSynchronize(D->d.stream));
  if (on && !D->d.timer) D->d.timer = new Timer();
  if (on) ((Timer*)D->d.timer)->mask = (unsigned int)on;
  if (!on) timer_free(D);
  return 0;
}

int les_domain_timing(les_domain* D, double* ms, long long* count) {
  if (!live_has(D)) return fail("null, unknown or destroyed domain");
  for (int c = 0; c < KC_N; c++) {
    ms[c] = 0.0;
    count[c] = 0;
  }
  Timer* T = (Timer*)D->d.timer;
  if (!T) return 0;
  HIP_TRY(hipStreamSynchronize(D->d.stream));
  for (size_t q = 0; q < T->used; q++) {
    float e = 0.0f;
    HIP_TRY(hipEventElapsedTime(&e, T->ev[2 * q], T->ev[2 * q + 1]));
    ms[T->cls[q]] += e;
    count[T->cls[q]] += T->n[q];
  }
  T->used = 0;
  T->chain = false;
  return 0;
}

int les_domain_set_timing_runs(les_domain* D, int on) {
  if (!live_has(D)) return fail("null, unknown or destroyed domain");
  Timer* T = (Timer*)D->d.timer;
  if (!T) return fail("timing is off");
  HIP_TRY(hipStreamSynchronize(D->d.stream));
  T->runs = on != 0;
  T->chain = false;
  return 0;
}

int les_kernel_classes(void) { return KC_N; }

const char* les_kernel_class_name(int c) { return (c >= 0 && c < KC_N) ? kClassNames[c] : ""; }

}  // extern "C"

struct les_batch {
  std::vector<les_domain*> doms;
  hipStream_t stream = nullptr;
  hipEvent_t fork = nullptr, sor_done = nullptr;
  std::vector<hipEvent_t> head_done, tail_done;
  SorBatchEntry* dtab = nullptr;  // device copy of the SOR table, then its block0 values
  const int* dblock0 = nullptr;
  int ntab = 0, nblocks = 0, kcmax = 0, lmax = 0, remap = 1;
  int sor_kc = 0, sor_ty = 16, sor_tw = 64;  // SOR k chunk, tile rows and width of the batched domains
  bool pairs = false;             // batched SOR iterations 2.. two per launch (k_sor_pair_batch)
  bool om1 = false;               // every batched domain has omega == 1
  // Domains whose whole step runs as batched stage launches (fused velfg and
  // les, batchable SOR); the others step on their own streams.
  std::vector<char> full;
  void* stage_mem = nullptr;      // device: BatchDom[nfull], then StageSpan[BS_N][nfull]
  StageBatch stages{};            // stages.nd == 0: no batched stages
  std::vector<DevDomain> key;     // domain states the table and graph were made for
  std::vector<float> skey;        // and their scalars
  hipGraphExec_t gexec = nullptr;
  int use_graph = 1;
};

namespace {

void batch_free(les_batch* B) {
  if (B->stream) hipStreamSynchronize(B->stream);
  for (les_domain* D : B->doms)
    if (D->d.stream) hipStreamSynchronize(D->d.stream);
  if (B->gexec) hipGraphExecDestroy(B->gexec);
  dev_release(B->dtab);
  dev_release(B->stage_mem);
  for (hipEvent_t e : B->head_done) hipEventDestroy(e);
  for (hipEvent_t e : B->tail_done) hipEventDestroy(e);
  if (B->fork) hipEventDestroy(B->fork);
  if (B->sor_done) hipEventDestroy(B->sor_done);
  stream_destroy(B->stream);
  delete B;
}

std::vector<float> batch_scalars(const les_batch* B) {
  std::vector<float> v;
  for (const les_domain* D : B->doms) v.insert(v.end(), {D->ro, D->vn, D->alpha, D->beta, D->dt});
  return v;
}

bool batch_key_same(const les_batch* B) {
  if (B->key.size() != B->doms.size()) return false;
  for (size_t q = 0; q < B->doms.size(); q++)
    if (memcmp(&B->key[q], &B->doms[q]->d, sizeof(DevDomain)) != 0) return false;
  return B->skey == batch_scalars(B);
}

// The SOR table and the stage tables for the domains' current states
// (synchronous uploads).
int batch_prepare(les_batch* B) {
  std::vector<SorBatchEntry> tab;
  std::vector<BatchDom> bdom;
  std::vector<std::array<StageSpan, BS_N>> spans;
  int block = 0;
  B->kcmax = 0;
  B->lmax = 0;
  B->om1 = true;
  B->full.assign(B->doms.size(), 0);
  std::vector<const DevDomain*> fd;
  long tiles = 0;  // (i,j) tiles of the stage kernels over the batched domains
  int ncu = 256;
  for (size_t q = 0; q < B->doms.size(); q++) {
    les_domain* D = B->doms[q];
    if (flush_metrics(D)) return 1;
    if (!sor_batchable(D->d) || !D->d.fuse_velfg || !D->d.fuse_les) continue;
    B->full[q] = 1;
    fd.push_back(&D->d);
    tiles += (long)((D->d.g.im + 63) / 64) * ((D->d.g.jm + 3) / 4);
    if (D->d.ncu > 0) ncu = D->d.ncu;
  }
  // one SOR k chunk for the batch (LES_BATCH_SOR_KCHUNK overrides), and k
  // splits of the copy-like stages giving about 36 blocks per CU in all
  B->sor_ty = env_int("LES_BATCH_SOR_TILE_ROWS", 16) == 32 ? 32 : 16;
  B->sor_tw = (B->sor_ty == 32 && env_int("LES_BATCH_SOR_TILE_WIDTH", 64) == 32) ? 32 : 64;
  // two iterations per launch measured no faster on batches of small domains
  // (24 domains of 96..160 x 96..160 x 64: 8.0 against 8.1 ms per step, and
  // 10.4 against 8.7 before tiles as wide as a domain lost their ring), so
  // the batch keeps one iteration per launch unless LES_BATCH_SOR_PAIRS=1
  B->pairs = env_int("LES_BATCH_SOR_PAIRS", 0) != 0;
  B->sor_kc = fd.empty() ? 0 : env_int("LES_BATCH_SOR_KCHUNK", 0);
  if (!fd.empty() && B->sor_kc <= 0)
    B->sor_kc = B->pairs ? sor_pair_batch_kchunk(fd.data(), (int)fd.size(), ncu)
                         : sor_batch_kchunk(fd.data(), (int)fd.size(), B->sor_ty, B->sor_tw, ncu);
  const int nzb = tiles > 0 ? (int)std::max(1L, 36L * ncu / tiles) : 0;
  for (size_t q = 0; q < B->doms.size(); q++) {
    if (!B->full[q]) continue;
    les_domain* D = B->doms[q];
    SorBatchEntry e;
    sor_batch_entry(D->d, B->sor_kc, B->sor_ty, B->sor_tw, &e, B->pairs);
    e.block0 = block;
    block += (e.nparts + 7) / 8 * 8;
    B->kcmax = std::max(B->kcmax, e.kchunk);
    B->lmax = std::max(B->lmax, e.nmaxp);
    tab.push_back(e);
    B->om1 = B->om1 && D->d.omega == 1.0f;
    BatchDom bd{};
    std::array<StageSpan, BS_N> sp{};
    bd.d = D->d;
    bd.ro = D->ro;
    bd.dt = D->dt;
    bd.vn = D->vn;
    bd.alpha = D->alpha;
    bd.beta = D->beta;
    step_batch_spans(D->d, nzb, &bd, sp.data());
    press_batch_spans(D->d, B->sor_kc, B->sor_ty, B->sor_tw, nzb, &bd, sp.data());
    bdom.push_back(bd);
    spans.push_back(sp);
  }
  B->ntab = (int)tab.size();
  B->nblocks = block;
  if (B->dtab || B->stage_mem) {
    HIP_TRY(hipStreamSynchronize(B->stream));
    dev_release(B->dtab);
    dev_release(B->stage_mem);
    B->dtab = nullptr;
    B->dblock0 = nullptr;
    B->stage_mem = nullptr;
  }
  B->stages = StageBatch{};
  if (!tab.empty()) {
    // the table, then its block0 values packed for the block-to-domain lookup
    const size_t ntb = tab.size() * sizeof(SorBatchEntry);
    std::vector<int> b0s;
    for (const SorBatchEntry& x : tab) b0s.push_back(x.block0);
    HIP_TRY(dev_alloc(&B->dtab, ntb + b0s.size() * sizeof(int)));
    HIP_TRY(copy_sync(B->dtab, tab.data(), ntb, hipMemcpyHostToDevice, B->stream));
    HIP_TRY(copy_sync((char*)B->dtab + ntb, b0s.data(), b0s.size() * sizeof(int), hipMemcpyHostToDevice, B->stream));
    B->dblock0 = (const int*)((const char*)B->dtab + ntb);
    // stage spans by stage, then domain; each domain's range of blocks starts
    // at a multiple of 8
    const int nd = (int)bdom.size();
    std::vector<StageSpan> flat((size_t)BS_N * nd);
    StageBatch& S = B->stages;
    S.fuse_sm = env_int("LES_FUSE_SM", 1) != 0;
    for (int s = 0; s < BS_N; s++) {
      int b0 = 0;
      for (int q = 0; q < nd; q++) {
        StageSpan x = spans[q][s];
        x.b0 = b0;
        b0 += (x.gx * x.gy * x.gz + 7) / 8 * 8;
        flat[(size_t)s * nd + q] = x;
      }
      S.nblocks[s] = b0;
    }
    for (const BatchDom& x : bdom) S.kc1max = std::max(S.kc1max, x.kc1);
    const size_t nbd = bdom.size() * sizeof(BatchDom);
    HIP_TRY(dev_alloc(&B->stage_mem, nbd + flat.size() * sizeof(StageSpan)));
    HIP_TRY(copy_sync(B->stage_mem, bdom.data(), nbd, hipMemcpyHostToDevice, B->stream));
    HIP_TRY(copy_sync((char*)B->stage_mem + nbd, flat.data(), flat.size() * sizeof(StageSpan),
                      hipMemcpyHostToDevice, B->stream));
    S.doms = (const BatchDom*)B->stage_mem;
    S.spans = (const StageSpan*)((const char*)B->stage_mem + nbd);
    S.nd = nd;
    S.remap = B->remap;
  }
  B->key.clear();
  for (les_domain* D : B->doms) B->key.push_back(D->d);
  B->skey = batch_scalars(B);
  return 0;
}

// One step of every domain, enqueued on the streams.  Under capture the
// domains batched by stage never enter the graph's streams: the graph runs
// on the batch stream, and les_batch_step orders it against their streams.
//  * Staged (every full domain at n > 1): the full domains' stages run as one
//    launch per stage on the batch stream, their SOR iterations 2..nmaxp as
//    one launch per iteration; the other domains step whole on their streams.
//  * Otherwise (a first step): every full domain runs its stages and SOR
//    iteration 1 on its own stream, the batched SOR iterations, then its own
//    press tail.
int batch_enqueue(les_batch* B, const int* ns, const int* nmaxs, bool capturing) {
  const size_t nd = B->doms.size();
  bool staged = B->stages.nd > 0;
  for (size_t q = 0; q < nd; q++)
    if (B->full[q] && ns[q] == 1) staged = false;
  HIP_TRY(hipEventRecord(B->fork, B->stream));
  std::vector<char> split(nd, 0);
  for (size_t q = 0; q < nd; q++) {
    les_domain* D = B->doms[q];
    D->rhs_lazy = D->d.press_mode == 0;  // a batched press also leaves the split rhs equal to rhs
    if (staged && B->full[q]) {
      if (flush_metrics(D)) return 1;
      if (!capturing) {  // after the work already on the domain's stream
        HIP_TRY(hipEventRecord(B->head_done[q], D->d.stream));
        HIP_TRY(hipStreamWaitEvent(B->stream, B->head_done[q], 0));
      }
      continue;
    }
    HIP_TRY(hipStreamWaitEvent(D->d.stream, B->fork, 0));
    if (B->full[q]) {
      bool sp = false;
      if (run_step_head(D, ns[q], nmaxs[q], B->sor_kc, B->sor_ty, B->sor_tw, &sp)) return 1;
      split[q] = sp;
    } else if (run_step(D, ns[q], nmaxs[q])) {
      return 1;
    }
    HIP_TRY(hipEventRecord(B->head_done[q], D->d.stream));
    HIP_TRY(hipStreamWaitEvent(B->stream, B->head_done[q], 0));
  }
  if (staged) {
    launch_step_batch(B->stages, B->stream);
    launch_press_batch(B->stages, 0, B->stream);
  }
  if (B->pairs && B->ntab > 0)
    launch_sor_pair_batch(B->dtab, B->dblock0, B->ntab, B->nblocks, B->lmax, B->kcmax, B->om1, B->remap, B->stream);
  for (int l = 2; !B->pairs && l <= B->lmax && B->ntab > 0; l++)
    launch_sor_batch(B->dtab, B->dblock0, B->ntab, B->nblocks, l, B->kcmax, B->sor_ty, B->sor_tw, B->om1, B->remap,
                     B->stream);
  if (staged) launch_press_batch(B->stages, 1, B->stream, B->pairs);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(B->sor_done, B->stream));
  for (size_t q = 0; q < nd; q++) {
    les_domain* D = B->doms[q];
    if (staged && B->full[q]) {
      if (!capturing) HIP_TRY(hipStreamWaitEvent(D->d.stream, B->sor_done, 0));
      continue;
    }
    if (split[q]) {
      HIP_TRY(hipStreamWaitEvent(D->d.stream, B->sor_done, 0));
      launch_press_part(D->d, D->dt, 1, B->sor_kc, B->sor_ty, B->sor_tw, B->pairs);
      HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(B->tail_done[q], D->d.stream));
    HIP_TRY(hipStreamWaitEvent(B->stream, B->tail_done[q], 0));
  }
  return 0;
}

}  // namespace

extern "C" {

int les_batch_create(les_domain* const* doms, int n, les_batch** out) {
  if (!doms || n < 1 || !out) return fail("les_batch_create: need at least one domain");
  for (int q = 0; q < n; q++) {
    if (!live_has(doms[q])) return fail("les_batch_create: null, unknown or destroyed domain");
    for (int r = 0; r < q; r++)
      if (doms[r] == doms[q]) return fail("les_batch_create: a domain appears twice");
    if (doms[q]->d.timer) return fail("les_batch_create: a domain has timing on");
  }
  les_batch* B = new les_batch();
  B->doms.assign(doms, doms + n);
  B->remap = doms[0]->d.xcd_remap;
  hipError_t e = stream_create(&B->stream);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&B->fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&B->sor_done, hipEventDisableTiming);
  for (int q = 0; q < n && e == hipSuccess; q++) {
    hipEvent_t a = nullptr, b = nullptr;
    e = hipEventCreateWithFlags(&a, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b, hipEventDisableTiming);
    if (a) B->head_done.push_back(a);
    if (b) B->tail_done.push_back(b);
  }
  if (e != hipSuccess) {
    batch_free(B);
    return fail(std::string("les_batch_create: ") + hipGetErrorString(e));
  }
  live_add(B);
  *out = B;
  return 0;
}

int les_batch_destroy(les_batch* B) {
  if (!B) return 0;
  if (!live_remove(B)) return fail("les_batch_destroy: unknown or already destroyed batch");
  batch_free(B);
  return 0;
}

int les_batch_set_graph(les_batch* B, int on) {
  if (!live_has(B)) return fail("null, unknown or destroyed batch");
  B->use_graph = on ? 1 : 0;
  return 0;
}

int les_batch_step(les_batch* B, const int* ns, const int* nmaxs) {
  if (!live_has(B) || !ns || !nmaxs) return fail("les_batch_step: null argument, or unknown or destroyed batch");
  for (les_domain* D : B->doms)  // a domain must outlive the batches holding it
    if (!live_has(D)) return fail("les_batch_step: a domain of the batch was destroyed");
  if (!batch_key_same(B)) {
    if (B->gexec) {
      HIP_TRY(hipStreamSynchronize(B->stream));
      HIP_TRY(hipGraphExecDestroy(B->gexec));
      B->gexec = nullptr;
    }
    if (batch_prepare(B)) return 1;
  }
  for (les_domain* D : B->doms) {
    if (flush_metrics(D)) return 1;  // synchronous: not under capture
    D->rhs_lazy = D->d.press_mode == 0;
  }
  bool replay = B->use_graph != 0;
  for (size_t q = 0; q < B->doms.size(); q++) replay = replay && ns[q] != 1;
  if (!replay) return batch_enqueue(B, ns, nmaxs, false);
  if (!B->gexec) {
    for (les_domain* D : B->doms)
      if (!D->d.fuse_velfg && ensure_cov(D)) return 1;  // allocations cannot be captured
    std::lock_guard<std::recursive_mutex> ck(g_capture_mu);
    HIP_TRY(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
    const int rc = batch_enqueue(B, ns, nmaxs, true);
    hipGraph_t gr = nullptr;
    const hipError_t e = hipStreamEndCapture(B->stream, &gr);
    if (rc) {
      if (gr) hipGraphDestroy(gr);
      return 1;
    }
    if (e != hipSuccess) return fail(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    const hipError_t ei = hipGraphInstantiate(&B->gexec, gr, nullptr, nullptr, 0);
    hipGraphDestroy(gr);
    if (ei != hipSuccess) {
      B->gexec = nullptr;
      return fail(std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    }
  }
  // a graph runs on the batch stream only: order it after the domains'
  // streams and their later work after it
  for (size_t q = 0; q < B->doms.size(); q++) {
    HIP_TRY(hipEventRecord(B->head_done[q], B->doms[q]->d.stream));
    HIP_TRY(hipStreamWaitEvent(B->stream, B->head_done[q], 0));
  }
  HIP_TRY(hipGraphLaunch(B->gexec, B->stream));
  HIP_TRY(hipEventRecord(B->sor_done, B->stream));
  for (les_domain* D : B->doms) HIP_TRY(hipStreamWaitEvent(D->d.stream, B->sor_done, 0));
  return 0;
}

int les_batch_query(les_batch* B) {
  if (!live_has(B)) return fail("null, unknown or destroyed batch");
  const hipError_t e = hipStreamQuery(B->stream);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  fail(std::string("hipStreamQuery: ") + hipGetErrorString(e));
  return -1;
}

int les_batch_synchronize(les_batch* B) {
  if (!live_has(B)) return fail("null, unknown or destroyed batch");
  HIP_TRY(hipStreamSynchronize(B->stream));
  return 0;
}

}  // extern "C"
