// libgpe_hip.so: C ABI (include/gpe_hip.h) over the fused gfx950 kernels.
//
// Host logic mirrors init/solve! of the reference:
//   resolve_fixed_timestepping   src/fixed_time_stepping.jl:14-24
//   solve! save/step loops       src/fixed_time_stepping.jl:26-54
//   init (tables, pump buffers)  src/strang_splitting.jl:32-67
//   step! ordering               src/strang_splitting.jl:86-90
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <functional>
#include <map>
#include <mutex>
#include <tuple>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gpe_hip.h"
#include "gpe_launch.hpp"

using namespace gpe;

// ------------------------------------------------------------------ errors
static thread_local std::string g_err;

static int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                               \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess) return fail(GPE_ERR_DEVICE, "%s: %s", #expr, hipGetErrorString(e_));     \
  } while (0)

// ------------------------------------------------------------------ handle
// Scratch of gpe_steady_state: two field snapshots and the reduction buffers,
// allocated on the first call and kept by the handle.
struct SteadyBufs {
  void* a[kMaxM] = {nullptr, nullptr};
  void* b[kMaxM] = {nullptr, nullptr};
  double* partial = nullptr;
  double* result = nullptr;
  double* host = nullptr;  // pinned pair (diff^2, norm^2)
};

constexpr int kPlanMaxCand = 6;  // (rows, columns) pairs a measured plan times
// Device form of a D(k) that couples M > 2 components (exp(-i dt D) row-major per k).
constexpr int kDFormCoupled = 5;
// Largest sampled-pump table per component (GPE_PUMP_SAMPLES): 32 GiB of HBM.
constexpr int64_t kMaxPumpSampleBytes = int64_t(32) << 30;

struct gpe_handle {
  gpe_problem_desc d;
  int M = 1, ndim = 1;
  // n1 is the product of the extents beyond dim 0 (the x-pass's lines per
  // grid); gy, gz are the extents of dims 1 and 2 (1 where absent)
  int64_t n0 = 0, n1 = 1, nsp = 0, batch = 1, npts = 0;
  int64_t gy = 1, gz = 1;
  bool f64 = false;
  size_t csz = 8;
  hipStream_t stream = nullptr;
  void* u[kMaxM] = {nullptr, nullptr};
  void* tw = nullptr;
  int64_t ntw = 0;
  // dispersion
  int disp_form = GPE_FORM_NONE;  // device form: 1 scalar, 2 diag, 3 full, 4 symmetric 2x2, 5 coupled M x M
  int ed_fold = 0;                // bit d: D is even in k_d and the table is folded
  int64_t ed_e0 = 0, ed_e1 = 0;   // compact table extents
  double* dD = nullptr;
  void* ed = nullptr;
  size_t dD_cap = 0, ed_cap = 0;  // bytes held (tables are reused while they fit)
  double ed_dt = NAN;
  // potential
  int pot_form = GPE_FORM_NONE;
  double* dV = nullptr;
  void* ev[kMaxM] = {nullptr, nullptr};
  double ev_dt = NAN;
  // nonlinearity
  int gmask = 0;
  int g_cplx = 0;
  int g_coupled = 0;
  double g[kMaxM][kMaxM][2] = {};
  double c[kMaxM][2] = {};
  // 2x2 matrix nonlinearity G_ab = gm_c[a][b] + sum_j gm_g[a][b][j] |u_j|^2
  bool gmat = false;
  double gm_g[2][2][2][2] = {};
  double gm_c[2][2][2] = {};
  // pump
  int pump_kind = GPE_PUMP_NONE;
  int64_t pump_samples = 0;  // GPE_PUMP_SAMPLES: sample planes per component
  void* pump_a[kMaxM] = {nullptr, nullptr};  // active profile per component (nullptr: F_c = 0)
  void* pump_b[kMaxM] = {nullptr, nullptr};
  void* pump_buf_a[kMaxM] = {nullptr, nullptr};  // owned storage behind pump_a / pump_b
  void* pump_buf_b[kMaxM] = {nullptr, nullptr};
  size_t pump_cap_a[kMaxM] = {0, 0}, pump_cap_b[kMaxM] = {0, 0};
  std::vector<double> env;  // complex samples
  double* d_env = nullptr;
  int64_t d_env_cap = 0;
  // noise
  int noise_kind = GPE_NOISE_NONE;
  double eta[kMaxM] = {0, 0};
  uint64_t seed = 0, call = 0;
  // staging for frames that go to host memory (or feed the observables only)
  void* stage[kMaxM] = {nullptr, nullptr};
  // ensemble observables: caller's outputs and the device accumulation buffer
  double* obs_n = nullptr;
  double* obs_n2 = nullptr;
  double* d_obs = nullptr;
  size_t d_obs_cap = 0;
  // last solve configuration (for gpe_advance / timing)
  bool configured = false;
  double dt = 0;
  float last_ms = 0;
  double event_pair_ms = 0;  // latency of an empty HIP event pair (gpe_profile_kernels)
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  SteadyBufs steady;
  // launch schedule of 2D solves (GPE_SCHEDULE_*) and the one last used
  int schedule = GPE_SCHEDULE_AUTO;
  int last_schedule = GPE_SCHEDULE_LAUNCHES;
  // launch shapes of the 2D passes (rows / columns per workgroup; 0: the
  // built-in rule), how they were chosen, and the last measurement: the
  // (rows, columns) pairs timed and their us per step
  int planner = GPE_PLAN_AUTO;
  int plan_w = 0, plan_c = 0;
  bool plan_fixed = false;
  uint64_t plan_key = 0;  // terms the measured plan was timed with (0: none)
  int plan_ncand = 0;
  int plan_cand[kPlanMaxCand][2] = {};
  double plan_us[kPlanMaxCand] = {};
  // scratch of gpe_kspace_g2 / gpe_windowed_g1 (transformed ensembles, sums)
  void* g2_buf = nullptr;
  size_t g2_cap = 0;
};

static bool is_pow2(int64_t n) { return n > 0 && (n & (n - 1)) == 0; }

static size_t field_bytes(const gpe_handle* h) { return (size_t)h->npts * h->csz; }

static bool is_device_ptr(const void* p) {
  hipPointerAttribute_t attr;
  hipError_t e = hipPointerGetAttributes(&attr, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

static hipError_t run_x_w(gpe_handle* h, const XPassArgs& a, int w) {
  const int64_t nlines = h->npts / h->n0;
  return h->f64 ? gpe_run_x_f64(h->M, h->n0, a, nlines, w, h->stream)
                : gpe_run_x_f32(h->M, h->n0, a, nlines, w, h->stream);
}
static hipError_t run_y_c(gpe_handle* h, const YPassArgs& a, int c) {
  return h->f64 ? gpe_run_y_f64(h->M, h->n1, a, c, h->stream) : gpe_run_y_f32(h->M, h->n1, a, c, h->stream);
}
static hipError_t run_x(gpe_handle* h, const XPassArgs& a) { return run_x_w(h, a, h->plan_w); }
static hipError_t run_y(gpe_handle* h, const YPassArgs& a) { return run_y_c(h, a, h->plan_c); }
static hipError_t run_col(gpe_handle* h, const YPassArgs& a, int64_t n) {
  return h->f64 ? gpe_run_y_f64(h->M, n, a, 0, h->stream) : gpe_run_y_f32(h->M, n, a, 0, h->stream);
}
static hipError_t run_1d(gpe_handle* h, const Step1DArgs& a) {
  return h->f64 ? gpe_run_1d_f64(h->M, h->n0, a, h->stream) : gpe_run_1d_f32(h->M, h->n0, a, h->stream);
}

// --------------------------------------------------------------- helpers
static int upload(void** dst, const void* src, size_t bytes) {
  if (*dst == nullptr) HIPCHK(hipMalloc(dst, bytes));
  HIPCHK(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
  return GPE_OK;
}

static void dfree(void*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

// Keep *p when it already holds `bytes`; reallocate otherwise.  Re-setting a
// term with a table of the same size then allocates and frees nothing.
static int ensure_buf(void** p, size_t* cap, size_t bytes) {
  if (*p && *cap >= bytes) return GPE_OK;
  dfree(*p);
  *cap = 0;
  HIPCHK(hipMalloc(p, bytes));
  *cap = bytes;
  return GPE_OK;
}

static int upload_into(void** p, size_t* cap, const void* src, size_t bytes) {
  int rc = ensure_buf(p, cap, bytes);
  if (rc) return rc;
  HIPCHK(hipMemcpy(*p, src, bytes, hipMemcpyHostToDevice));
  return GPE_OK;
}

static void free_steady(SteadyBufs& sb) {
  for (int c = 0; c < kMaxM; ++c) {
    dfree(sb.a[c]);
    dfree(sb.b[c]);
  }
  void* p = sb.partial;
  dfree(p);
  p = sb.result;
  dfree(p);
  if (sb.host) (void)hipHostFree(sb.host);
  sb = SteadyBufs();
}

// Twiddle table w[m] = exp(-2 pi i m / n), m = 0..n-1, in FP64: the angle is
// reduced exactly to a quadrant and reflected to [0, pi/4] so every value is
// one libm sin or cos of a small argument.
static void twiddles_fp64(int64_t n, std::vector<double>& w) {
  w.resize(2 * n);
  for (int64_t m = 0; m < n; ++m) {
    const int64_t q4 = 4 * m;
    const int64_t quad = q4 / n;         // 0..3
    const int64_t rem = q4 - quad * n;   // angle within quadrant = rem/n * pi/2
    double c, s;
    if (2 * rem <= n) {
      const double phi = (double)rem / (double)n * (M_PI / 2);
      c = std::cos(phi);
      s = std::sin(phi);
    } else {
      const double phi = (double)(n - rem) / (double)n * (M_PI / 2);
      c = std::sin(phi);
      s = std::cos(phi);
    }
    double cq, sq;
    switch (quad) {
      case 0: cq = c; sq = s; break;
      case 1: cq = -s; sq = c; break;
      case 2: cq = -c; sq = -s; break;
      default: cq = s; sq = -c; break;
    }
    w[2 * m] = cq;
    w[2 * m + 1] = -sq;  // exp(-i angle)
  }
}

// ------------------------------------------------------------------- ABI
extern "C" int gpe_version(void) { return GPE_ABI_VERSION; }

extern "C" const char* gpe_last_error(void) { return g_err.c_str(); }

extern "C" int gpe_create(const gpe_problem_desc* desc, gpe_handle** out) {
  if (!desc || !out) return fail(GPE_ERR_INVALID, "gpe_create: null argument");
  *out = nullptr;
  const gpe_problem_desc& d = *desc;
  if (d.ndim < 1 || d.ndim > 3) return fail(GPE_ERR_UNSUPPORTED, "gpe_create: ndim=%d (supported: 1, 2, 3)", d.ndim);
  if (d.ncomp < 1 || d.ncomp > kMaxM)
    return fail(GPE_ERR_UNSUPPORTED, "gpe_create: ncomp=%d (supported: 1..%d)", d.ncomp, kMaxM);
  if (d.dtype != GPE_C64 && d.dtype != GPE_C128) return fail(GPE_ERR_INVALID, "gpe_create: dtype=%d", d.dtype);
  if (d.batch < 1) return fail(GPE_ERR_INVALID, "gpe_create: batch=%lld", (long long)d.batch);
  for (int i = 0; i < d.ndim; ++i) {
    if (!is_pow2(d.n[i]) || d.n[i] < 16 || d.n[i] > 4096)
      return fail(GPE_ERR_UNSUPPORTED, "gpe_create: n[%d]=%lld (supported: powers of two 16..4096)", i, (long long)d.n[i]);
    if (!(d.L[i] > 0) || !std::isfinite(d.L[i])) return fail(GPE_ERR_INVALID, "gpe_create: L[%d]=%g", i, d.L[i]);
  }
  // a pass holds the M components of a line in one workgroup: M * N/8
  // threads and two LDS buffers of M padded lines
  for (int i = 0; i < d.ndim; ++i) {
    const bool f64 = d.dtype == GPE_C128;
    if (d.ncomp * (d.n[i] / kE) > 1024 || pass_lds_bytes((int)d.n[i], d.ncomp, f64) > 160 * 1024)
      return fail(GPE_ERR_UNSUPPORTED, "gpe_create: n[%d]=%lld is too long for %d %s components (%s)", i,
                  (long long)d.n[i], d.ncomp, f64 ? "FP64" : "FP32",
                  d.ncomp * (d.n[i] / kE) > 1024 ? "threads per workgroup" : "LDS");
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(GPE_ERR_DEVICE, "gpe_create: no HIP device");
  if (d.device < 0 || d.device >= ndev) return fail(GPE_ERR_INVALID, "gpe_create: device=%d of %d", d.device, ndev);
  HIPCHK(hipSetDevice(d.device));
  gpe_handle* h = new gpe_handle();
  h->d = d;
  h->M = d.ncomp;
  h->ndim = d.ndim;
  h->n0 = d.n[0];
  h->gy = d.ndim >= 2 ? d.n[1] : 1;
  h->gz = d.ndim == 3 ? d.n[2] : 1;
  h->n1 = h->gy * h->gz;
  h->nsp = h->n0 * h->n1;
  h->batch = d.batch;
  h->npts = h->nsp * h->batch;
  h->f64 = d.dtype == GPE_C128;
  h->csz = h->f64 ? 16 : 8;
  int rc = GPE_OK;
  auto cleanup = [&](int code) {
    gpe_destroy(h);
    return code;
  };
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup(fail(GPE_ERR_DEVICE, "gpe_create: stream"));
  for (int c = 0; c < h->M; ++c)
    if (hipMalloc(&h->u[c], field_bytes(h)) != hipSuccess)
      return cleanup(fail(GPE_ERR_NOMEM, "gpe_create: field allocation of %zu bytes", field_bytes(h)));
  h->ntw = std::max(h->n0, std::max(h->gy, h->gz));
  std::vector<double> w;
  twiddles_fp64(h->ntw, w);
  if (h->f64) {
    rc = upload(&h->tw, w.data(), w.size() * sizeof(double));
  } else {
    std::vector<float> wf(w.begin(), w.end());
    rc = upload(&h->tw, wf.data(), wf.size() * sizeof(float));
  }
  if (rc) return cleanup(rc);
  if (hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess)
    return cleanup(fail(GPE_ERR_DEVICE, "gpe_create: events"));
  *out = h;
  return GPE_OK;
}

extern "C" int gpe_destroy(gpe_handle* h) {
  if (!h) return GPE_OK;
  (void)hipSetDevice(h->d.device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (int c = 0; c < kMaxM; ++c) {
    dfree(h->u[c]);
    dfree(h->ev[c]);
    dfree(h->pump_buf_a[c]);
    dfree(h->pump_buf_b[c]);
    dfree(h->stage[c]);
  }
  dfree(h->tw);
  void* po = h->d_obs;
  dfree(po);
  void* p = h->dD;
  dfree(p);
  p = h->dV;
  dfree(p);
  dfree(h->ed);
  p = h->d_env;
  dfree(p);
  dfree(h->g2_buf);
  free_steady(h->steady);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return GPE_OK;
}

static int per_point(int form, int M) { return form == GPE_FORM_SCALAR ? 1 : (form == GPE_FORM_DIAG ? M : M * M); }

// D(k) is uploaded in a compact form: a 1x1 matrix or 1-vector becomes the
// scalar form, a 2x2 matrix with D01 == D10 everywhere is stored as (a, b, d),
// and when D(k) == D(-k_d) exactly along dimension d only k_d in [0, N_d/2]
// is stored.
extern "C" int gpe_set_dispersion(gpe_handle* h, int32_t form, const double* D) {
  if (!h) return fail(GPE_ERR_INVALID, "gpe_set_dispersion: null handle");
  HIPCHK(hipSetDevice(h->d.device));
  if (form == GPE_FORM_NONE) {
    h->disp_form = GPE_FORM_NONE;
    h->ed_dt = NAN;
    return GPE_OK;
  }
  if (form < GPE_FORM_SCALAR || form > GPE_FORM_FULL) return fail(GPE_ERR_INVALID, "gpe_set_dispersion: form=%d", form);
  if (!D) return fail(GPE_ERR_INVALID, "gpe_set_dispersion: null table");
  const int M = h->M;
  std::vector<double> diag_only;
  bool coupled = false;  // M > 2 with off-diagonal entries: device form 5
  if (M > 2 && form == GPE_FORM_FULL) {
    // a matrix D(k) whose off-diagonal entries are all zero is applied as its
    // diagonal (no exchange between components in the k-space pass)
    diag_only.resize((size_t)2 * M * h->nsp);
    for (int64_t k = 0; k < h->nsp && !coupled; ++k)
      for (int a = 0; a < M; ++a)
        for (int b = 0; b < M; ++b) {
          const double* e = D + 2 * ((size_t)k * M * M + a + (size_t)M * b);
          if (a == b) {
            diag_only[2 * ((size_t)k * M + a)] = e[0];
            diag_only[2 * ((size_t)k * M + a) + 1] = e[1];
          } else if (e[0] != 0.0 || e[1] != 0.0) {
            coupled = true;
          }
        }
    if (!coupled) {
      D = diag_only.data();
      form = GPE_FORM_DIAG;
    }
  }
  const int per_in = per_point(form, M);
  const int64_t n0 = h->n0, n1 = h->n1;
  // device form and values per entry
  int dform = form;
  if (M == 1) dform = GPE_FORM_SCALAR;
  if (coupled) dform = kDFormCoupled;
  if (dform == GPE_FORM_FULL) {
    bool sym = true;
    for (int64_t k = 0; k < h->nsp && sym; ++k) {
      const double* e = D + 2 * (size_t)k * 4;
      sym = e[2] == e[4] && e[3] == e[5];
    }
    if (sym) dform = 4;
  }
  auto entry = [&](int64_t k0, int64_t k1) { return D + 2 * (size_t)(k0 + n0 * k1) * per_in; };
  auto same = [&](const double* x, const double* y) {
    for (int i = 0; i < 2 * per_in; ++i)
      if (x[i] != y[i]) return false;
    return true;
  };
  int fold = 0;
  bool even0 = true, even1 = (h->ndim == 2);
  for (int64_t k1 = 0; k1 < n1 && (even0 || even1); ++k1)
    for (int64_t k0 = 0; k0 < n0 && (even0 || even1); ++k0) {
      if (even0 && !same(entry(k0, k1), entry((n0 - k0) % n0, k1))) even0 = false;
      if (even1 && !same(entry(k0, k1), entry(k0, (n1 - k1) % n1))) even1 = false;
    }
  if (even0) fold |= 1;
  if (even1) fold |= 2;
  const int64_t e0 = even0 ? n0 / 2 + 1 : n0, e1 = even1 ? n1 / 2 + 1 : n1;
  const int per_out = dform == GPE_FORM_SCALAR  ? 1
                      : dform == GPE_FORM_DIAG  ? M
                      : dform == GPE_FORM_FULL  ? 4
                      : dform == kDFormCoupled ? M * M
                                                : 3;
  std::vector<double> compact((size_t)e0 * e1 * per_out * 2);
  for (int64_t k1 = 0; k1 < e1; ++k1)
    for (int64_t k0 = 0; k0 < e0; ++k0) {
      const double* src = entry(k0, k1);
      double* dst = compact.data() + 2 * (size_t)(k0 + e0 * k1) * per_out;
      if (dform == 4) {
        const int pick[3] = {0, 1, 3};
        for (int i = 0; i < 3; ++i) {
          dst[2 * i] = src[2 * pick[i]];
          dst[2 * i + 1] = src[2 * pick[i] + 1];
        }
      } else if (dform == GPE_FORM_SCALAR && per_in != 1) {
        dst[0] = src[0];
        dst[1] = src[1];
      } else {
        for (int i = 0; i < 2 * per_out; ++i) dst[i] = src[i];
      }
    }
  // launches already queued on the stream may still read the current tables
  HIPCHK(hipStreamSynchronize(h->stream));
  h->disp_form = GPE_FORM_NONE;
  h->ed_dt = NAN;
  void* p = h->dD;
  int rc = upload_into(&p, &h->dD_cap, compact.data(), compact.size() * sizeof(double));
  h->dD = (double*)p;
  if (rc) return rc;
  // device records: four values per k for the two-component forms (see load_ed),
  // M x M (row-major) for the coupled form
  const int per_dev = dform == kDFormCoupled ? M * M : (dform >= 3 ? 4 : per_out);
  rc = ensure_buf(&h->ed, &h->ed_cap, (size_t)e0 * e1 * per_dev * h->csz);
  if (rc) return rc;
  h->disp_form = dform;
  h->ed_fold = fold;
  h->ed_e0 = e0;
  h->ed_e1 = e1;
  h->ed_dt = NAN;
  return GPE_OK;
}

extern "C" int gpe_set_potential(gpe_handle* h, int32_t form, const double* V) {
  if (!h) return fail(GPE_ERR_INVALID, "gpe_set_potential: null handle");
  HIPCHK(hipSetDevice(h->d.device));
  HIPCHK(hipStreamSynchronize(h->stream));  // queued launches may still read the tables
  void* p = h->dV;
  dfree(p);
  h->dV = nullptr;
  for (int c = 0; c < kMaxM; ++c) dfree(h->ev[c]);
  h->ev_dt = NAN;
  if (form == GPE_FORM_NONE) {
    h->pot_form = GPE_FORM_NONE;
    return GPE_OK;
  }
  if (form != GPE_FORM_SCALAR && form != GPE_FORM_DIAG)
    return fail(GPE_ERR_UNSUPPORTED, "gpe_set_potential: form=%d (supported: scalar, diagonal)", form);
  if (!V) return fail(GPE_ERR_INVALID, "gpe_set_potential: null table");
  int rc = upload(&p, V, (size_t)h->nsp * per_point(form, h->M) * 2 * sizeof(double));
  h->dV = (double*)p;
  if (rc) return rc;
  for (int c = 0; c < h->M; ++c) HIPCHK(hipMalloc(&h->ev[c], (size_t)h->nsp * h->csz));
  h->pot_form = form;
  return GPE_OK;
}

extern "C" int gpe_set_nonlinearity(gpe_handle* h, const double* g, const double* c) {
  if (!h) return fail(GPE_ERR_INVALID, "gpe_set_nonlinearity: null handle");
  const int M = h->M;
  h->gmask = 0;
  h->g_cplx = 0;
  for (int i = 0; i < M; ++i) {
    bool any = false;
    for (int j = 0; j < M; ++j)
      for (int r = 0; r < 2; ++r) {
        h->g[i][j][r] = g ? g[2 * (i + M * j) + r] : 0.0;
        any = any || h->g[i][j][r] != 0.0;
        if (r == 1 && h->g[i][j][r] != 0.0) h->g_cplx = 1;
      }
    for (int r = 0; r < 2; ++r) {
      h->c[i][r] = c ? c[2 * i + r] : 0.0;
      any = any || h->c[i][r] != 0.0;
      if (r == 1 && h->c[i][r] != 0.0) h->g_cplx = 1;
    }
    if (any) h->gmask |= 1 << i;
  }
  h->g_coupled = 0;
  for (int i = 0; i < M; ++i)
    for (int j = 0; j < M; ++j)
      if (i != j && (h->g[i][j][0] != 0.0 || h->g[i][j][1] != 0.0)) h->g_coupled = 1;
  h->gmat = false;
  return GPE_OK;
}

extern "C" int gpe_set_nonlinearity_matrix(gpe_handle* h, const double* g, const double* c) {
  if (!h || !g || !c) return fail(GPE_ERR_INVALID, "gpe_set_nonlinearity_matrix: null argument");
  if (h->M != 2) return fail(GPE_ERR_UNSUPPORTED, "gpe_set_nonlinearity_matrix: two components only (M=%d)", h->M);
  if (h->stream) HIPCHK(hipStreamSynchronize(h->stream));
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      for (int q = 0; q < 2; ++q) h->gm_c[a][b][q] = c[2 * (a + 2 * b) + q];
      for (int j = 0; j < 2; ++j)
        for (int q = 0; q < 2; ++q) h->gm_g[a][b][j][q] = g[2 * ((a + 2 * b) + 4 * j) + q];
    }
  // the vector-form G of the fused x-pass is off while the matrix form is set
  h->gmask = 0;
  h->g_cplx = 0;
  h->g_coupled = 0;
  memset(h->g, 0, sizeof h->g);
  memset(h->c, 0, sizeof h->c);
  h->gmat = true;
  return GPE_OK;
}

extern "C" int gpe_set_pump(gpe_handle* h, int32_t kind, const double* profile, const double* envelope, int64_t n_env) {
  if (!h) return fail(GPE_ERR_INVALID, "gpe_set_pump: null handle");
  HIPCHK(hipSetDevice(h->d.device));
  HIPCHK(hipStreamSynchronize(h->stream));  // queued launches may still read the profiles
  for (int c = 0; c < kMaxM; ++c) {
    h->pump_a[c] = nullptr;
    h->pump_b[c] = nullptr;
  }
  h->env.clear();
  h->pump_kind = GPE_PUMP_NONE;
  h->pump_samples = 0;
  if (kind == GPE_PUMP_NONE) return GPE_OK;
  if (kind != GPE_PUMP_TABLE && kind != GPE_PUMP_SEPARABLE && kind != GPE_PUMP_SAMPLES)
    return fail(GPE_ERR_INVALID, "gpe_set_pump: kind=%d", kind);
  if (kind == GPE_PUMP_SAMPLES) {
    if (envelope) return fail(GPE_ERR_INVALID, "gpe_set_pump: a sampled pump takes no envelope");
    if (n_env < 1) return fail(GPE_ERR_INVALID, "gpe_set_pump: a sampled pump needs n_env >= 1 samples");
    if ((double)n_env * (double)h->nsp * (double)h->csz > (double)kMaxPumpSampleBytes)
      return fail(GPE_ERR_UNSUPPORTED, "gpe_set_pump: %lld samples of %lld points exceed the %lld-byte limit per component",
                  (long long)n_env, (long long)h->nsp, (long long)kMaxPumpSampleBytes);
  }
  if (kind == GPE_PUMP_SEPARABLE && h->ndim == 3)
    return fail(GPE_ERR_UNSUPPORTED, "gpe_set_pump: 3D grids take the table form");
  if (!profile) return fail(GPE_ERR_INVALID, "gpe_set_pump: null profile");
  if (kind != GPE_PUMP_SAMPLES && (n_env < 0 || (n_env > 0 && !envelope)))
    return fail(GPE_ERR_INVALID, "gpe_set_pump: envelope");
  const int64_t per = kind == GPE_PUMP_TABLE     ? h->nsp
                    : kind == GPE_PUMP_SAMPLES ? n_env * h->nsp
                                               : (h->n0 + (h->ndim == 2 ? h->n1 : 0));
  for (int c = 0; c < h->M; ++c) {
    const double* pc = profile + 2 * per * c;
    bool nonzero = false;
    for (int64_t i = 0; i < 2 * per && !nonzero; ++i) nonzero = pc[i] != 0.0;
    if (!nonzero) continue;  // component not pumped: F_c = 0 exactly
    std::vector<double> tmp;
    auto put = [&](void** dst, size_t* cap, const double* src, int64_t n) -> int {
      if (h->f64) return upload_into(dst, cap, src, (size_t)n * 16);
      std::vector<float> f(2 * n);
      for (int64_t i = 0; i < 2 * n; ++i) f[i] = (float)src[i];
      return upload_into(dst, cap, f.data(), (size_t)n * 8);
    };
    int rc;
    if (kind == GPE_PUMP_TABLE || kind == GPE_PUMP_SAMPLES) {
      rc = put(&h->pump_buf_a[c], &h->pump_cap_a[c], pc, per);
    } else {
      rc = put(&h->pump_buf_a[c], &h->pump_cap_a[c], pc, h->n0);
      if (!rc && h->ndim == 2) rc = put(&h->pump_buf_b[c], &h->pump_cap_b[c], pc + 2 * h->n0, h->n1);
    }
    if (rc) return rc;
    h->pump_a[c] = h->pump_buf_a[c];
    h->pump_b[c] = (kind == GPE_PUMP_SEPARABLE && h->ndim == 2) ? h->pump_buf_b[c] : nullptr;
  }
  if (kind == GPE_PUMP_SAMPLES)
    h->pump_samples = n_env;
  else
    h->env.assign(envelope, envelope + 2 * n_env);
  h->pump_kind = kind;
  return GPE_OK;
}

extern "C" int gpe_set_noise(gpe_handle* h, int32_t kind, const double* eta, uint64_t seed, uint64_t call0) {
  if (!h) return fail(GPE_ERR_INVALID, "gpe_set_noise: null handle");
  if (kind != GPE_NOISE_NONE && kind != GPE_NOISE_COMPLEX && kind != GPE_NOISE_REAL)
    return fail(GPE_ERR_INVALID, "gpe_set_noise: kind=%d", kind);
  if (kind != GPE_NOISE_NONE && !eta) return fail(GPE_ERR_INVALID, "gpe_set_noise: null eta");
  h->noise_kind = kind;
  for (int c = 0; c < kMaxM; ++c) h->eta[c] = (kind != GPE_NOISE_NONE && c < h->M) ? eta[c] : 0.0;
  h->seed = seed;
  h->call = call0;
  return GPE_OK;
}

extern "C" int gpe_resolve_timestepping(double dt, double t0, double t1, int64_t nsaves, double* dt_out,
                                        int64_t* steps_per_save) {
  if (!(dt > 0) || nsaves < 1 || !dt_out || !steps_per_save || !(t1 > t0))
    return fail(GPE_ERR_INVALID, "gpe_resolve_timestepping: dt=%g t0=%g t1=%g nsaves=%lld", dt, t0, t1, (long long)nsaves);
  const double dT = (t1 - t0) / (double)nsaves;
  const double q = std::ceil(dT / dt);
  if (!(q >= 1) || q > 9.0e15) return fail(GPE_ERR_INVALID, "gpe_resolve_timestepping: steps_per_save=%g", q);
  *steps_per_save = (int64_t)q;
  *dt_out = dT / (double)(int64_t)q;
  return GPE_OK;
}

extern "C" int gpe_pump_times(double t0, double t1, double dt, int64_t nsaves, double* times, int64_t n) {
  double rdt;
  int64_t sps;
  int rc = gpe_resolve_timestepping(dt, t0, t1, nsaves, &rdt, &sps);
  if (rc) return rc;
  const int64_t S = sps * nsaves;
  if (!times || n != 1 + 2 * S) return fail(GPE_ERR_INVALID, "gpe_pump_times: n=%lld, expected %lld", (long long)n, (long long)(1 + 2 * S));
  double t = t0;
  times[0] = t0;
  for (int64_t s = 1; s <= S; ++s) {
    t += rdt;
    times[2 * s - 1] = t + rdt / 2;
    times[2 * s] = t + rdt;
  }
  return GPE_OK;
}

// Build exp(-i dt D)/(prod n) and exp(-i dt/2 V) for the resolved dt.
static int prepare_tables(gpe_handle* h, double dt) {
  if (h->disp_form != GPE_FORM_NONE && !(h->ed_dt == dt)) {
    EdArgs a;
    a.D = h->dD;
    a.out = h->ed;
    a.npts = h->ed_e0 * h->ed_e1;
    a.e0 = h->ed_e0;
    a.e1 = h->ed_e1;
    a.form = h->disp_form;
    a.M = h->M;
    a.dt = dt;
    a.scale = 1.0 / (double)h->nsp;
    const unsigned blocks = (unsigned)((a.npts + 255) / 256);
    if (h->f64)
      hipLaunchKernelGGL(ed_table_kernel<double>, dim3(blocks), dim3(256), 0, h->stream, a);
    else
      hipLaunchKernelGGL(ed_table_kernel<float>, dim3(blocks), dim3(256), 0, h->stream, a);
    HIPCHK(hipGetLastError());
    h->ed_dt = dt;
  }
  if (h->pot_form != GPE_FORM_NONE && !(h->ev_dt == dt)) {
    EvArgs a;
    a.V = h->dV;
    for (int c = 0; c < kMaxM; ++c) a.out[c] = h->ev[c];
    a.npts = h->nsp;
    a.form = h->pot_form;
    a.M = h->M;
    a.tau = dt / 2;
    const unsigned blocks = (unsigned)((h->nsp + 255) / 256);
    if (h->f64)
      hipLaunchKernelGGL(ev_table_kernel<double>, dim3(blocks), dim3(256), 0, h->stream, a);
    else
      hipLaunchKernelGGL(ev_table_kernel<float>, dim3(blocks), dim3(256), 0, h->stream, a);
    HIPCHK(hipGetLastError());
    h->ev_dt = dt;
  }
  return GPE_OK;
}

static PointTerms point_terms(const gpe_handle* h) {
  PointTerms pt;
  memset(&pt, 0, sizeof pt);
  pt.gmask = h->gmask;
  pt.g_cplx = h->g_cplx;
  pt.g_coupled = h->g_coupled;
  memcpy(pt.g, h->g, sizeof pt.g);
  memcpy(pt.c, h->c, sizeof pt.c);
  pt.pump_kind = h->pump_kind;
  pt.pump_stride = h->nsp;
  for (int c = 0; c < kMaxM; ++c) {
    pt.pump_a[c] = h->pump_a[c];
    pt.pump_b[c] = h->pump_b[c];
    pt.pot[c] = h->ev[c];
  }
  pt.pot_kind = h->pot_form != GPE_FORM_NONE;
  pt.noise_kind = h->noise_kind;
  for (int c = 0; c < kMaxM; ++c) pt.eta[c] = h->eta[c];
  pt.seed_lo = (uint32_t)h->seed;
  pt.seed_hi = (uint32_t)(h->seed >> 32);
  return pt;
}

// envelope sample i (complex); 1 when the pump is static
static void env_at(const gpe_handle* h, int64_t i, double out[2]) {
  if (h->env.empty()) {
    out[0] = 1.0;
    out[1] = 0.0;
  } else {
    i = std::min<int64_t>(i, (int64_t)h->env.size() / 2 - 1);
    out[0] = h->env[2 * i];
    out[1] = h->env[2 * i + 1];
  }
}

// Half-step `half` (0: at t_s + dt/2, 1: at t_s + dt) of step s (1-based).
static HalfStep make_half(const gpe_handle* h, double dt, int64_t s, int half, uint64_t call0) {
  HalfStep hs;
  hs.delta = dt / 2;
  const double q = hs.delta / 2;
  double en[2], ex[2];
  env_at(h, 2 * s - 2 + half, en);
  env_at(h, 2 * s - 1 + half, ex);
  hs.fnow[0] = q * en[0];
  hs.fnow[1] = q * en[1];
  hs.fnext[0] = q * ex[0];
  hs.fnext[1] = q * ex[1];
  hs.call = call0 + 2 * (uint64_t)(s - 1) + (uint64_t)half;
  // sample indices stay inside the sampled table also for the plan
  // measurement's steps 1 and 2 of a one-step solve
  const int64_t jmax = h->pump_samples > 0 ? h->pump_samples - 1 : 0;
  hs.jnow = std::min<int64_t>(2 * s - 2 + half, jmax);
  hs.jnext = std::min<int64_t>(2 * s - 1 + half, jmax);
  return hs;
}

static XPassArgs x_args(const gpe_handle* h) {
  XPassArgs a;
  memset(&a, 0, sizeof a);
  for (int c = 0; c < kMaxM; ++c) a.u[c] = h->u[c];
  a.nlines = h->npts / h->n0;
  a.ny = h->n1;
  a.tw = h->tw;
  a.tws = (int)(h->ntw / h->n0);
  a.pt = point_terms(h);
  return a;
}

static EdDesc ed_desc(const gpe_handle* h) {
  EdDesc d;
  memset(&d, 0, sizeof d);
  d.form = h->disp_form == GPE_FORM_NONE ? 0 : h->disp_form;
  d.fold = h->ed_fold;
  d.tab = h->ed;
  d.n0 = h->n0;
  d.n1 = h->n1;
  d.e0 = h->ed_e0;
  d.e1 = h->ed_e1;
  d.scale = 1.0 / (double)h->nsp;
  return d;
}

static YPassArgs y_args(const gpe_handle* h) {
  YPassArgs a;
  memset(&a, 0, sizeof a);
  for (int c = 0; c < kMaxM; ++c) a.u[c] = h->u[c];
  const int C = y_cols_per_wg((int)h->n1);
  a.nx = h->n0;
  a.ncolblk = h->n0 / C;
  a.nblocks = a.ncolblk * h->batch;
  a.xcd_remap = (a.nblocks % 8 == 0) ? 1 : 0;
  a.ed = ed_desc(h);
  a.tw = h->tw;
  a.tws = (int)(h->ntw / h->n1);
  return a;
}

// Column passes of a 3D grid (x fastest, then y, then z), with the built-in
// column count: which = 1 FFT_y, 2 IFFT_y (columns of every (z, trajectory)
// slab), 3 FFT_z -> exp(-i dt D) / prod(N) -> IFFT_z (columns over (x, y)).
static YPassArgs col_args_3d(const gpe_handle* h, int which) {
  YPassArgs a;
  memset(&a, 0, sizeof a);
  for (int c = 0; c < kMaxM; ++c) a.u[c] = h->u[c];
  const int64_t n = which == 3 ? h->gz : h->gy;
  const int C = y_cols_per_wg((int)n);
  a.nx = which == 3 ? h->n0 * h->gy : h->n0;
  a.ncolblk = a.nx / C;
  a.nblocks = a.ncolblk * (which == 3 ? h->batch : h->gz * h->batch);
  a.xcd_remap = (a.nblocks % 8 == 0) ? 1 : 0;
  a.mode = which == 3 ? 0 : which;
  if (which == 3) {
    a.ed = ed_desc(h);
    a.kx_n = h->n0;
    a.ky_n = h->gy;
  }
  a.tw = h->tw;
  a.tws = (int)(h->ntw / n);
  return a;
}

// The k-space part of a 2D or 3D step: the y-pass, or the three column
// passes of a 3D grid.
static hipError_t run_k(gpe_handle* h, const YPassArgs& ya) {
  if (h->ndim != 3) return run_y(h, ya);
  hipError_t e = run_col(h, col_args_3d(h, 1), h->gy);
  if (e == hipSuccess) e = run_col(h, col_args_3d(h, 3), h->gz);
  if (e == hipSuccess) e = run_col(h, col_args_3d(h, 2), h->gy);
  return e;
}

static Step1DArgs s1_args(gpe_handle* h, double dt) {
  Step1DArgs a;
  memset(&a, 0, sizeof a);
  for (int c = 0; c < kMaxM; ++c) a.u[c] = h->u[c];
  a.nlines = h->batch;
  a.delta = dt / 2;
  a.ed = ed_desc(h);
  a.tw = h->tw;
  a.tws = (int)(h->ntw / h->n0);
  a.pt = point_terms(h);
  a.env = h->env.empty() ? nullptr : h->d_env;
  return a;
}

// ---------------------------------------------------------------- planner
// The measured counterpart of FFTW's planner (the reference plans with
// FFTW.jl's default, src/misc.jl:53-58): the kernels' arithmetic is fixed, so
// the plan is the launch shape of the two 2D passes.  GPE_PLAN_MEASURE times
// every valid shape on the handle itself, with its tables and terms, before
// the first solve (and again when the terms that select kernel code change),
// and keeps the fastest.  The shapes change no arithmetic: results are
// identical bit for bit under every plan.
// Disjoint bit fields: noise 0, coupled G 1, complex G 2, G mask 3-6 (four
// components), dispersion form 7-9, potential form 10-11, pump kind 12-13,
// matrix G 14; bit 15 marks a key as set.
static uint64_t plan_key_of(const gpe_handle* h) {
  return (1ull << 15) | (uint64_t)(h->noise_kind != GPE_NOISE_NONE) | ((uint64_t)(h->g_coupled != 0) << 1) |
         ((uint64_t)(h->g_cplx != 0) << 2) | (((uint64_t)h->gmask & 15u) << 3) |
         (((uint64_t)h->disp_form & 7u) << 7) | (((uint64_t)h->pot_form & 3u) << 10) |
         (((uint64_t)h->pump_kind & 3u) << 12) | ((uint64_t)(h->gmat ? 1 : 0) << 14);
}

// Mean device time of one call of fn (the launches it issues), from HIP
// events on the handle's stream: two warm-up calls, a three-call estimate,
// then enough calls for about 2 ms (3..200).
static int time_launch(gpe_handle* h, const std::function<hipError_t()>& fn, double* us) {
  for (int i = 0; i < 2; ++i) HIPCHK(fn());
  HIPCHK(hipEventRecord(h->ev0, h->stream));
  for (int i = 0; i < 3; ++i) HIPCHK(fn());
  HIPCHK(hipEventRecord(h->ev1, h->stream));
  HIPCHK(hipEventSynchronize(h->ev1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  const double est = std::max((double)ms / 3, 1e-4);
  const int reps = (int)std::min(200.0, std::max(3.0, 2.0 / est));
  HIPCHK(hipEventRecord(h->ev0, h->stream));
  for (int i = 0; i < reps; ++i) HIPCHK(fn());
  HIPCHK(hipEventRecord(h->ev1, h->stream));
  HIPCHK(hipEventSynchronize(h->ev1));
  HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  *us = 1e3 * ms / reps;
  return GPE_OK;
}

// Measured plans of this process, keyed by what the timing depends on
// (device, grid, components, precision, batch, terms), so a new handle of a
// shape measured before reuses the result, as FFTW keeps its wisdom.
using PlanKey = std::tuple<int, int64_t, int64_t, int, int, int64_t, uint64_t>;
static std::mutex g_plan_mu;
static std::map<PlanKey, std::pair<int, int>> g_plans;

// A measurement times 6 shape pairs x 2 rounds x ~2 ms; it is skipped (the
// built-in shapes are used, nothing is cached) for runs of fewer steps,
// whose whole duration it would exceed several times.
constexpr int64_t kPlanMinSteps = 1000;

// Called with the tables prepared and before the start field is copied in:
// the candidates run in place on a zeroed field.  nsteps: the steps the
// caller is about to run.
static int ensure_plan(gpe_handle* h, double dt, int64_t nsteps) {
  if (h->ndim != 2 || h->planner == GPE_PLAN_ESTIMATE || h->plan_fixed) return GPE_OK;
  const uint64_t key = plan_key_of(h);
  if (h->plan_key == key) return GPE_OK;
  const PlanKey pk{h->d.device, h->n0, h->n1, h->M, h->f64 ? 1 : 0, h->batch, key};
  if (h->planner == GPE_PLAN_AUTO) {
    std::lock_guard<std::mutex> lock(g_plan_mu);
    auto it = g_plans.find(pk);
    if (it != g_plans.end()) {
      h->plan_w = it->second.first;
      h->plan_c = it->second.second;
      h->plan_key = key;
      h->plan_ncand = 0;
      return GPE_OK;
    }
  }
  if (h->planner == GPE_PLAN_AUTO && nsteps < kPlanMinSteps) {
    h->plan_w = h->plan_c = 0;
    return GPE_OK;
  }
  for (int c = 0; c < h->M; ++c) HIPCHK(hipMemsetAsync(h->u[c], 0, field_bytes(h), h->stream));
  XPassArgs xa = x_args(h);
  xa.flags = kInK | kEnd | kStart | kOutK;
  xa.h_end = make_half(h, dt, 1, 1, h->call);
  xa.h_start = make_half(h, dt, 2, 0, h->call);
  const YPassArgs ya = y_args(h);
  const int n0 = (int)h->n0, n1 = (int)h->n1;
  const int W = x_lines_per_wg(n0), C = y_cols_per_wg(n1);
  const int cw[2] = {W, 2 * W};
  const int cc[3] = {C > 1 ? C / 2 : 0, C, 2 * C};
  // Whole steps (column pass, then row pass) are timed, not the passes
  // alone: a row pass repeated without the column pass's 1/(Nx Ny) scale
  // lets the field grow until the half-step's sine/cosine leaves its fast
  // range, which is not the work of a real step.
  int n = 0;
  for (int i = 0; i < 2; ++i) {
    if (!x_plan_valid(n0, h->M, h->f64, cw[i])) continue;
    for (int k = 0; k < 3; ++k) {
      if (cc[k] < 1 || !y_plan_valid(n1, h->M, h->f64, cc[k]) || h->n0 % cc[k] != 0) continue;
      h->plan_cand[n][0] = cw[i];
      h->plan_cand[n][1] = cc[k];
      h->plan_us[n] = INFINITY;
      ++n;
    }
  }
  h->plan_ncand = n;
  // Two rounds over all candidates, keeping each one's faster round: the
  // first launches on a freshly allocated field run slow, which would
  // otherwise count against whichever shape is timed first.
  for (int round = 0; round < 2; ++round) {
    for (int i = 0; i < n; ++i) {
      const int w = h->plan_cand[i][0], c = h->plan_cand[i][1];
      double us = 0;
      int rc = time_launch(h, [&] {
        const hipError_t e = run_y_c(h, ya, c);
        return e != hipSuccess ? e : run_x_w(h, xa, w);
      }, &us);
      if (rc) return rc;
      h->plan_us[i] = std::min(h->plan_us[i], us);
    }
  }
  int best_w = W, best_c = y_rule_cols(n1, h->M, h->f64, h->n0, h->batch);
  double best = INFINITY;
  for (int i = 0; i < n; ++i)
    if (h->plan_us[i] < best) {
      best = h->plan_us[i];
      best_w = h->plan_cand[i][0];
      best_c = h->plan_cand[i][1];
    }
  h->plan_w = best_w;
  h->plan_c = best_c;
  h->plan_key = key;
  std::lock_guard<std::mutex> lock(g_plan_mu);
  g_plans[pk] = {best_w, best_c};
  return GPE_OK;
}

extern "C" int gpe_set_planner(gpe_handle* h, int32_t planner) {
  if (!h) return fail(GPE_ERR_INVALID, "gpe_set_planner: null handle");
  if (planner != GPE_PLAN_ESTIMATE && planner != GPE_PLAN_MEASURE && planner != GPE_PLAN_AUTO)
    return fail(GPE_ERR_INVALID, "gpe_set_planner: planner=%d", planner);
  h->planner = planner;
  h->plan_fixed = false;
  h->plan_key = 0;
  h->plan_w = h->plan_c = 0;
  return GPE_OK;
}

extern "C" int gpe_set_plan(gpe_handle* h, int32_t x_rows, int32_t y_cols) {
  if (!h) return fail(GPE_ERR_INVALID, "gpe_set_plan: null handle");
  if (h->ndim != 2 && (x_rows != 0 || y_cols != 0))
    return fail(GPE_ERR_UNSUPPORTED, "gpe_set_plan: launch shapes are chosen for 2D grids only");
  if (x_rows != 0 && !x_plan_valid((int)h->n0, h->M, h->f64, x_rows))
    return fail(GPE_ERR_INVALID, "gpe_set_plan: %d rows per workgroup is not a valid x-pass shape for n0=%lld",
                x_rows, (long long)h->n0);
  if (y_cols != 0 && (!y_plan_valid((int)h->n1, h->M, h->f64, y_cols) || h->n0 % y_cols != 0))
    return fail(GPE_ERR_INVALID, "gpe_set_plan: %d columns per workgroup is not a valid y-pass shape for n1=%lld",
                y_cols, (long long)h->n1);
  h->plan_w = x_rows;
  h->plan_c = y_cols;
  h->plan_fixed = true;
  return GPE_OK;
}

extern "C" int gpe_get_plan(gpe_handle* h, int32_t* x_rows, int32_t* y_cols, int32_t* cand, double* us) {
  if (!h || !x_rows || !y_cols) return fail(GPE_ERR_INVALID, "gpe_get_plan: null argument");
  if (h->ndim != 2) {
    *x_rows = x_lines_per_wg((int)h->n0);
    *y_cols = 0;
  } else {
    *x_rows = h->plan_w ? h->plan_w : x_lines_per_wg((int)h->n0);
    *y_cols = h->plan_c ? h->plan_c : y_rule_cols((int)h->n1, h->M, h->f64, h->n0, h->batch);
  }
  if (cand && us) {
    for (int i = 0; i < kPlanMaxCand; ++i) {
      const bool have = i < h->plan_ncand;
      cand[2 * i] = have ? h->plan_cand[i][0] : 0;
      cand[2 * i + 1] = have ? h->plan_cand[i][1] : 0;
      us[i] = have ? h->plan_us[i] : 0;
    }
  }
  return GPE_OK;
}

static int ensure_stage(gpe_handle* h) {
  for (int c = 0; c < h->M; ++c)
    if (!h->stage[c]) HIPCHK(hipMalloc(&h->stage[c], field_bytes(h)));
  return GPE_OK;
}

// Frame destination pointers for frame index f of result[], staging when the
// caller's memory is on the host.
struct FrameSink {
  bool to_host = false;
  void* kernel_dst[kMaxM] = {nullptr, nullptr};
  char* host_dst[kMaxM] = {nullptr, nullptr};
};

extern "C" int gpe_set_observables(gpe_handle* h, double* n, double* n2) {
  if (!h) return fail(GPE_ERR_INVALID, "gpe_set_observables: null handle");
  if ((n == nullptr) != (n2 == nullptr)) return fail(GPE_ERR_INVALID, "gpe_set_observables: give both n and n2, or neither");
  h->obs_n = n;
  h->obs_n2 = n2;
  return GPE_OK;
}

// The 2D step loop as a HIP graph: steps_per_graph repetitions of (y-pass,
// x-pass) with the arguments of a step inside a save block.  Those arguments
// are the same for every such step when no noise is drawn and the envelope is
// static (the Philox counter in HalfStep::call is then unused), so one graph,
// replayed once per save block, issues exactly the launches of the per-step
// loop.  The graph is built node by node from described launches rather than
// by capturing the stream: a capture in progress on one handle's stream makes
// synchronous runtime calls from other threads (other handles, the caller's
// own device work) fail, and distinct handles may run concurrently.
static int build_steps_graph(gpe_handle* h, const XPassArgs& xa, const YPassArgs& ya, int64_t steps_per_graph,
                             hipGraphExec_t* exec) {
  LaunchDesc dx, dy;
  const int64_t nlines = h->npts / h->n0;
  hipError_t e = h->f64 ? gpe_run_x_f64(h->M, h->n0, xa, nlines, h->plan_w, h->stream, &dx)
                        : gpe_run_x_f32(h->M, h->n0, xa, nlines, h->plan_w, h->stream, &dx);
  if (e == hipSuccess)
    e = h->f64 ? gpe_run_y_f64(h->M, h->n1, ya, h->plan_c, h->stream, &dy)
               : gpe_run_y_f32(h->M, h->n1, ya, h->plan_c, h->stream, &dy);
  if (e != hipSuccess) return fail(GPE_ERR_DEVICE, "gpe_solve: step graph: %s", hipGetErrorString(e));
  hipGraph_t g = nullptr;
  HIPCHK(hipGraphCreate(&g, 0));
  hipGraphNode_t prev = nullptr;
  auto add = [&](LaunchDesc& d) -> hipError_t {
    hipKernelNodeParams p;
    memset(&p, 0, sizeof p);
    void* args[] = {d.arg};
    p.func = const_cast<void*>(d.func);
    p.gridDim = d.grid;
    p.blockDim = d.block;
    p.sharedMemBytes = d.lds;
    p.kernelParams = args;
    p.extra = nullptr;
    hipGraphNode_t node = nullptr;
    const hipError_t r = hipGraphAddKernelNode(&node, g, prev ? &prev : nullptr, prev ? 1 : 0, &p);
    if (r == hipSuccess) prev = node;
    return r;
  };
  for (int64_t s = 0; s < steps_per_graph && e == hipSuccess; ++s) {
    e = add(dy);
    if (e == hipSuccess) e = add(dx);
  }
  if (e == hipSuccess) e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return fail(GPE_ERR_DEVICE, "gpe_solve: step graph: %s", hipGetErrorString(e));
  return GPE_OK;
}

// Releases a step graph on every return path, after the work queued on its
// stream (a replay may still be in flight when an error ends the loop).
struct GraphExecGuard {
  hipGraphExec_t& e;
  hipStream_t stream;
  ~GraphExecGuard() {
    if (e) {
      (void)hipStreamSynchronize(stream);
      (void)hipGraphExecDestroy(e);
    }
  }
};

// Whether a 2D solve replays a step graph.  Building it costs about as much
// host time as issuing the launches of its steps once, and each
// replay saves ~0.3-0.9 us per step on the device (measured on MI355X from
// 64x64 to 512x512), so AUTO replays when there are at least kGraphMinSaves
// save blocks.
constexpr int64_t kGraphMinSaves = 8;

static bool graph_eligible(const gpe_handle* h, int64_t sps) {
  return h->ndim == 2 && h->noise_kind == GPE_NOISE_NONE && h->env.empty() && h->pump_kind != GPE_PUMP_SAMPLES &&
         sps >= 2 && !h->gmat;
}

static bool use_step_graph(const gpe_handle* h, int64_t sps, int64_t nsaves) {
  if (!graph_eligible(h, sps) || h->schedule == GPE_SCHEDULE_LAUNCHES) return false;
  if (h->schedule == GPE_SCHEDULE_GRAPH) return true;
  return nsaves >= kGraphMinSaves;
}

extern "C" int gpe_solve(gpe_handle* h, const void* const* u0, double t0, double t1, double dt, int64_t nsaves,
                         int32_t save_start, void* const* result, double* ts) {
  if (!h || !u0) return fail(GPE_ERR_INVALID, "gpe_solve: null argument");
  for (int c = 0; c < h->M; ++c) {
    if (!u0[c]) return fail(GPE_ERR_INVALID, "gpe_solve: null field pointer for component %d", c);
    if (result && !result[c]) return fail(GPE_ERR_INVALID, "gpe_solve: null result pointer for component %d", c);
  }
  HIPCHK(hipSetDevice(h->d.device));
  double rdt;
  int64_t sps;
  int rc = gpe_resolve_timestepping(dt, t0, t1, nsaves, &rdt, &sps);
  if (rc) return rc;
  const int64_t S = sps * nsaves;
  if (h->pump_kind != GPE_PUMP_NONE && !h->env.empty() && (int64_t)h->env.size() != 2 * (1 + 2 * S))
    return fail(GPE_ERR_STATE, "gpe_solve: pump envelope has %lld samples, this solve needs %lld (gpe_pump_times)",
                (long long)(h->env.size() / 2), (long long)(1 + 2 * S));
  if (h->pump_kind == GPE_PUMP_SAMPLES && h->pump_samples != 1 + 2 * S)
    return fail(GPE_ERR_STATE, "gpe_solve: the sampled pump has %lld samples, this solve needs %lld (gpe_pump_times)",
                (long long)h->pump_samples, (long long)(1 + 2 * S));
  if (h->ndim == 1 && !h->env.empty()) {
    const int64_t need = (int64_t)h->env.size();
    if (need > h->d_env_cap) {
      void* p = h->d_env;
      dfree(p);
      h->d_env = nullptr;
      HIPCHK(hipMalloc(&p, need * sizeof(double)));
      h->d_env = (double*)p;
      h->d_env_cap = need;
    }
    HIPCHK(hipMemcpyAsync(h->d_env, h->env.data(), need * sizeof(double), hipMemcpyHostToDevice, h->stream));
  }
  const bool graph = use_step_graph(h, sps, nsaves);
  if (h->schedule == GPE_SCHEDULE_GRAPH && !graph)
    return fail(GPE_ERR_UNSUPPORTED,
                "gpe_solve: the graph schedule needs a 2D grid, no noise, a static envelope and >= 2 steps per save");
  rc = prepare_tables(h, rdt);
  if (rc) return rc;
  rc = ensure_plan(h, rdt, S);
  if (rc) return rc;
  const size_t fb = field_bytes(h);
  for (int c = 0; c < h->M; ++c) HIPCHK(hipMemcpyAsync(h->u[c], u0[c], fb, hipMemcpyDefault, h->stream));
  const int64_t nfr = nsaves + (save_start ? 1 : 0);
  const bool obs = h->obs_n != nullptr;
  bool res_dev = result != nullptr;
  for (int c = 0; c < h->M && res_dev; ++c) res_dev = is_device_ptr(result[c]);
  if (!res_dev) {
    rc = ensure_stage(h);
    if (rc) return rc;
  }
  const size_t obs_per_frame = (size_t)h->M * h->nsp;
  if (obs) {
    const size_t need = 2 * obs_per_frame * nfr;
    if (need > h->d_obs_cap) {
      void* p = h->d_obs;
      dfree(p);
      h->d_obs = nullptr;
      HIPCHK(hipMalloc(&p, need * sizeof(double)));
      h->d_obs = (double*)p;
      h->d_obs_cap = need;
    }
  }
  // observables of frame f from the real-space field at `src`
  auto observe = [&](void* const* src, int64_t f) -> int {
    if (!obs) return GPE_OK;
    ObsArgs oa;
    memset(&oa, 0, sizeof oa);
    for (int c = 0; c < h->M; ++c) {
      oa.u[c] = src[c];
      oa.n1[c] = h->d_obs + ((size_t)c * nfr + f) * h->nsp;
      oa.n2[c] = h->d_obs + obs_per_frame * nfr + ((size_t)c * nfr + f) * h->nsp;
    }
    oa.nsp = h->nsp;
    oa.batch = h->batch;
    oa.M = h->M;
    const unsigned blocks = (unsigned)((h->nsp + 255) / 256);
    if (h->f64)
      hipLaunchKernelGGL(observables_kernel<double>, dim3(blocks), dim3(256), 0, h->stream, oa);
    else
      hipLaunchKernelGGL(observables_kernel<float>, dim3(blocks), dim3(256), 0, h->stream, oa);
    HIPCHK(hipGetLastError());
    return GPE_OK;
  };
  if (save_start && result)
    for (int c = 0; c < h->M; ++c) HIPCHK(hipMemcpyAsync(result[c], u0[c], fb, hipMemcpyDefault, h->stream));
  if (save_start) {
    rc = observe(h->u, 0);
    if (rc) return rc;
  }
  const bool want_frames = result != nullptr || obs;
  auto frame_ptr = [&](int c, int64_t f) -> void* {
    if (!want_frames) return nullptr;
    return res_dev ? (void*)((char*)result[c] + (size_t)f * fb) : h->stage[c];
  };
  auto flush_frame = [&](int64_t f) -> int {
    void* src[kMaxM];
    for (int c = 0; c < h->M; ++c) src[c] = frame_ptr(c, f);
    int r = observe(src, f);
    if (r) return r;
    if (res_dev || !result) return GPE_OK;
    for (int c = 0; c < h->M; ++c)
      HIPCHK(hipMemcpyAsync((char*)result[c] + (size_t)f * fb, h->stage[c], fb, hipMemcpyDeviceToHost, h->stream));
    return GPE_OK;
  };
  const uint64_t call0 = h->call;
  double t = t0;
  if (ts) ts[0] = t0;
  hipGraphExec_t gexec = nullptr;
  if (graph) {
    XPassArgs xg = x_args(h);
    xg.flags = kInK | kEnd | kStart | kOutK;
    xg.h_end = make_half(h, rdt, 1, 1, call0);
    xg.h_start = make_half(h, rdt, 2, 0, call0);
    rc = build_steps_graph(h, xg, y_args(h), sps - 1, &gexec);
    if (rc) return rc;
  }
  GraphExecGuard exec_guard{gexec, h->stream};
  HIPCHK(hipEventRecord(h->ev0, h->stream));
  if (graph) {
    // per save block: the replayed steps, then the block's last step with its
    // frame (and, at the end of the solve, without the start half-step)
    XPassArgs xa = x_args(h);
    const YPassArgs ya = y_args(h);
    xa.flags = kStart | kOutK;
    xa.h_start = make_half(h, rdt, 1, 0, call0);
    HIPCHK(run_x(h, xa));
    for (int64_t n = 0; n < nsaves; ++n) {
      HIPCHK(hipGraphLaunch(gexec, h->stream));
      const int64_t s = (n + 1) * sps;
      for (int64_t k = 0; k < sps; ++k) t += rdt;
      HIPCHK(run_y(h, ya));
      const int64_t f = n + (save_start ? 1 : 0);
      xa.h_end = make_half(h, rdt, s, 1, call0);
      for (int c = 0; c < kMaxM; ++c) xa.frame[c] = c < h->M ? frame_ptr(c, f) : nullptr;
      if (s < S) {
        xa.flags = kInK | kEnd | kStart | kOutK;
        xa.h_start = make_half(h, rdt, s + 1, 0, call0);
      } else {
        xa.flags = kInK | kEnd;
      }
      HIPCHK(run_x(h, xa));
      rc = flush_frame(f);
      if (rc) return rc;
      if (ts) ts[n + (save_start ? 1 : 0)] = t;
    }
  } else if (h->gmat) {
    // matrix G: the x-passes only transform; gmat_half_kernel applies the
    // half-steps (and writes the frames) in real space between them
    XPassArgs xa = x_args(h);
    // (1D handles have no column pass: y_args would divide by n1 / 8 = 0)
    const YPassArgs ya = h->ndim == 2 ? y_args(h) : YPassArgs{};
    const bool grid = h->ndim >= 2;
    GmatArgs ga;
    memset(&ga, 0, sizeof ga);
    for (int c = 0; c < 2; ++c) {
      ga.u[c] = h->u[c];
      ga.pump_a[c] = h->pump_a[c];
      ga.pump_b[c] = h->pump_b[c];
    }
    ga.npts = h->npts;
    ga.n0 = h->n0;
    ga.nsp = h->nsp;
    ga.pump_kind = h->pump_kind;
    ga.noise_kind = h->noise_kind;
    ga.pot_kind = h->pot_form == GPE_FORM_NONE ? 0 : (h->pot_form == GPE_FORM_SCALAR ? 1 : 2);
    for (int c = 0; c < 2; ++c) ga.pot[c] = h->ev[c];
    for (int c = 0; c < 2; ++c) ga.eta[c] = h->eta[c];
    ga.seed_lo = (uint32_t)h->seed;
    ga.seed_hi = (uint32_t)(h->seed >> 32);
    memcpy(ga.g, h->gm_g, sizeof ga.g);
    memcpy(ga.c, h->gm_c, sizeof ga.c);
    const unsigned gblocks = (unsigned)((h->npts + 255) / 256);
    auto run_gm = [&]() -> hipError_t {
      if (h->f64)
        hipLaunchKernelGGL(gmat_half_kernel<double>, dim3(gblocks), dim3(256), 0, h->stream, ga);
      else
        hipLaunchKernelGGL(gmat_half_kernel<float>, dim3(gblocks), dim3(256), 0, h->stream, ga);
      return hipGetLastError();
    };
    // 1D: the whole-step kernel with every real-space term off does the
    // k-space part of one step (its half-steps are then the identity)
    Step1DArgs ka = s1_args(h, rdt);
    memset(&ka.pt, 0, sizeof ka.pt);
    ka.env = nullptr;
    ka.nsteps = 1;
    for (int c = 0; c < kMaxM; ++c) ka.frame[c] = nullptr;
    ga.flags = kStart;
    ga.h_start = make_half(h, rdt, 1, 0, call0);
    HIPCHK(run_gm());
    if (grid) {
      xa.flags = kOutK;
      HIPCHK(run_x(h, xa));
    }
    for (int64_t s = 1; s <= S; ++s) {
      t += rdt;
      if (grid) {
        HIPCHK(run_k(h, ya));
        xa.flags = kInK;
        HIPCHK(run_x(h, xa));
      } else {
        HIPCHK(run_1d(h, ka));
      }
      const bool save = (s % sps) == 0;
      const int64_t f = s / sps - 1 + (save_start ? 1 : 0);
      for (int c = 0; c < kMaxM; ++c) ga.frame[c] = save ? frame_ptr(c, f) : nullptr;
      ga.flags = kEnd | (s < S ? kStart : 0);
      ga.h_end = make_half(h, rdt, s, 1, call0);
      if (s < S) ga.h_start = make_half(h, rdt, s + 1, 0, call0);
      HIPCHK(run_gm());
      if (s < S && grid) {
        xa.flags = kOutK;
        HIPCHK(run_x(h, xa));
      }
      if (save) {
        rc = flush_frame(f);
        if (rc) return rc;
        if (ts) ts[s / sps - (save_start ? 0 : 1)] = t;
      }
    }
  } else if (h->ndim >= 2) {
    XPassArgs xa = x_args(h);
    const YPassArgs ya = h->ndim == 2 ? y_args(h) : YPassArgs{};
    xa.flags = kStart | kOutK;
    xa.h_start = make_half(h, rdt, 1, 0, call0);
    HIPCHK(run_x(h, xa));
    for (int64_t s = 1; s <= S; ++s) {
      t += rdt;
      HIPCHK(run_k(h, ya));
      const bool save = (s % sps) == 0;
      const int64_t f = s / sps - 1 + (save_start ? 1 : 0);
      xa.h_end = make_half(h, rdt, s, 1, call0);
      for (int c = 0; c < kMaxM; ++c) xa.frame[c] = (save && c < h->M) ? frame_ptr(c, f) : nullptr;
      if (s < S) {
        xa.flags = kInK | kEnd | kStart | kOutK;
        xa.h_start = make_half(h, rdt, s + 1, 0, call0);
      } else {
        xa.flags = kInK | kEnd;
      }
      HIPCHK(run_x(h, xa));
      if (save) {
        rc = flush_frame(f);
        if (rc) return rc;
        if (ts) ts[s / sps - (save_start ? 0 : 1)] = t;
      }
    }
  } else {
    Step1DArgs a = s1_args(h, rdt);
    for (int64_t n = 0; n < nsaves; ++n) {
      a.nsteps = sps;
      a.env_base = 2 * n * sps;
      a.call0 = call0 + 2 * (uint64_t)(n * sps);
      const int64_t f = n + (save_start ? 1 : 0);
      for (int c = 0; c < kMaxM; ++c) a.frame[c] = c < h->M ? frame_ptr(c, f) : nullptr;
      HIPCHK(run_1d(h, a));
      rc = flush_frame(f);
      if (rc) return rc;
      for (int64_t k = 0; k < sps; ++k) t += rdt;
      if (ts) ts[n + (save_start ? 1 : 0)] = t;
    }
  }
  HIPCHK(hipEventRecord(h->ev1, h->stream));
  h->call = call0 + 2 * (uint64_t)S;
  if (obs) {
    HIPCHK(hipMemcpyAsync(h->obs_n, h->d_obs, obs_per_frame * nfr * sizeof(double), hipMemcpyDefault, h->stream));
    HIPCHK(hipMemcpyAsync(h->obs_n2, h->d_obs + obs_per_frame * nfr, obs_per_frame * nfr * sizeof(double),
                          hipMemcpyDefault, h->stream));
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipEventElapsedTime(&h->last_ms, h->ev0, h->ev1));
  h->configured = true;
  h->dt = rdt;
  if (h->ndim >= 2) h->last_schedule = graph ? GPE_SCHEDULE_GRAPH : GPE_SCHEDULE_LAUNCHES;
  return GPE_OK;
}

extern "C" int gpe_set_schedule(gpe_handle* h, int32_t schedule) {
  if (!h) return fail(GPE_ERR_INVALID, "gpe_set_schedule: null handle");
  if (schedule != GPE_SCHEDULE_AUTO && schedule != GPE_SCHEDULE_LAUNCHES && schedule != GPE_SCHEDULE_GRAPH)
    return fail(GPE_ERR_INVALID, "gpe_set_schedule: schedule=%d", schedule);
  h->schedule = schedule;
  return GPE_OK;
}

extern "C" int gpe_get_schedule(gpe_handle* h, int32_t* schedule) {
  if (!h || !schedule) return fail(GPE_ERR_INVALID, "gpe_get_schedule: null argument");
  *schedule = h->last_schedule;
  return GPE_OK;
}

extern "C" int gpe_kspace_g2(gpe_handle* h, const void* frame, double commutator, double* nk, double* g2) {
  if (!h || !frame || !nk || !g2) return fail(GPE_ERR_INVALID, "gpe_kspace_g2: null argument");
  if (h->ndim != 1)
    return fail(GPE_ERR_UNSUPPORTED, "gpe_kspace_g2: 1D ensembles only (a 2D grid's pair matrix has prod(n)^2 entries)");
  if (!(commutator >= 0) || !std::isfinite(commutator))
    return fail(GPE_ERR_INVALID, "gpe_kspace_g2: commutator=%g", commutator);
  HIPCHK(hipSetDevice(h->d.device));
  const int64_t n = h->n0, nb = h->batch;
  const void* src = frame;
  if (!is_device_ptr(frame)) {
    int rc = ensure_stage(h);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(h->stage[0], frame, field_bytes(h), hipMemcpyHostToDevice, h->stream));
    src = h->stage[0];
  }
  const size_t need = (size_t)(nb * n + 2 * n + n * n) * sizeof(double);
  int rc = ensure_buf(&h->g2_buf, &h->g2_cap, need);
  if (rc) return rc;
  double* a = (double*)h->g2_buf;
  double* nbar = a + nb * n;
  double* dnk = nbar + n;
  double* dg2 = dnk + n;
  KAbs2Args ka;
  ka.u = src;
  ka.a = a;
  ka.batch = nb;
  ka.tw = h->tw;
  ka.tws = (int)(h->ntw / n);
  HIPCHK(h->f64 ? gpe_run_kabs2_f64(n, ka, h->stream) : gpe_run_kabs2_f32(n, ka, h->stream));
  hipLaunchKernelGGL(kspace_mean_kernel<0>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, (const double*)a, n,
                     nb, commutator, nbar, dnk);
  hipLaunchKernelGGL(kspace_g2_kernel<0>, dim3((unsigned)(n / 16), (unsigned)(n / 16)), dim3(256), 0, h->stream,
                     (const double*)a, (const double*)nbar, n, nb, commutator, dg2);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(nk, dnk, n * sizeof(double), hipMemcpyDefault, h->stream));
  HIPCHK(hipMemcpyAsync(g2, dg2, (size_t)n * n * sizeof(double), hipMemcpyDefault, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return GPE_OK;
}

extern "C" int gpe_windowed_g1(gpe_handle* h, const void* frame, const double* w1, const double* w2, double* g1) {
  if (!h || !frame || !w1 || !w2 || !g1) return fail(GPE_ERR_INVALID, "gpe_windowed_g1: null argument");
  if (h->ndim != 1)
    return fail(GPE_ERR_UNSUPPORTED, "gpe_windowed_g1: 1D ensembles only (a 2D grid's pair matrix has prod(n)^2 entries)");
  HIPCHK(hipSetDevice(h->d.device));
  const int64_t n = h->n0, nb = h->batch;
  const void* src = frame;
  if (!is_device_ptr(frame)) {
    int rc = ensure_stage(h);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(h->stage[0], frame, field_bytes(h), hipMemcpyHostToDevice, h->stream));
    src = h->stage[0];
  }
  // complex FP64: two transformed ensembles, the windows, the matrix
  const size_t need = (size_t)(2 * 2 * nb * n + 2 * 2 * n + 2 * n * n) * sizeof(double);
  int rc = ensure_buf(&h->g2_buf, &h->g2_cap, need);
  if (rc) return rc;
  double* x1 = (double*)h->g2_buf;
  double* x2 = x1 + 2 * nb * n;
  double* dw1 = x2 + 2 * nb * n;
  double* dw2 = dw1 + 2 * n;
  double* dg1 = dw2 + 2 * n;
  HIPCHK(hipMemcpyAsync(dw1, w1, 2 * n * sizeof(double), hipMemcpyDefault, h->stream));
  HIPCHK(hipMemcpyAsync(dw2, w2, 2 * n * sizeof(double), hipMemcpyDefault, h->stream));
  KWinArgs ka;
  ka.u = src;
  ka.w1 = dw1;
  ka.w2 = dw2;
  ka.x1 = x1;
  ka.x2 = x2;
  ka.batch = nb;
  ka.tw = h->tw;
  ka.tws = (int)(h->ntw / n);
  HIPCHK(h->f64 ? gpe_run_kwin_f64(n, ka, h->stream) : gpe_run_kwin_f32(n, ka, h->stream));
  hipLaunchKernelGGL(kspace_g1_kernel<0>, dim3((unsigned)(n / 16), (unsigned)(n / 16)), dim3(256), 0, h->stream,
                     (const double*)x1, (const double*)x2, n, nb, dg1);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(g1, dg1, 2 * (size_t)n * n * sizeof(double), hipMemcpyDefault, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return GPE_OK;
}

extern "C" int gpe_density_g2_x(gpe_handle* h, const void* frame, double commutator, double* g2r, double* nphys) {
  if (!h || !frame || !g2r || !nphys) return fail(GPE_ERR_INVALID, "gpe_density_g2_x: null argument");
  if (!(commutator >= 0) || !std::isfinite(commutator))
    return fail(GPE_ERR_INVALID, "gpe_density_g2_x: commutator=%g", commutator);
  HIPCHK(hipSetDevice(h->d.device));
  const int64_t n = h->n0, nlines = h->npts / h->n0;
  const void* src = frame;
  if (!is_device_ptr(frame)) {
    int rc = ensure_stage(h);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(h->stage[0], frame, field_bytes(h), hipMemcpyHostToDevice, h->stream));
    src = h->stage[0];
  }
  const int nblk = gpe_rowcorr_blocks(n, nlines);
  const size_t need = (size_t)(nblk * (n + 1) + n + 1) * sizeof(double);
  int rc = ensure_buf(&h->g2_buf, &h->g2_cap, need);
  if (rc) return rc;
  double* partial = (double*)h->g2_buf;
  double* dg2 = partial + (size_t)nblk * (n + 1);
  double* dnp = dg2 + n;
  RowCorrArgs ra;
  ra.u = src;
  ra.nlines = nlines;
  ra.partial = partial;
  ra.tw = h->tw;
  ra.tws = (int)(h->ntw / n);
  HIPCHK(h->f64 ? gpe_run_rowcorr_f64(n, ra, nblk, h->stream) : gpe_run_rowcorr_f32(n, ra, nblk, h->stream));
  hipLaunchKernelGGL(row_autocorr_final_kernel<0>, dim3((unsigned)((n + 1 + 255) / 256)), dim3(256), 0, h->stream,
                     (const double*)partial, nblk, n, nlines, commutator, dg2, dnp);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(g2r, dg2, n * sizeof(double), hipMemcpyDefault, h->stream));
  HIPCHK(hipMemcpyAsync(nphys, dnp, sizeof(double), hipMemcpyDefault, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return GPE_OK;
}

extern "C" int gpe_get_field(gpe_handle* h, void* const* u) {
  if (!h || !u) return fail(GPE_ERR_INVALID, "gpe_get_field: null argument");
  HIPCHK(hipSetDevice(h->d.device));
  for (int c = 0; c < h->M; ++c) HIPCHK(hipMemcpyAsync(u[c], h->u[c], field_bytes(h), hipMemcpyDefault, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return GPE_OK;
}

extern "C" int gpe_last_solve_ms(gpe_handle* h, float* ms) {
  if (!h || !ms) return fail(GPE_ERR_INVALID, "gpe_last_solve_ms: null argument");
  *ms = h->last_ms;
  return GPE_OK;
}

// Continue from the current field with the last solve's dt and the envelope
// frozen at its last sample, no frame output.  With per_kernel, HIP events
// bracket every launch and the mean durations are returned.
static int advance_impl(gpe_handle* h, int64_t nsteps, bool per_kernel, double* ms_x, double* ms_y) {
  if (!h->configured) return fail(GPE_ERR_STATE, "gpe_advance: no previous gpe_solve");
  if (nsteps < 1) return fail(GPE_ERR_INVALID, "gpe_advance: nsteps=%lld", (long long)nsteps);
  if (h->gmat) return fail(GPE_ERR_UNSUPPORTED, "gpe_advance: matrix nonlinearity not supported");
  if (h->pump_kind == GPE_PUMP_SAMPLES)
    return fail(GPE_ERR_UNSUPPORTED, "gpe_advance: a sampled pump covers the times of one solve only");
  HIPCHK(hipSetDevice(h->d.device));
  const std::vector<double> saved_env = h->env;
  if (!h->env.empty()) {
    const double re = h->env[h->env.size() - 2], im = h->env[h->env.size() - 1];
    h->env.assign(2 * (1 + 2 * nsteps), 0.0);
    for (int64_t i = 0; i < 1 + 2 * nsteps; ++i) {
      h->env[2 * i] = re;
      h->env[2 * i + 1] = im;
    }
    if (h->ndim == 1) {
      const int64_t need = (int64_t)h->env.size();
      if (need > h->d_env_cap) {
        void* p = h->d_env;
        dfree(p);
        h->d_env = nullptr;
        HIPCHK(hipMalloc(&p, need * sizeof(double)));
        h->d_env = (double*)p;
        h->d_env_cap = need;
      }
      HIPCHK(hipMemcpyAsync(h->d_env, h->env.data(), need * sizeof(double), hipMemcpyHostToDevice, h->stream));
    }
  }
  std::vector<hipEvent_t> evs;
  std::vector<int> kinds;
  auto launch = [&](int kind, auto&& fn) -> int {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (per_kernel) {
      HIPCHK(hipEventCreate(&e0));
      HIPCHK(hipEventCreate(&e1));
      HIPCHK(hipEventRecord(e0, h->stream));
    }
    HIPCHK(fn());
    if (per_kernel) {
      HIPCHK(hipEventRecord(e1, h->stream));
      evs.push_back(e0);
      evs.push_back(e1);
      kinds.push_back(kind);
    }
    return GPE_OK;
  };
  int rc = GPE_OK;
  const uint64_t call0 = h->call;
  HIPCHK(hipEventRecord(h->ev0, h->stream));
  if (h->ndim >= 2) {
    XPassArgs xa = x_args(h);
    const YPassArgs ya = h->ndim == 2 ? y_args(h) : YPassArgs{};
    xa.flags = kStart | kOutK;
    xa.h_start = make_half(h, h->dt, 1, 0, call0);
    rc = launch(0, [&] { return run_x(h, xa); });
    for (int64_t s = 1; s <= nsteps && rc == GPE_OK; ++s) {
      rc = launch(1, [&] { return run_k(h, ya); });
      if (rc) break;
      xa.h_end = make_half(h, h->dt, s, 1, call0);
      if (s < nsteps) {
        xa.flags = kInK | kEnd | kStart | kOutK;
        xa.h_start = make_half(h, h->dt, s + 1, 0, call0);
      } else {
        xa.flags = kInK | kEnd;
      }
      rc = launch(0, [&] { return run_x(h, xa); });
    }
  } else {
    Step1DArgs a = s1_args(h, h->dt);
    a.nsteps = nsteps;
    a.env_base = 0;
    a.call0 = call0;
    rc = launch(2, [&] { return run_1d(h, a); });
  }
  h->env = saved_env;
  if (rc) return rc;
  HIPCHK(hipEventRecord(h->ev1, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipEventElapsedTime(&h->last_ms, h->ev0, h->ev1));
  h->call = call0 + 2 * (uint64_t)nsteps;
  if (per_kernel) {
    double acc[3] = {0, 0, 0};
    int64_t cnt[3] = {0, 0, 0};
    for (size_t i = 0; i < kinds.size(); ++i) {
      float ms;
      HIPCHK(hipEventElapsedTime(&ms, evs[2 * i], evs[2 * i + 1]));
      acc[kinds[i]] += ms;
      cnt[kinds[i]]++;
    }
    for (hipEvent_t e : evs) (void)hipEventDestroy(e);
    // latency of an event pair with nothing between it, on the same stream
    double empty = 0;
    const int npair = 64;
    std::vector<hipEvent_t> pe(2 * npair);
    for (auto& e : pe) HIPCHK(hipEventCreate(&e));
    for (int i = 0; i < npair; ++i) {
      HIPCHK(hipEventRecord(pe[2 * i], h->stream));
      HIPCHK(hipEventRecord(pe[2 * i + 1], h->stream));
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int i = 0; i < npair; ++i) {
      float ms;
      HIPCHK(hipEventElapsedTime(&ms, pe[2 * i], pe[2 * i + 1]));
      empty += ms;
    }
    for (auto& e : pe) (void)hipEventDestroy(e);
    empty /= npair;
    h->event_pair_ms = empty;
    if (ms_x) *ms_x = h->ndim >= 2 ? (cnt[0] ? acc[0] / cnt[0] : 0) : (cnt[2] ? acc[2] / cnt[2] : 0);
    if (ms_y) *ms_y = cnt[1] ? acc[1] / cnt[1] : 0;
  }
  return GPE_OK;
}

extern "C" int gpe_advance(gpe_handle* h, int64_t nsteps, float* ms) {
  if (!h) return fail(GPE_ERR_INVALID, "gpe_advance: null handle");
  int rc = advance_impl(h, nsteps, false, nullptr, nullptr);
  if (rc) return rc;
  if (ms) *ms = h->last_ms;
  return GPE_OK;
}

extern "C" int gpe_profile_kernels(gpe_handle* h, int64_t nsteps, double* ms_x, double* ms_y, double* ms_event) {
  if (!h) return fail(GPE_ERR_INVALID, "gpe_profile_kernels: null handle");
  int rc = advance_impl(h, nsteps, true, ms_x, ms_y);
  if (rc) return rc;
  if (ms_event) *ms_event = h->event_pair_ms;
  return GPE_OK;
}

// ------------------------------------------------------------------ sweeps
// Independent 1D single-component instances of uneven length, advanced
// together (one launch per distinct length and save block).
struct gpe_sweep {
  int32_t device = 0;
  bool f64 = false;
  size_t csz = 8;
  int64_t ninst = 0, total = 0, nmax = 0;
  std::vector<int64_t> n, off;
  std::vector<double> L;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  float last_ms = 0;
  void* u = nullptr;          // concatenated field
  void* tw = nullptr;         // twiddles of length nmax
  double* dD = nullptr;       // FP64 D tables
  void* ed = nullptr;         // exp(-i dt D) (field precision)
  double ed_dt = NAN;
  void* pump = nullptr;       // concatenated profiles (field precision) or null
  std::vector<int64_t> pump_off;
  std::vector<double> env;
  double* d_env = nullptr;
  int64_t d_env_cap = 0;
  std::vector<double> g, c;   // complex per instance
  int noise_kind = GPE_NOISE_NONE;
  std::vector<double> eta;
  uint64_t seed = 0, call = 0;
  SweepInst* d_inst = nullptr;
  std::vector<int64_t> sizes;                 // distinct lengths
  std::vector<int32_t*> d_ids;                // instance ids per length
  std::vector<int64_t> nids;
  void* d_res = nullptr;      // device result staging when the caller's result is on the host
  size_t d_res_cap = 0;
};

extern "C" int gpe_sweep_destroy(gpe_sweep* s) {
  if (!s) return GPE_OK;
  (void)hipSetDevice(s->device);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  dfree(s->u);
  dfree(s->tw);
  void* p = s->dD;
  dfree(p);
  dfree(s->ed);
  dfree(s->pump);
  p = s->d_env;
  dfree(p);
  p = s->d_inst;
  dfree(p);
  for (auto& q : s->d_ids) {
    void* v = q;
    dfree(v);
  }
  dfree(s->d_res);
  if (s->ev0) (void)hipEventDestroy(s->ev0);
  if (s->ev1) (void)hipEventDestroy(s->ev1);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return GPE_OK;
}

extern "C" int gpe_sweep_create(int64_t ninst, const int64_t* n, const double* L, int32_t dtype, int32_t device,
                                gpe_sweep** out) {
  if (!out || !n || !L || ninst < 1) return fail(GPE_ERR_INVALID, "gpe_sweep_create: bad argument");
  *out = nullptr;
  if (dtype != GPE_C64 && dtype != GPE_C128) return fail(GPE_ERR_INVALID, "gpe_sweep_create: dtype=%d", dtype);
  for (int64_t i = 0; i < ninst; ++i) {
    if (!is_pow2(n[i]) || n[i] < 16 || n[i] > 4096)
      return fail(GPE_ERR_UNSUPPORTED, "gpe_sweep_create: n[%lld]=%lld (supported: powers of two 16..4096)",
                  (long long)i, (long long)n[i]);
    if (!(L[i] > 0) || !std::isfinite(L[i])) return fail(GPE_ERR_INVALID, "gpe_sweep_create: L[%lld]", (long long)i);
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(GPE_ERR_DEVICE, "gpe_sweep_create: no HIP device");
  if (device < 0 || device >= ndev) return fail(GPE_ERR_INVALID, "gpe_sweep_create: device=%d", device);
  HIPCHK(hipSetDevice(device));
  gpe_sweep* s = new gpe_sweep();
  auto bail = [&](int code) {
    gpe_sweep_destroy(s);
    return code;
  };
  s->device = device;
  s->f64 = dtype == GPE_C128;
  s->csz = s->f64 ? 16 : 8;
  s->ninst = ninst;
  s->n.assign(n, n + ninst);
  s->L.assign(L, L + ninst);
  s->off.resize(ninst);
  for (int64_t i = 0; i < ninst; ++i) {
    s->off[i] = s->total;
    s->total += n[i];
    s->nmax = std::max(s->nmax, n[i]);
  }
  s->g.assign(2 * ninst, 0.0);
  s->c.assign(2 * ninst, 0.0);
  s->eta.assign(ninst, 0.0);
  s->pump_off.assign(ninst, -1);
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&s->ev0) != hipSuccess || hipEventCreate(&s->ev1) != hipSuccess)
    return bail(fail(GPE_ERR_DEVICE, "gpe_sweep_create: stream/events"));
  if (hipMalloc(&s->u, (size_t)s->total * s->csz) != hipSuccess)
    return bail(fail(GPE_ERR_NOMEM, "gpe_sweep_create: field of %lld points", (long long)s->total));
  std::vector<double> w;
  twiddles_fp64(s->nmax, w);
  int rc;
  if (s->f64) {
    rc = upload(&s->tw, w.data(), w.size() * sizeof(double));
  } else {
    std::vector<float> wf(w.begin(), w.end());
    rc = upload(&s->tw, wf.data(), wf.size() * sizeof(float));
  }
  if (rc) return bail(rc);
  // instance ids grouped by length
  for (int64_t i = 0; i < ninst; ++i)
    if (std::find(s->sizes.begin(), s->sizes.end(), n[i]) == s->sizes.end()) s->sizes.push_back(n[i]);
  std::sort(s->sizes.begin(), s->sizes.end(), std::greater<int64_t>());
  for (int64_t N : s->sizes) {
    std::vector<int32_t> ids;
    for (int64_t i = 0; i < ninst; ++i)
      if (n[i] == N) ids.push_back((int32_t)i);
    void* p = nullptr;
    rc = upload(&p, ids.data(), ids.size() * sizeof(int32_t));
    if (rc) return bail(rc);
    s->d_ids.push_back((int32_t*)p);
    s->nids.push_back((int64_t)ids.size());
  }
  *out = s;
  return GPE_OK;
}

extern "C" int gpe_sweep_set_dispersion(gpe_sweep* s, const double* D) {
  if (!s || !D) return fail(GPE_ERR_INVALID, "gpe_sweep_set_dispersion: null argument");
  HIPCHK(hipSetDevice(s->device));
  void* p = s->dD;
  int rc = upload(&p, D, (size_t)s->total * 2 * sizeof(double));
  s->dD = (double*)p;
  if (rc) return rc;
  if (!s->ed) HIPCHK(hipMalloc(&s->ed, (size_t)s->total * s->csz));
  s->ed_dt = NAN;
  return GPE_OK;
}

extern "C" int gpe_sweep_set_nonlinearity(gpe_sweep* s, const double* g, const double* c) {
  if (!s) return fail(GPE_ERR_INVALID, "gpe_sweep_set_nonlinearity: null handle");
  for (int64_t i = 0; i < 2 * s->ninst; ++i) {
    s->g[i] = g ? g[i] : 0.0;
    s->c[i] = c ? c[i] : 0.0;
  }
  return GPE_OK;
}

extern "C" int gpe_sweep_set_pump(gpe_sweep* s, const double* profile, const double* envelope, int64_t n_env) {
  if (!s) return fail(GPE_ERR_INVALID, "gpe_sweep_set_pump: null handle");
  if (n_env < 0 || (n_env > 0 && !envelope)) return fail(GPE_ERR_INVALID, "gpe_sweep_set_pump: envelope");
  HIPCHK(hipSetDevice(s->device));
  dfree(s->pump);
  s->pump_off.assign(s->ninst, -1);
  s->env.clear();
  if (!profile) return GPE_OK;
  for (int64_t i = 0; i < s->ninst; ++i) {
    bool nz = false;
    for (int64_t k = 0; k < 2 * s->n[i] && !nz; ++k) nz = profile[2 * s->off[i] + k] != 0.0;
    if (nz) s->pump_off[i] = s->off[i];
  }
  int rc;
  if (s->f64) {
    rc = upload(&s->pump, profile, (size_t)s->total * 16);
  } else {
    std::vector<float> f(2 * s->total);
    for (int64_t k = 0; k < 2 * s->total; ++k) f[k] = (float)profile[k];
    rc = upload(&s->pump, f.data(), (size_t)s->total * 8);
  }
  if (rc) return rc;
  s->env.assign(envelope, envelope + 2 * n_env);
  return GPE_OK;
}

extern "C" int gpe_sweep_set_noise(gpe_sweep* s, int32_t kind, const double* eta, uint64_t seed, uint64_t call0) {
  if (!s) return fail(GPE_ERR_INVALID, "gpe_sweep_set_noise: null handle");
  if (kind != GPE_NOISE_NONE && kind != GPE_NOISE_COMPLEX && kind != GPE_NOISE_REAL)
    return fail(GPE_ERR_INVALID, "gpe_sweep_set_noise: kind=%d", kind);
  if (kind != GPE_NOISE_NONE && !eta) return fail(GPE_ERR_INVALID, "gpe_sweep_set_noise: null eta");
  s->noise_kind = kind;
  for (int64_t i = 0; i < s->ninst; ++i) s->eta[i] = kind != GPE_NOISE_NONE ? eta[i] : 0.0;
  s->seed = seed;
  s->call = call0;
  return GPE_OK;
}

extern "C" int gpe_sweep_solve(gpe_sweep* s, const void* u0, double t0, double t1, double dt, int64_t nsaves,
                               int32_t save_start, void* result, double* ts) {
  if (!s || !u0) return fail(GPE_ERR_INVALID, "gpe_sweep_solve: null argument");
  if (!s->dD) return fail(GPE_ERR_STATE, "gpe_sweep_solve: no dispersion set");
  HIPCHK(hipSetDevice(s->device));
  double rdt;
  int64_t sps;
  int rc = gpe_resolve_timestepping(dt, t0, t1, nsaves, &rdt, &sps);
  if (rc) return rc;
  const int64_t S = sps * nsaves;
  if (!s->env.empty() && (int64_t)s->env.size() != 2 * (1 + 2 * S))
    return fail(GPE_ERR_STATE, "gpe_sweep_solve: envelope has %lld samples, this solve needs %lld",
                (long long)(s->env.size() / 2), (long long)(1 + 2 * S));
  if (!s->env.empty()) {
    const int64_t need = (int64_t)s->env.size();
    if (need > s->d_env_cap) {
      void* p = s->d_env;
      dfree(p);
      s->d_env = nullptr;
      HIPCHK(hipMalloc(&p, need * sizeof(double)));
      s->d_env = (double*)p;
      s->d_env_cap = need;
    }
    HIPCHK(hipMemcpyAsync(s->d_env, s->env.data(), need * sizeof(double), hipMemcpyHostToDevice, s->stream));
  }
  if (!(s->ed_dt == rdt)) {
    EdArgs a;
    a.D = s->dD;
    a.out = s->ed;
    a.npts = s->total;
    a.e0 = s->total;
    a.e1 = 1;
    a.form = 1;
    a.M = 1;
    a.dt = rdt;
    a.scale = 1.0;
    const unsigned blocks = (unsigned)((s->total + 255) / 256);
    if (s->f64)
      hipLaunchKernelGGL(ed_table_kernel<double>, dim3(blocks), dim3(256), 0, s->stream, a);
    else
      hipLaunchKernelGGL(ed_table_kernel<float>, dim3(blocks), dim3(256), 0, s->stream, a);
    HIPCHK(hipGetLastError());
    s->ed_dt = rdt;
  }
  const int64_t nfr = nsaves + (save_start ? 1 : 0);
  // instance records
  std::vector<SweepInst> inst(s->ninst);
  bool g_cplx = false;
  for (int64_t i = 0; i < s->ninst; ++i) {
    inst[i].off = s->off[i];
    inst[i].pump_off = s->pump_off[i];
    inst[i].res_off = s->off[i] * nfr;
    inst[i].g[0] = s->g[2 * i];
    inst[i].g[1] = s->g[2 * i + 1];
    inst[i].c[0] = s->c[2 * i];
    inst[i].c[1] = s->c[2 * i + 1];
    inst[i].eta = s->eta[i];
    g_cplx = g_cplx || inst[i].g[1] != 0.0 || inst[i].c[1] != 0.0;
  }
  if (!s->d_inst) HIPCHK(hipMalloc((void**)&s->d_inst, s->ninst * sizeof(SweepInst)));
  HIPCHK(hipMemcpyAsync(s->d_inst, inst.data(), s->ninst * sizeof(SweepInst), hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipMemcpyAsync(s->u, u0, (size_t)s->total * s->csz, hipMemcpyDefault, s->stream));
  const size_t res_bytes = (size_t)s->total * nfr * s->csz;
  void* res = nullptr;
  bool res_host = false;
  if (result) {
    if (is_device_ptr(result)) {
      res = result;
    } else {
      res_host = true;
      if (res_bytes > s->d_res_cap) {
        dfree(s->d_res);
        HIPCHK(hipMalloc(&s->d_res, res_bytes));
        s->d_res_cap = res_bytes;
      }
      res = s->d_res;
    }
  }
  SweepArgs a;
  memset(&a, 0, sizeof a);
  a.u = s->u;
  a.res = res;
  a.inst = s->d_inst;
  a.env = s->env.empty() ? nullptr : s->d_env;
  a.delta = rdt / 2;
  a.ed = s->ed;
  a.pump = s->pump;
  a.noise_kind = s->noise_kind;
  a.seed_lo = (uint32_t)s->seed;
  a.seed_hi = (uint32_t)(s->seed >> 32);
  a.g_cplx = g_cplx ? 1 : 0;
  // One launch per length, issued in turn.  Running the lengths on
  // concurrent streams was measured slower (727 against 535 us per 20-step
  // block of the 4096-instance sweep): the short lengths take SIMDs from the
  // longest one instead of filling idle ones.
  auto launch_all = [&](int64_t nsteps, int64_t frame, int64_t env_base, uint64_t call0) -> int {
    for (size_t k = 0; k < s->sizes.size(); ++k) {
      SweepArgs b = a;
      b.ids = s->d_ids[k];
      b.ninst = s->nids[k];
      b.nsteps = nsteps;
      b.frame = frame;
      b.env_base = env_base;
      b.call0 = call0;
      b.tw = s->tw;
      b.tws = (int)(s->nmax / s->sizes[k]);
      if (frame < 0) b.res = nullptr;
      HIPCHK(s->f64 ? gpe_run_sweep_f64(s->sizes[k], b, s->stream) : gpe_run_sweep_f32(s->sizes[k], b, s->stream));
    }
    return GPE_OK;
  };
  const uint64_t call0 = s->call;
  double t = t0;
  if (ts && save_start) ts[0] = t0;
  HIPCHK(hipEventRecord(s->ev0, s->stream));
  if (save_start && res) {
    rc = launch_all(0, 0, 0, call0);  // writes u0 as frame 0
    if (rc) return rc;
  }
  for (int64_t blk = 0; blk < nsaves; ++blk) {
    rc = launch_all(sps, res ? blk + (save_start ? 1 : 0) : -1, 2 * blk * sps, call0 + 2 * (uint64_t)(blk * sps));
    if (rc) return rc;
    for (int64_t k = 0; k < sps; ++k) t += rdt;
    if (ts) ts[blk + (save_start ? 1 : 0)] = t;
  }
  HIPCHK(hipEventRecord(s->ev1, s->stream));
  if (res_host) HIPCHK(hipMemcpyAsync(result, s->d_res, res_bytes, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  HIPCHK(hipEventElapsedTime(&s->last_ms, s->ev0, s->ev1));
  s->call = call0 + 2 * (uint64_t)S;
  return GPE_OK;
}

extern "C" int gpe_sweep_last_ms(gpe_sweep* s, float* ms) {
  if (!s || !ms) return fail(GPE_ERR_INVALID, "gpe_sweep_last_ms: null argument");
  *ms = s->last_ms;
  return GPE_OK;
}

// ------------------------------------------------------------ steady state
// Step from u0 until r = ||u_{n+1} - u_n|| / ||u_{n+1}|| <= tol, checking
// every `check_every` steps on two consecutive real-space fields.

extern "C" int gpe_steady_state(gpe_handle* h, const void* const* u0, double dt, double tol, int64_t check_every,
                                int64_t max_steps, int64_t* steps, double* residual, double* history,
                                int64_t max_hist, int64_t* nhist, float* ms) {
  if (!h || !u0 || !steps || !residual) return fail(GPE_ERR_INVALID, "gpe_steady_state: null argument");
  if (!(dt > 0) || !(tol > 0) || check_every < 2 || max_steps < check_every)
    return fail(GPE_ERR_INVALID, "gpe_steady_state: dt=%g tol=%g check_every=%lld max_steps=%lld", dt, tol,
                (long long)check_every, (long long)max_steps);
  if (!h->env.empty() || h->pump_kind == GPE_PUMP_SAMPLES)
    return fail(GPE_ERR_STATE, "gpe_steady_state: the pump envelope must be static");
  if (h->gmat) return fail(GPE_ERR_UNSUPPORTED, "gpe_steady_state: matrix nonlinearity not supported");
  HIPCHK(hipSetDevice(h->d.device));
  int rc = prepare_tables(h, dt);
  if (rc) return rc;
  rc = ensure_plan(h, dt, max_steps);
  if (rc) return rc;
  const size_t fb = field_bytes(h);
  SteadyBufs& sb = h->steady;
  const int nblk = 512;
  auto cleanup = [&](int code) { return code; };
  if (!sb.host) {  // first call on this handle: allocate the scratch it keeps
    bool ok = true;
    for (int c = 0; c < h->M && ok; ++c)
      ok = hipMalloc(&sb.a[c], fb) == hipSuccess && hipMalloc(&sb.b[c], fb) == hipSuccess;
    ok = ok && hipMalloc((void**)&sb.partial, 4 * nblk * sizeof(double)) == hipSuccess &&
         hipMalloc((void**)&sb.result, 2 * sizeof(double)) == hipSuccess &&
         hipHostMalloc((void**)&sb.host, 2 * sizeof(double)) == hipSuccess;
    if (!ok) {
      free_steady(sb);
      return fail(GPE_ERR_NOMEM, "gpe_steady_state: scratch buffers");
    }
  }
  for (int c = 0; c < h->M; ++c)
    if (hipMemcpyAsync(h->u[c], u0[c], fb, hipMemcpyDefault, h->stream) != hipSuccess)
      return cleanup(fail(GPE_ERR_DEVICE, "gpe_steady_state: upload"));
  auto residual_of = [&](double* r) -> int {
    ResidualArgs ra;
    memset(&ra, 0, sizeof ra);
    for (int c = 0; c < h->M; ++c) {
      ra.a[c] = sb.b[c];
      ra.b[c] = sb.a[c];
    }
    ra.npts = h->npts;
    ra.M = h->M;
    ra.partial = sb.partial;
    if (h->f64)
      hipLaunchKernelGGL(residual_partial_kernel<double>, dim3(nblk), dim3(256), 0, h->stream, ra);
    else
      hipLaunchKernelGGL(residual_partial_kernel<float>, dim3(nblk), dim3(256), 0, h->stream, ra);
    hipLaunchKernelGGL(residual_final_kernel<0>, dim3(1), dim3(64), 0, h->stream, (const double*)sb.partial, nblk,
                       sb.result);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(sb.host, sb.result, 2 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    *r = sb.host[1] > 0 ? std::sqrt(sb.host[0] / sb.host[1]) : (sb.host[0] > 0 ? INFINITY : 0.0);
    return GPE_OK;
  };
  const uint64_t call0 = h->call;
  int64_t s = 0, nh = 0;
  double r = INFINITY;
  HIPCHK(hipEventRecord(h->ev0, h->stream));
  if (h->ndim >= 2) {
    XPassArgs xa = x_args(h);
    const YPassArgs ya = h->ndim == 2 ? y_args(h) : YPassArgs{};
    // the steps of a check block before the two that write the compared
    // fields are identical without noise: one prebuilt graph replays them
    hipGraphExec_t gexec = nullptr;
    GraphExecGuard exec_guard{gexec, h->stream};
    const bool graph = h->ndim == 2 && h->noise_kind == GPE_NOISE_NONE && check_every >= 4 &&
                       h->schedule != GPE_SCHEDULE_LAUNCHES;
    if (graph) {
      XPassArgs xg = x_args(h);
      xg.flags = kInK | kEnd | kStart | kOutK;
      xg.h_end = make_half(h, dt, 1, 1, call0);
      xg.h_start = make_half(h, dt, 2, 0, call0);
      rc = build_steps_graph(h, xg, ya, check_every - 2, &gexec);
      if (rc) return cleanup(rc);
    }
    h->last_schedule = graph ? GPE_SCHEDULE_GRAPH : GPE_SCHEDULE_LAUNCHES;
    xa.flags = kStart | kOutK;
    xa.h_start = make_half(h, dt, 1, 0, call0);
    rc = run_x(h, xa) == hipSuccess ? GPE_OK : fail(GPE_ERR_DEVICE, "gpe_steady_state: launch");
    while (!rc) {
      if (gexec && (s % check_every) == 0) {
        if (hipGraphLaunch(gexec, h->stream) != hipSuccess) {
          rc = fail(GPE_ERR_DEVICE, "gpe_steady_state: graph launch");
          break;
        }
        s += check_every - 2;
      }
      ++s;
      if (run_k(h, ya) != hipSuccess) {
        rc = fail(GPE_ERR_DEVICE, "gpe_steady_state: launch");
        break;
      }
      const bool prev = (s % check_every) == check_every - 1;
      const bool check = (s % check_every) == 0;
      xa.h_end = make_half(h, dt, s, 1, call0);
      xa.h_start = make_half(h, dt, s + 1, 0, call0);
      xa.flags = kInK | kEnd | kStart | kOutK;
      for (int c = 0; c < kMaxM; ++c) xa.frame[c] = c < h->M ? (prev ? sb.a[c] : (check ? sb.b[c] : nullptr)) : nullptr;
      if (run_x(h, xa) != hipSuccess) {
        rc = fail(GPE_ERR_DEVICE, "gpe_steady_state: launch");
        break;
      }
      if (check) {
        rc = residual_of(&r);
        if (rc) break;
        if (history && nh < max_hist) history[nh] = r;
        ++nh;
        if (r <= tol || s + check_every > max_steps) break;
      }
    }
    // the real-space field after the last checked step becomes the state
    if (!rc)
      for (int c = 0; c < h->M; ++c)
        if (hipMemcpyAsync(h->u[c], sb.b[c], fb, hipMemcpyDeviceToDevice, h->stream) != hipSuccess)
          rc = fail(GPE_ERR_DEVICE, "gpe_steady_state: copy");
  } else {
    Step1DArgs a = s1_args(h, dt);
    a.env = nullptr;
    a.env_base = 0;
    while (!rc) {
      a.nsteps = check_every - 1;
      a.call0 = call0 + 2 * (uint64_t)s;
      for (int c = 0; c < kMaxM; ++c) a.frame[c] = c < h->M ? sb.a[c] : nullptr;
      if (run_1d(h, a) != hipSuccess) {
        rc = fail(GPE_ERR_DEVICE, "gpe_steady_state: launch");
        break;
      }
      a.nsteps = 1;
      a.call0 = call0 + 2 * (uint64_t)(s + check_every - 1);
      for (int c = 0; c < kMaxM; ++c) a.frame[c] = c < h->M ? sb.b[c] : nullptr;
      if (run_1d(h, a) != hipSuccess) {
        rc = fail(GPE_ERR_DEVICE, "gpe_steady_state: launch");
        break;
      }
      s += check_every;
      rc = residual_of(&r);
      if (rc) break;
      if (history && nh < max_hist) history[nh] = r;
      ++nh;
      if (r <= tol || s + check_every > max_steps) break;
    }
  }
  if (rc) return cleanup(rc);
  HIPCHK(hipEventRecord(h->ev1, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipEventElapsedTime(&h->last_ms, h->ev0, h->ev1));
  h->call = call0 + 2 * (uint64_t)s;
  h->configured = true;
  h->dt = dt;
  *steps = s;
  *residual = r;
  if (nhist) *nhist = nh;
  if (ms) *ms = h->last_ms;
  return cleanup(GPE_OK);
}
