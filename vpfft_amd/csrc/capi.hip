// capi.hip -- C ABI of the VPFFT strain-step engine: grid handle, device
// arena, upload/download and the strain-step orchestration
// (MaterialGrid::RunSingleStrainStep, Src/MaterialGrid.cpp:897-1084).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <utility>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/vpfft.h"
#include "capi_internal.h"
#include "device_common.h"
#include "fft_green.h"
#include "voxel_kernels.h"

namespace vpfft {

static thread_local std::string g_err;

int set_error(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIP_TRY(expr)                                                                            \
  do {                                                                                           \
    hipError_t e_ = (expr);                                                                      \
    if (e_ != hipSuccess) return set_error(VPFFT_E_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

__global__ void k_fill(double* p, double v, long n) {
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) p[t] = v;
}

__global__ void k_schmid_aos(const double* __restrict__ R, long n, long offset, long count, double* __restrict__ out) {
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < count; t += (long)gridDim.x * blockDim.x) {
    const long v = offset + t;
    double r[9], m[kSlips][5];
    for (int e = 0; e < 9; e++) r[e] = R[e * n + v];
    schmid_from_R(r, m);
    for (int k = 0; k < kSlips; k++)
      for (int i = 0; i < 5; i++) out[60 * t + 5 * k + i] = m[k][i];
  }
}

// the material constants the constitutive law reads, passed by value
struct RateLaw {
  double gdot0[kSlips];
  int nexp[kSlips];
};

// Solvers::ApplyConstitutiveEquations (Solvers.cpp:255-274) at every voxel of
// [offset, offset + count): strain rate from the stress, orientation and
// CRSS fields, AoS (5 per voxel) into out.
__global__ void k_strain_rate_aos(const double* __restrict__ R, const double* __restrict__ crss,
                                  const double* __restrict__ sigma, long n, long offset, long count, RateLaw law,
                                  double* __restrict__ out) {
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < count; t += (long)gridDim.x * blockDim.x) {
    const long v = offset + t;
    double r[9], m[kSlips][5], s[5], e[5] = {0, 0, 0, 0, 0};
    for (int q = 0; q < 9; q++) r[q] = R[q * n + v];
    for (int i = 0; i < 5; i++) s[i] = sigma[i * n + v];
    schmid_from_R(r, m);
    for (int k = 0; k < kSlips; k++) {
      const double rss = dot5(m[k], s) / crss[k * n + v];
      const double g = law.nexp[k] > 0 ? law.gdot0[k] * ipow(fabs(rss), law.nexp[k]) : law.gdot0[k];
      for (int i = 0; i < 5; i++) e[i] += m[k][i] * g * sgn(rss);
    }
    for (int i = 0; i < 5; i++) out[5 * t + i] = e[i];
  }
}

static int fill_blocks(long n) {
  long b = (n + 255) / 256;
  return (int)std::max(1L, std::min(b, 8192L));
}

enum KernelId {
  K_PROLOGUE = 0, K_L0_FINALIZE, K_Z_R2C, K_Y_FWD, K_X_GREEN, K_Y_INV, K_Z_C2R, K_NR, K_ITER_FINALIZE,
  K_EPILOGUE, K_ZNR, kNumKernels
};
static const char* kKernelNames[kNumKernels] = {"prologue", "L0_finalize", "fft_z_r2c", "fft_y_fwd",
                                                "fft_x_green", "fft_y_inv", "fft_z_c2r", "nr_update",
                                                "iter_finalize", "epilogue", "z_nr_fused"};

struct Grid {
  unsigned long serial = 0;  // unique per created grid (keys of captured batch graphs)
  int nx = 0, ny = 0, nz = 0, nzh = 0, nzp = 0;
  long n = 0, nc = 0;
  int prec = 0;
  double lx = 1, ly = 1, lz = 1;
  vpfft_material mat{};
  hipStream_t st = nullptr;
  double *R = nullptr, *crss = nullptr, *acc = nullptr, *sigma = nullptr, *lambda = nullptr;
  void *phi = nullptr, *dtil = nullptr, *spec = nullptr;  // sized for FP64
  // Nz = 256 or 512: the loop runs the fused inverse-z / Newton / forward-z kernel,
  // alternating the half spectrum between spec and spec2 (VPFFT_FUSE_Z=0
  // turns it off; the separate kernels then take the fused kernel's runs of
  // 64 voxels, so both forms give the same bits)
  void* spec2 = nullptr;
  bool fuse_z = false;
  void* tw64[3] = {nullptr, nullptr, nullptr};
  void* tw32[3] = {nullptr, nullptr, nullptr};
  AxisPlan plan[3];
  FftGeom geom{};
  StepConst* dC = nullptr;
  GreenConst<double>* dG64 = nullptr;
  GreenConst<float>* dG32 = nullptr;
  double* partial = nullptr;
  IterRecord* drec = nullptr;
  int rec_cap = 0;
  PrologueRecord* dprol = nullptr;
  HealthRecord* dhealth = nullptr;
  HealthRecord* hhealth = nullptr;
  StepConst* hC = nullptr;
  IterRecord* hrec = nullptr;
  PrologueRecord* hprol = nullptr;
  double* staging = nullptr;
  long staging_elems = 0;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  bool uploaded = false;
  int nexp_uniform = 0;  // the exponent every slip shares (Newton kernel specialisation), else 0
  int fftB[2][3] = {{0, 0, 0}, {0, 0, 0}};  // FFT block shape per precision and pass (z, y, x); 0 = default
  // every voxel's 12 CRSS values are equal (uploaded uniform and kept so by a
  // hardening law whose matrix rows and Voce constants are slip-independent)
  bool crss_uniform = false;
  bool hardening_uniform = false;
  bool gdot0_uniform = false;  // every slip system has the same reference slip rate
  // the Newton kernel's uniform variant: one CRSS and one reference slip rate
  // for all slip systems of a voxel
  bool uniform_slips() const { return crss_uniform && gdot0_uniform; }
  // fixed-point loop: device loop state and a captured graph of several
  // iterations (VPFFT_GRAPH_ITERS, 0 = launch every kernel from the host)
  LoopState* dls = nullptr;
  LoopState* hls = nullptr;
  int graph_iters = 4;
  hipGraphExec_t gexec = nullptr;
  struct GraphKey {
    int prec, max_iter, nexp, uniform;
    double tol;
    bool operator==(const GraphKey& o) const {
      return prec == o.prec && max_iter == o.max_iter && nexp == o.nexp && uniform == o.uniform && tol == o.tol;
    }
  } gkey{-1, 0, 0, 0, 0.0};
  int phi_prec = 0, dtil_prec = 0;  // precision the field buffers currently hold
  // device copy of the state a strain step changes (vpfft_grid_save_state)
  double* saved = nullptr;
  double saved_len[3] = {1, 1, 1};
  bool saved_crss_uniform = false, has_saved = false;
  // per-kernel timing with HIP events on the grid's stream
  bool profile = false;
  hipEvent_t kev[kNumKernels][2] = {};
  bool kev_used[kNumKernels] = {};
  double kms[kNumKernels] = {};
  long kcount[kNumKernels] = {};
  std::vector<void*> dev_allocs;
  std::vector<void*> host_allocs;
};

static int dalloc(Grid* g, void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return set_error(VPFFT_E_NOMEM, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
  }
  g->dev_allocs.push_back(*p);
  return VPFFT_OK;
}
static int halloc(Grid* g, void** p, size_t bytes) {
  hipError_t e = hipHostMalloc(p, bytes, hipHostMallocDefault);
  if (e != hipSuccess) {
    *p = nullptr;
    return set_error(VPFFT_E_NOMEM, "hipHostMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
  }
  g->host_allocs.push_back(*p);
  return VPFFT_OK;
}

static void grid_free(Grid* g) {
  if (!g) return;
  if (g->st) (void)hipStreamSynchronize(g->st);
  for (void* p : g->dev_allocs) (void)hipFree(p);
  for (void* p : g->host_allocs) (void)hipHostFree(p);
  for (auto& e : g->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& pr : g->kev)
    for (auto& e : pr)
      if (e) (void)hipEventDestroy(e);
  if (g->gexec) (void)hipGraphExecDestroy(g->gexec);
  if (g->st) (void)hipStreamDestroy(g->st);
  delete g;
}

static int make_twiddles(Grid* g, int axis, int n) {
  std::vector<double2> t64(n);
  std::vector<float2> t32(n);
  for (int t = 0; t < n; t++) {
    const long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)t / (long double)n;
    t64[t].x = (double)cosl(a);
    t64[t].y = (double)sinl(a);
    t32[t].x = (float)t64[t].x;
    t32[t].y = (float)t64[t].y;
  }
  int rc;
  if ((rc = dalloc(g, &g->tw64[axis], sizeof(double2) * n))) return rc;
  if ((rc = dalloc(g, &g->tw32[axis], sizeof(float2) * n))) return rc;
  HIP_TRY(hipMemcpy(g->tw64[axis], t64.data(), sizeof(double2) * n, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(g->tw32[axis], t32.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
  return VPFFT_OK;
}

static int ensure_records(Grid* g, int cap) {
  if (cap <= g->rec_cap) return VPFFT_OK;
  int newcap = std::max(cap, 2 * g->rec_cap);
  void *d = nullptr, *h = nullptr;
  int rc;
  if ((rc = dalloc(g, &d, sizeof(IterRecord) * newcap))) return rc;
  if ((rc = halloc(g, &h, sizeof(IterRecord) * newcap))) return rc;
  g->drec = (IterRecord*)d;
  g->hrec = (IterRecord*)h;
  g->rec_cap = newcap;
  return VPFFT_OK;
}

// basis diagonal of EigenRep::ToMatrixRep (LinearAlgebra.h:352-360)
static void edot_diag(const double e[5], double d[3]) {
  const double s2 = 1.0 / std::sqrt(2.0), s6 = 1.0 / std::sqrt(6.0);
  d[0] = 0.0 + (-1.0 * s2) * e[0] + (-1.0 * s6) * e[1];
  d[1] = 0.0 + (1.0 * s2) * e[0] + (-1.0 * s6) * e[1];
  d[2] = 0.0 + (0.0 * s2) * e[0] + (2.0 * s6) * e[1];
}

static void fill_host_const(Grid* g, const double edot[5], double dt) {
  StepConst& C = *g->hC;
  std::memset(&C, 0, sizeof C);
  for (int i = 0; i < 5; i++) C.edot[i] = edot[i];
  int nmin = g->mat.rate_sensitivity[0];
  C.uniform_n = 1;
  for (int k = 0; k < kSlips; k++) {
    C.gdot0[k] = g->mat.gamma_dot_base[k];
    C.nexp[k] = g->mat.rate_sensitivity[k];
    if (C.nexp[k] != C.nexp[0]) C.uniform_n = 0;
    nmin = std::min(nmin, C.nexp[k]);
    C.tau0[k] = g->mat.tau0[k];
    C.tau1[k] = g->mat.tau1[k];
    C.theta0[k] = g->mat.theta0[k];
    C.theta1[k] = g->mat.theta1[k];
  }
  for (int e = 0; e < kSlips * kSlips; e++) C.hmat[e] = g->mat.hardening[e];
  // MaterialGrid.cpp:162-166
  double ee = 0;
  for (int i = 0; i < 5; i++) ee += edot[i] * edot[i];
  double imposed = 2.0 * std::pow(ee / C.gdot0[0], 1.0 / (double)(float)nmin);
  C.imposed_max_rss = std::max(2.0, imposed);
  C.inv_n = 1.0 / (double)g->n;
  C.dt = dt;
  C.kscale[0] = 1.0 / ((double)g->nx * g->lx);
  C.kscale[1] = 1.0 / ((double)g->ny * g->ly);
  C.kscale[2] = 1.0 / ((double)g->nz * g->lz);
  C.nx = g->nx;
  C.ny = g->ny;
  C.nz = g->nz;
  C.nzh = g->nzh;
}

// max(err_stress, err_strain_rate) of the last record over the same five
// records earlier
static double err_ratio_5(const IterRecord* rec, int it) {
  if (it < 6) return 1.0;
  const double a = std::max(rec[it - 1].err_stress, rec[it - 1].err_strain_rate);
  const double b = std::max(rec[it - 6].err_stress, rec[it - 6].err_strain_rate);
  return b > 0 ? a / b : 1.0;
}

// Bracket one kernel launch with events when profiling is on.
static inline void kbegin(Grid* g, int k) {
  if (g->profile) {
    (void)hipEventRecord(g->kev[k][0], g->st);
    g->kev_used[k] = true;
  }
}
static inline void kend(Grid* g, int k) {
  if (g->profile) (void)hipEventRecord(g->kev[k][1], g->st);
}
// After a stream synchronisation: fold the recorded intervals into the totals.
static void kcollect(Grid* g) {
  if (!g->profile) return;
  for (int k = 0; k < kNumKernels; k++) {
    if (!g->kev_used[k]) continue;
    float ms = 0;
    if (hipEventElapsedTime(&ms, g->kev[k][0], g->kev[k][1]) == hipSuccess) {
      g->kms[k] += ms;
      g->kcount[k] += 1;
    }
    g->kev_used[k] = false;
  }
}

static int run_fft_green(Grid* g, int prec, const int* done) {
  void* const* tw = prec == 0 ? g->tw64 : g->tw32;
  const void* gc = prec == 0 ? (const void*)g->dG64 : (const void*)g->dG32;
  kbegin(g, K_Z_R2C);
  HIP_TRY(launch_z_r2c(prec, g->phi, g->spec, g->geom, g->plan[2], tw[2], g->fftB[prec][0], done, g->st));
  kend(g, K_Z_R2C);
  kbegin(g, K_Y_FWD);
  HIP_TRY(launch_y(prec, g->spec, g->geom, g->plan[1], tw[1], -1, g->fftB[prec][1], done, g->st));
  kend(g, K_Y_FWD);
  kbegin(g, K_X_GREEN);
  HIP_TRY(launch_x_green(prec, g->spec, g->geom, g->plan[0], tw[0], gc, g->fftB[prec][2], done, g->st));
  kend(g, K_X_GREEN);
  kbegin(g, K_Y_INV);
  HIP_TRY(launch_y(prec, g->spec, g->geom, g->plan[1], tw[1], +1, g->fftB[prec][1], done, g->st));
  kend(g, K_Y_INV);
  kbegin(g, K_Z_C2R);
  HIP_TRY(launch_z_c2r(prec, g->spec, g->dtil, g->geom, g->plan[2], tw[2], 1.0 / (double)g->n, g->fftB[prec][0], done,
                       g->st));
  kend(g, K_Z_C2R);
  g->dtil_prec = prec;
  return VPFFT_OK;
}

// host AoS (count x width) -> device SoA (width x n), through the staging buffer
static int upload_aos(Grid* g, const double* host, double* dev, int width) {
  const long chunk = g->staging_elems / width;
  for (long off = 0; off < g->n; off += chunk) {
    const long cnt = std::min(chunk, g->n - off);
    HIP_TRY(hipMemcpyAsync(g->staging, host + off * width, sizeof(double) * cnt * width, hipMemcpyHostToDevice, g->st));
    HIP_TRY(launch_aos_to_soa(g->staging, dev, cnt, width, g->n, off, g->st));
  }
  HIP_TRY(hipStreamSynchronize(g->st));
  return VPFFT_OK;
}
static int download_aos(Grid* g, const double* dev, double* host, int width) {
  const long chunk = g->staging_elems / width;
  for (long off = 0; off < g->n; off += chunk) {
    const long cnt = std::min(chunk, g->n - off);
    HIP_TRY(launch_soa_to_aos(dev, g->staging, cnt, width, g->n, off, g->st));
    HIP_TRY(hipMemcpyAsync(host + off * width, g->staging, sizeof(double) * cnt * width, hipMemcpyDeviceToHost, g->st));
    HIP_TRY(hipStreamSynchronize(g->st));
  }
  return VPFFT_OK;
}
// 5 x N field in precision prec <-> host doubles
static int upload_field5(Grid* g, const double* host, void* dev, int prec) {
  if (prec == 0) {
    HIP_TRY(hipMemcpyAsync(dev, host, sizeof(double) * 5 * g->n, hipMemcpyHostToDevice, g->st));
  } else {
    const long chunk = g->staging_elems;
    for (long off = 0; off < 5 * g->n; off += chunk) {
      const long cnt = std::min(chunk, 5 * g->n - off);
      HIP_TRY(hipMemcpyAsync(g->staging, host + off, sizeof(double) * cnt, hipMemcpyHostToDevice, g->st));
      HIP_TRY(launch_convert(1, (float*)dev + off, 0, g->staging, cnt, g->st));
    }
  }
  HIP_TRY(hipStreamSynchronize(g->st));
  return VPFFT_OK;
}
static int download_field5(Grid* g, const void* dev, double* host, int prec) {
  if (prec == 0) {
    HIP_TRY(hipMemcpyAsync(host, dev, sizeof(double) * 5 * g->n, hipMemcpyDeviceToHost, g->st));
    HIP_TRY(hipStreamSynchronize(g->st));
  } else {
    const long chunk = g->staging_elems;
    for (long off = 0; off < 5 * g->n; off += chunk) {
      const long cnt = std::min(chunk, 5 * g->n - off);
      HIP_TRY(launch_convert(0, g->staging, 1, (const float*)dev + off, cnt, g->st));
      HIP_TRY(hipMemcpyAsync(host + off, g->staging, sizeof(double) * cnt, hipMemcpyDeviceToHost, g->st));
      HIP_TRY(hipStreamSynchronize(g->st));
    }
  }
  return VPFFT_OK;
}

// true when every voxel's 12 CRSS values are bitwise equal
static bool crss_rows_uniform(const double* crss, long n) {
  for (long v = 0; v < n; v++)
    for (int k = 1; k < kSlips; k++)
      if (crss[12 * v + k] != crss[12 * v]) return false;
  return true;
}

static int check_material(const vpfft_material* m) {
  for (int k = 0; k < kSlips; k++) {
    if (!(m->gamma_dot_base[k] > 0) || !std::isfinite(m->gamma_dot_base[k]))
      return set_error(VPFFT_E_INVALID, "gamma_dot_base[%d] must be positive and finite", k);
    if (m->rate_sensitivity[k] < 1 || m->rate_sensitivity[k] > 200)
      return set_error(VPFFT_E_INVALID, "rate_sensitivity[%d]=%d out of range [1,200]", k, m->rate_sensitivity[k]);
  }
  return VPFFT_OK;
}

// Device bytes grid_create allocates for these dimensions (records at their
// initial capacity; the saved-state copy not included).
static size_t grid_bytes(const vpfft_grid_desc* d) {
  const long n = (long)d->nx * d->ny * d->nz;
  const long nc = (long)d->nx * d->ny * padded_half(d->nz);
  const bool fused = d->nz == 256 || d->nz == 512;
  size_t b = sizeof(double) * 42 * (size_t)n + sizeof(double2) * 5 * (size_t)nc * (fused ? 2 : 1);
  b += sizeof(StepConst) + sizeof(GreenConst<double>) + sizeof(GreenConst<float>) + sizeof(PrologueRecord) +
       sizeof(HealthRecord) + sizeof(LoopState) + sizeof(IterRecord) * 128;
  b += sizeof(double) * (size_t)std::max<long>(2L * 4096 * 18, nr_partial_rows(n) * kPartialStride);
  b += sizeof(double) * (size_t)std::min<long>(60L * n, 8L << 20);
  b += (sizeof(double2) + sizeof(float2)) * (size_t)(d->nx + d->ny + d->nz);
  return b;
}

static int grid_create(const vpfft_grid_desc* d, const vpfft_material* mat, Grid** out) {
  if (!d || !mat || !out) return set_error(VPFFT_E_INVALID, "null argument");
  *out = nullptr;
  if (d->nx < 2 || d->ny < 2 || d->nz < 2 || d->nx > 4096 || d->ny > 4096 || d->nz > 4096)
    return set_error(VPFFT_E_INVALID, "grid edges must be in [2, 4096]");
  if (d->precision != VPFFT_FP64 && d->precision != VPFFT_MIXED)
    return set_error(VPFFT_E_INVALID, "precision must be VPFFT_FP64 or VPFFT_MIXED");
  int rc = check_material(mat);
  if (rc) return rc;
  static unsigned long next_serial = 0;
  Grid* g = new Grid();
  g->serial = ++next_serial;
  g->nx = d->nx;
  g->ny = d->ny;
  g->nz = d->nz;
  g->nzh = d->nz / 2 + 1;
  g->nzp = padded_half(d->nz);
  g->n = (long)d->nx * d->ny * d->nz;
  g->nc = (long)d->nx * d->ny * g->nzp;
  g->prec = d->precision;
  g->mat = *mat;
  if (const char* k = std::getenv("VPFFT_GRAPH_ITERS")) g->graph_iters = std::max(0, std::min(64, std::atoi(k)));
  g->gdot0_uniform = true;
  for (int k = 0; k < kSlips; k++)
    if (mat->gamma_dot_base[k] != mat->gamma_dot_base[0]) g->gdot0_uniform = false;
  g->hardening_uniform = true;
  for (int i = 0; i < kSlips; i++) {
    if (mat->tau0[i] != mat->tau0[0] || mat->tau1[i] != mat->tau1[0] || mat->theta0[i] != mat->theta0[0] ||
        mat->theta1[i] != mat->theta1[0])
      g->hardening_uniform = false;
    for (int j = 0; j < kSlips; j++)
      if (mat->hardening[i * kSlips + j] != mat->hardening[j]) g->hardening_uniform = false;
  }
  // Newton kernels are specialised for the common shared exponents
  g->nexp_uniform = mat->rate_sensitivity[0];
  for (int k = 0; k < kSlips; k++)
    if (mat->rate_sensitivity[k] != mat->rate_sensitivity[0]) g->nexp_uniform = 0;
  if (g->nexp_uniform != 5 && g->nexp_uniform != 10 && g->nexp_uniform != 20) g->nexp_uniform = 0;
  const int dims[3] = {g->nx, g->ny, g->nz};
  for (int a = 0; a < 3; a++)
    if (!make_axis_plan(dims[a], &g->plan[a])) {
      grid_free(g);
      return set_error(VPFFT_E_UNSUPPORTED, "grid edge %d has a prime factor other than 2, 3, 5, 7", dims[a]);
    }
  // LDS limits of the FFT passes (2 ping-pong buffers of one pencil batch in FP64)
  if ((long)4 * g->nz * 16 * 2 > 160 * 1024 || (long)g->ny * 16 * 2 > 160 * 1024 ||
      (long)g->nx * 16 * 8 * 2 > 160 * 1024) {
    grid_free(g);
    return set_error(VPFFT_E_UNSUPPORTED, "grid edge too long for the LDS-resident FFT");
  }
  g->geom.n = g->n;
  g->geom.nc = g->nc;
  g->geom.nrows = (long)g->nx * g->ny;
  g->geom.nx = g->nx;
  g->geom.ny = g->ny;
  g->geom.nz = g->nz;
  g->geom.nzh = g->nzh;
  g->geom.nzp = g->nzp;
  hipError_t e = hipStreamCreateWithFlags(&g->st, hipStreamNonBlocking);
  if (e != hipSuccess) {
    grid_free(g);
    return set_error(VPFFT_E_DEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
  }
  for (auto& ev : g->ev)
    if (hipEventCreate(&ev) != hipSuccess) {
      grid_free(g);
      return set_error(VPFFT_E_DEVICE, "hipEventCreate failed");
    }
  for (auto& pr : g->kev)
    for (auto& ev : pr)
      if (hipEventCreate(&ev) != hipSuccess) {
        grid_free(g);
        return set_error(VPFFT_E_DEVICE, "hipEventCreate failed");
      }
  const long n = g->n;
  void* p;
#define ALLOC(field, bytes)                          \
  if ((rc = dalloc(g, &p, (bytes)))) {               \
    grid_free(g);                                    \
    return rc;                                       \
  }                                                  \
  field = (decltype(field))p;
  ALLOC(g->R, sizeof(double) * 9 * n);
  ALLOC(g->crss, sizeof(double) * 12 * n);
  ALLOC(g->acc, sizeof(double) * n);
  ALLOC(g->sigma, sizeof(double) * 5 * n);
  ALLOC(g->lambda, sizeof(double) * 5 * n);
  ALLOC(g->phi, sizeof(double) * 5 * n);
  ALLOC(g->dtil, sizeof(double) * 5 * n);
  ALLOC(g->spec, sizeof(double2) * 5 * g->nc);
  if (g->nz == 256 || g->nz == 512) {
    ALLOC(g->spec2, sizeof(double2) * 5 * g->nc);
    const char* f = std::getenv("VPFFT_FUSE_Z");
    g->fuse_z = !(f && std::atoi(f) == 0);
  }
  ALLOC(g->dC, sizeof(StepConst));
  ALLOC(g->dG64, sizeof(GreenConst<double>));
  ALLOC(g->dG32, sizeof(GreenConst<float>));
  ALLOC(g->partial, sizeof(double) * std::max<long>(2L * 4096 * 18, nr_partial_rows(n) * kPartialStride));
  ALLOC(g->dprol, sizeof(PrologueRecord));
  ALLOC(g->dhealth, sizeof(HealthRecord));
  ALLOC(g->dls, sizeof(LoopState));
  g->staging_elems = std::min<long>(60L * n, 8L << 20);
  ALLOC(g->staging, sizeof(double) * g->staging_elems);
#undef ALLOC
  if ((rc = halloc(g, &p, sizeof(StepConst)))) { grid_free(g); return rc; }
  g->hC = (StepConst*)p;
  if ((rc = halloc(g, &p, sizeof(PrologueRecord)))) { grid_free(g); return rc; }
  g->hprol = (PrologueRecord*)p;
  if ((rc = halloc(g, &p, sizeof(LoopState)))) { grid_free(g); return rc; }
  g->hls = (LoopState*)p;
  if ((rc = halloc(g, &p, sizeof(HealthRecord)))) { grid_free(g); return rc; }
  g->hhealth = (HealthRecord*)p;
  std::memset(g->hhealth, 0, sizeof(HealthRecord));
  if ((rc = ensure_records(g, 128))) { grid_free(g); return rc; }
  for (int a = 0; a < 3; a++)
    if ((rc = make_twiddles(g, a, dims[a]))) { grid_free(g); return rc; }
  *out = g;
  return VPFFT_OK;
}

// One fixed-point iteration: FFT + Green + inverse FFT, the Newton kernel,
// and the device-side convergence test.  Every kernel returns at entry once
// the loop state says done, so the same sequence can be captured in a graph.
// The fused loop applies to this step: the grid is eligible and its z pass
// uses the 256-point register form (the default block shape).
static bool fused_step(const Grid* g, int prec) { return g->fuse_z && g->fftB[prec][0] == 0; }
// The Newton waves' voxels for the grids that can run the fused loop, fused
// or not, so both forms sum the same way: a z row of 256 (one partial row per
// wave) at Nz = 256, runs of 64 consecutive voxels at Nz = 512; else 4 x 16
// tiles (launch_nr's row_runs)
static int row_runs(const Grid* g) { return g->nz == 256 ? 4 : (g->nz == 512 ? 1 : 0); }

// Iteration k of a fused loop: y, x + Green, y inverse on the half spectrum
// the previous iteration left (spec for even k, spec2 for odd k), then the
// fused kernel writes the next half spectrum to the other buffer.
static void fused_buffers(Grid* g, int k, void** in, void** out) {
  *in = (k & 1) ? g->spec2 : g->spec;
  *out = (k & 1) ? g->spec : g->spec2;
}

static int enqueue_iteration(Grid* g, int prec, double tol, int max_iter, int k) {
  const int* done = &g->dls->done;
  int rc;
  if (fused_step(g, prec)) {
    void* const* tw = prec == 0 ? g->tw64 : g->tw32;
    const void* gc = prec == 0 ? (const void*)g->dG64 : (const void*)g->dG32;
    void *sin, *sout;
    fused_buffers(g, k, &sin, &sout);
    kbegin(g, K_Y_FWD);
    HIP_TRY(launch_y(prec, sin, g->geom, g->plan[1], tw[1], -1, g->fftB[prec][1], done, g->st));
    kend(g, K_Y_FWD);
    kbegin(g, K_X_GREEN);
    HIP_TRY(launch_x_green(prec, sin, g->geom, g->plan[0], tw[0], gc, g->fftB[prec][2], done, g->st));
    kend(g, K_X_GREEN);
    kbegin(g, K_Y_INV);
    HIP_TRY(launch_y(prec, sin, g->geom, g->plan[1], tw[1], +1, g->fftB[prec][1], done, g->st));
    kend(g, K_Y_INV);
    kbegin(g, K_ZNR);
    HIP_TRY(launch_zfused(prec, g->nexp_uniform, g->uniform_slips(), g->n, sin, sout, g->geom, tw[2], 1.0 / (double)g->n,
                          g->sigma, g->lambda, g->R, g->crss, g->dC, g->partial, g->dls, g->dhealth, g->st));
    kend(g, K_ZNR);
  } else {
    if ((rc = run_fft_green(g, prec, done))) return rc;
    kbegin(g, K_NR);
    HIP_TRY(launch_nr(prec, g->nexp_uniform, g->uniform_slips(), g->n, g->dtil, g->phi, g->sigma, g->lambda, g->R,
                      g->crss, g->dC, g->partial, g->dls, g->dhealth, row_runs(g), g->st));
    kend(g, K_NR);
  }
  g->phi_prec = prec;
  kbegin(g, K_ITER_FINALIZE);
  HIP_TRY(launch_iter_finalize(g->n, row_runs(g), g->partial, g->dC, tol, max_iter, g->dls, g->drec, g->st));
  kend(g, K_ITER_FINALIZE);
  return VPFFT_OK;
}

// Cholesky factorisation of the symmetric part of a column-major 5x5:
// true when every pivot is positive (the matrix is positive definite).
static bool spd5(const double A[25]) {
  double L[5][5] = {};
  for (int j = 0; j < 5; j++) {
    double d = A[j * 5 + j];
    for (int k = 0; k < j; k++) d -= L[j][k] * L[j][k];
    if (!(d > 0) || !std::isfinite(d)) return false;
    L[j][j] = std::sqrt(d);
    for (int i = j + 1; i < 5; i++) {
      double t = 0.5 * (A[j * 5 + i] + A[i * 5 + j]);
      for (int k = 0; k < j; k++) t -= L[i][k] * L[j][k];
      L[i][j] = t / L[j][j];
    }
  }
  return true;
}

// The status and message of a health record that has its flag set.
static int health_error(const Grid* g, const HealthRecord& h, const char* where) {
  const int code = (h.kernel == HK_PROLOGUE && h.field == HF_SECANT_INVERSE && std::isfinite(h.value) && h.value == 0.0)
                       ? VPFFT_E_DEGENERATE
                       : VPFFT_E_NONFINITE;
  if (h.voxel >= 0) {
    const long yz = (long)g->ny * g->nz;
    const long x = h.voxel / yz, y = (h.voxel / g->nz) % g->ny, z = h.voxel % g->nz;
    char it[32] = "";
    if (h.iter > 0) std::snprintf(it, sizeof it, ", iteration %d", h.iter);
    return set_error(code, "%s: %s %s at voxel %ld (x=%ld, y=%ld, z=%ld)%s (value %g)", where,
                     health_kernel_name(h.kernel), health_field_name(h.field), h.voxel, x, y, z, it, h.value);
  }
  return set_error(code, "%s: %s %s (value %g)", where, health_kernel_name(h.kernel), health_field_name(h.field),
                   h.value);
}

static int strain_step(Grid* g, const double edot[5], int max_iter, double dt, double tol, double avg[5],
                       vpfft_step_info* info) {
  if (!g->uploaded) return set_error(VPFFT_E_STATE, "run_strain_step before upload");
  if (max_iter < 1) return set_error(VPFFT_E_INVALID, "max_iter must be >= 1");
  if (!(tol >= 0) || !std::isfinite(dt)) return set_error(VPFFT_E_INVALID, "bad tolerance or time step");
  double ee = 0;
  for (int i = 0; i < 5; i++) {
    if (!std::isfinite(edot[i])) return set_error(VPFFT_E_NONFINITE, "non-finite strain rate");
    ee += edot[i] * edot[i];
  }
  if (!(ee > 0)) return set_error(VPFFT_E_DEGENERATE, "zero strain rate: max resolved shear stress < 1e-10");
  int rc = ensure_records(g, max_iter);
  if (rc) return rc;
  const int prec = g->prec;
  const long n = g->n;
  const int nb = point_blocks(n);
  fill_host_const(g, edot, dt);
  HIP_TRY(hipMemcpyAsync(g->dC, g->hC, sizeof(StepConst), hipMemcpyHostToDevice, g->st));
  HIP_TRY(launch_health_reset(g->dhealth, g->st));
  HIP_TRY(hipEventRecord(g->ev[0], g->st));
  // prologue: MaterialGrid.cpp:903-913
  kbegin(g, K_PROLOGUE);
  HIP_TRY(launch_prologue(prec, g->crss_uniform, n, g->R, g->crss, g->dC, g->phi, g->sigma, g->lambda, g->partial,
                          g->dhealth, g->st));
  kend(g, K_PROLOGUE);
  g->phi_prec = prec;
  kbegin(g, K_L0_FINALIZE);
  HIP_TRY(launch_L0_finalize(n, nb, g->partial, g->dC, g->dG64, g->dG32, g->dprol, g->dhealth, g->st));
  kend(g, K_L0_FINALIZE);
  HIP_TRY(hipMemcpyAsync(g->hprol, g->dprol, sizeof(PrologueRecord), hipMemcpyDeviceToHost, g->st));
  HIP_TRY(hipMemcpyAsync(g->hhealth, g->dhealth, sizeof(HealthRecord), hipMemcpyDeviceToHost, g->st));
  HIP_TRY(hipEventRecord(g->ev[1], g->st));
  HIP_TRY(hipStreamSynchronize(g->st));
  kcollect(g);
  if (g->hprol->degenerate > 0)
    return set_error(VPFFT_E_DEGENERATE, "%.0f voxels with max resolved shear stress < 1e-10 (initialization)",
                     g->hprol->degenerate);
  if (g->hhealth->flag) return health_error(g, *g->hhealth, "initialization");
  // the reference medium is an average of positive definite inverses: a
  // medium that is not positive definite means the prologue went wrong
  if (!spd5(g->hprol->L0) || !spd5(g->hprol->L0inv))
    return set_error(VPFFT_E_NONFINITE, "reference medium L0 (or its inverse) is not symmetric positive definite");
  // fixed-point loop: MaterialGrid.cpp:930-1068
  HIP_TRY(launch_loop_reset(g->dls, g->st));
  const bool fused = fused_step(g, prec);
  if (fused) {  // the fused loop starts from the half spectrum of the prologue's phi
    void* const* tw = prec == 0 ? g->tw64 : g->tw32;
    kbegin(g, K_Z_R2C);
    HIP_TRY(launch_z_r2c(prec, g->phi, g->spec, g->geom, g->plan[2], tw[2], 0, nullptr, g->st));
    kend(g, K_Z_R2C);
  }
  int it = 0;
  double nr_total = 0, wave_total = 0;
  const bool use_graph = g->graph_iters > 0 && !g->profile;
  // a fused loop alternates two spectrum buffers, so a replayed graph holds an
  // even number of iterations
  const int graph_iters = fused ? (g->graph_iters + 1) / 2 * 2 : g->graph_iters;
  if (use_graph) {
    const Grid::GraphKey key{prec, max_iter, g->nexp_uniform, (g->uniform_slips() ? 1 : 0) | (fused ? 2 : 0), tol};
    if (!g->gexec || !(g->gkey == key)) {
      if (g->gexec) {
        HIP_TRY(hipGraphExecDestroy(g->gexec));
        g->gexec = nullptr;
      }
      hipGraph_t graph;
      HIP_TRY(hipStreamBeginCapture(g->st, hipStreamCaptureModeThreadLocal));
      int crc = VPFFT_OK;
      for (int k = 0; k < graph_iters && crc == VPFFT_OK; k++) crc = enqueue_iteration(g, prec, tol, max_iter, k);
      hipError_t ce = hipStreamEndCapture(g->st, &graph);
      if (crc) return crc;
      HIP_TRY(ce);
      hipError_t ie = hipGraphInstantiate(&g->gexec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      HIP_TRY(ie);
      g->gkey = key;
    }
    for (;;) {
      HIP_TRY(hipGraphLaunch(g->gexec, g->st));
      HIP_TRY(hipMemcpyAsync(g->hls, g->dls, sizeof(LoopState), hipMemcpyDeviceToHost, g->st));
      HIP_TRY(hipStreamSynchronize(g->st));
      if (g->hls->done) break;
    }
    it = g->hls->iter;
    HIP_TRY(hipMemcpyAsync(g->hrec, g->drec, sizeof(IterRecord) * it, hipMemcpyDeviceToHost, g->st));
    HIP_TRY(hipStreamSynchronize(g->st));
    for (int i = 0; i < it; i++) {
      nr_total += g->hrec[i].nr_steps;
      wave_total += g->hrec[i].nr_wave_steps;
      if (g->hrec[i].degenerate > 0)
        return set_error(VPFFT_E_DEGENERATE, "%.0f voxels with max resolved shear stress < 1e-10 (iteration %d)",
                         g->hrec[i].degenerate, i + 1);
    }
  } else {
    for (;;) {
      if ((rc = enqueue_iteration(g, prec, tol, max_iter, it))) return rc;
      HIP_TRY(hipMemcpyAsync(&g->hrec[it], &g->drec[it], sizeof(IterRecord), hipMemcpyDeviceToHost, g->st));
      HIP_TRY(hipStreamSynchronize(g->st));
      kcollect(g);
      const IterRecord& r = g->hrec[it];
      nr_total += r.nr_steps;
      wave_total += r.nr_wave_steps;
      if (r.degenerate > 0)
        return set_error(VPFFT_E_DEGENERATE, "%.0f voxels with max resolved shear stress < 1e-10 (iteration %d)",
                         r.degenerate, it + 1);
      it++;
      if (r.done) break;
    }
  }
  if (fused) {
    // d~ and phi of the last iteration lived in LDS only: rebuild them from
    // the half spectrum that iteration read (the same inverse transform) and
    // from the final stress (the same polarization)
    void *sin, *sout;
    fused_buffers(g, it - 1, &sin, &sout);
    void* const* tw = prec == 0 ? g->tw64 : g->tw32;
    HIP_TRY(launch_z_c2r(prec, sin, g->dtil, g->geom, g->plan[2], tw[2], 1.0 / (double)g->n, 0, nullptr, g->st));
    HIP_TRY(launch_phi_rebuild(prec, g->n, g->dtil, g->phi, g->sigma, g->dC, g->st));
    g->dtil_prec = prec;
  }
  HIP_TRY(hipEventRecord(g->ev[2], g->st));
  const IterRecord& last = g->hrec[it - 1];
  HIP_TRY(hipMemcpyAsync(g->hhealth, g->dhealth, sizeof(HealthRecord), hipMemcpyDeviceToHost, g->st));
  HIP_TRY(hipStreamSynchronize(g->st));
  const bool loop_clean = !g->hhealth->flag;
  // epilogue: MaterialGrid.cpp:1074-1079 (skipped for dt <= 0, where the
  // reference's hardening divides 0 by 0, and after a non-finite iteration:
  // the kernel returns at entry then)
  if (dt > 0 && loop_clean) {
    kbegin(g, K_EPILOGUE);
    HIP_TRY(launch_epilogue(n, g->R, g->crss, g->acc, g->sigma, g->dC, g->dhealth, g->st));
    kend(g, K_EPILOGUE);
    g->crss_uniform = g->crss_uniform && g->hardening_uniform;
    double d[3];
    edot_diag(edot, d);
    g->lx *= (1.0 + d[0] * dt);
    g->ly *= (1.0 + d[1] * dt);
    g->lz *= (1.0 + d[2] * dt);
  }
  HIP_TRY(hipEventRecord(g->ev[3], g->st));
  HIP_TRY(hipMemcpyAsync(g->hhealth, g->dhealth, sizeof(HealthRecord), hipMemcpyDeviceToHost, g->st));
  HIP_TRY(hipStreamSynchronize(g->st));
  kcollect(g);
  bool finite = !g->hhealth->flag;
  for (int i = 0; i < 5; i++) finite = finite && std::isfinite(last.avg_stress[i]);
  finite = finite && std::isfinite(last.err_stress) && std::isfinite(last.err_strain_rate);
  if (avg)
    for (int i = 0; i < 5; i++) avg[i] = last.avg_stress[i];
  if (info) {
    std::memset(info, 0, sizeof *info);
    info->iterations = it;
    info->converged = (last.err_stress <= tol && last.err_strain_rate <= tol) ? 1 : 0;
    info->precision = prec;
    info->err_stress = last.err_stress;
    info->err_strain_rate = last.err_strain_rate;
    for (int i = 0; i < 5; i++) {
      info->avg_stress[i] = last.avg_stress[i];
      info->avg_strain_rate[i] = last.avg_strain_rate[i];
    }
    for (int e = 0; e < 25; e++) info->L0[e] = g->hprol->L0[e];
    info->max_nr_error = last.max_nr_error;
    info->max_nr_steps = last.max_nr_steps;
    info->nr_steps_last = last.nr_steps;
    info->nr_steps_total = nr_total;
    info->nr_wave_steps_total = wave_total;
    info->err_ratio_5 = err_ratio_5(g->hrec, it);
    info->taylor_steps_total = g->hprol->taylor_steps;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, g->ev[0], g->ev[1]);
    info->ms_prologue = ms;
    (void)hipEventElapsedTime(&ms, g->ev[1], g->ev[2]);
    info->ms_loop = ms;
    (void)hipEventElapsedTime(&ms, g->ev[2], g->ev[3]);
    info->ms_epilogue = ms;
    info->status = finite ? VPFFT_OK : VPFFT_E_NONFINITE;
  }
  if (!finite) {
    if (g->hhealth->flag) return health_error(g, *g->hhealth, loop_clean ? "epilogue" : "fixed-point loop");
    return set_error(VPFFT_E_NONFINITE, "non-finite average stress or error norm");
  }
  return VPFFT_OK;
}


// ---------------------------------------------------------------------------
// Batched strain steps.  Many grids (independent RVEs of any size) advance
// through one strain step together: every pass of the fixed-point loop is one
// launch per group of grids that share a kernel variant (precision, FFT block
// shape, Newton specialisation), so an iteration of the whole batch is a
// handful of launches however many grids it holds, and the iteration is
// replayed as a HIP graph.  Each grid keeps its own reference medium, loop
// flag and iteration count; a converged grid's blocks return at entry.  The
// FFT block shapes come from the defaults or from a tuner that times the
// candidates per (precision, pass, grid shape) once; a block shape only
// changes how pencils are grouped into blocks, never the arithmetic, so every
// grid's result is the same bits alone and in any batch.
// ---------------------------------------------------------------------------
enum LaunchKind { L_PRO, L_ZF, L_YF, L_X, L_YI, L_ZI, L_NR, L_EPI };

struct BLaunch {
  int kind, prec, B, E, nexp, uniform;
  long blocks;
  size_t lds;
  int ids_at, off_at, count;  // slices of the id / offset pools
};

// The radix orders the tuner tries, as orders of preference for
// make_axis_plan_ordered; the first is make_axis_plan's.
static const int kRadixOrders[][6] = {{8, 4, 2, 3, 5, 7}, {4, 8, 2, 3, 5, 7}, {4, 2, 8, 3, 5, 7},
                                      {2, 4, 8, 3, 5, 7}, {3, 5, 7, 8, 4, 2}, {7, 5, 3, 4, 8, 2}};
constexpr int kNumRadixOrders = (int)(sizeof(kRadixOrders) / sizeof(kRadixOrders[0]));

// An FFT pass's schedule for one grid shape: pencils per block and the
// radix order of the pass's axis (an index into kRadixOrders).
struct Sched {
  int B = 0;
  int order = 0;
};

struct Batch {
  hipStream_t st = nullptr;
  int cap = 0;
  FftJob* dfft = nullptr;
  VoxJob* dvox = nullptr;
  LoopState* dls = nullptr;
  int* dids = nullptr;
  long* doff = nullptr;
  int pool_cap = 0;
  // pinned host mirrors
  FftJob* hfft = nullptr;
  VoxJob* hvox = nullptr;
  LoopState* hls = nullptr;
  int* hids = nullptr;
  long* hoff = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  int graph_iters = 4;
  int tune = 0;
  // tune: 0 default schedules, 1 block shapes (the bits do not change),
  // 2 block shapes and radix orders (the rounding changes)
  std::map<std::tuple<int, int, int, int, int>, Sched> tuned;  // (prec, pass, nx, ny, nz) -> schedule
  hipGraphExec_t gexec = nullptr;
  std::vector<unsigned long> gkey_grids;
  std::vector<int> gkey_vals;
  double gkey_tol = 0;
  std::vector<BLaunch> launches;
};

static void batch_free_tables(Batch* b) {
  for (void* p : {(void*)b->dfft, (void*)b->dvox, (void*)b->dls, (void*)b->dids, (void*)b->doff})
    if (p) (void)hipFree(p);
  for (void* p : {(void*)b->hfft, (void*)b->hvox, (void*)b->hls, (void*)b->hids, (void*)b->hoff})
    if (p) (void)hipHostFree(p);
  b->dfft = nullptr; b->dvox = nullptr; b->dls = nullptr; b->dids = nullptr; b->doff = nullptr;
  b->hfft = nullptr; b->hvox = nullptr; b->hls = nullptr; b->hids = nullptr; b->hoff = nullptr;
  b->cap = b->pool_cap = 0;
}

static void batch_destroy(Batch* b) {
  if (!b) return;
  if (b->st) (void)hipStreamSynchronize(b->st);
  batch_free_tables(b);
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  for (auto& e : b->ev)
    if (e) (void)hipEventDestroy(e);
  if (b->st) (void)hipStreamDestroy(b->st);
  delete b;
}

static int batch_reserve(Batch* b, int count) {
  const int pool = 16 * count + 64;  // ids and offsets of every launch of one step
  if (count <= b->cap && pool <= b->pool_cap) return VPFFT_OK;
  if (b->gexec) {
    (void)hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
    b->gkey_grids.clear();
  }
  batch_free_tables(b);
  const int cap = std::max(count, 8);
  const int pcap = 16 * cap + 64;
  auto dal = [](void** p, size_t bytes) { return hipMalloc(p, bytes) == hipSuccess; };
  auto hal = [](void** p, size_t bytes) { return hipHostMalloc(p, bytes, hipHostMallocDefault) == hipSuccess; };
  bool ok = dal((void**)&b->dfft, sizeof(FftJob) * cap) && dal((void**)&b->dvox, sizeof(VoxJob) * cap) &&
            dal((void**)&b->dls, sizeof(LoopState) * cap) && dal((void**)&b->dids, sizeof(int) * pcap) &&
            dal((void**)&b->doff, sizeof(long) * pcap) && hal((void**)&b->hfft, sizeof(FftJob) * cap) &&
            hal((void**)&b->hvox, sizeof(VoxJob) * cap) && hal((void**)&b->hls, sizeof(LoopState) * cap) &&
            hal((void**)&b->hids, sizeof(int) * pcap) && hal((void**)&b->hoff, sizeof(long) * pcap);
  if (!ok) {
    batch_free_tables(b);
    return set_error(VPFFT_E_NOMEM, "batch tables for %d grids", count);
  }
  b->cap = cap;
  b->pool_cap = pcap;
  return VPFFT_OK;
}

static void fill_fft_job(const Grid* g, const int* done, FftJob* J) {
  const int prec = g->prec;
  J->g = g->geom;
  J->ax = g->plan[0];
  J->ay = g->plan[1];
  J->az = g->plan[2];
  for (int a = 0; a < 3; a++) J->tw[a] = prec == 0 ? g->tw64[a] : g->tw32[a];
  J->field = g->phi;
  J->dtil = g->dtil;
  J->spec = g->spec;
  J->gc = prec == 0 ? (const void*)g->dG64 : (const void*)g->dG32;
  J->done = done;
}

static void fill_vox_job(Grid* g, LoopState* ls, VoxJob* J) {
  J->n = g->n;
  J->dtil = g->dtil;
  J->phi = g->phi;
  J->sigma = g->sigma;
  J->lambda = g->lambda;
  J->R = g->R;
  J->crss = g->crss;
  J->acc = g->acc;
  J->C = g->dC;
  J->g64 = g->dG64;
  J->g32 = g->dG32;
  J->partial = g->partial;
  J->ls = ls;
  J->rec = g->drec;
  J->prol = g->dprol;
  J->health = g->dhealth;
}

static long pass_blocks(const Grid* g, int kind, int B) {
  if (kind == L_NR) return nr_blocks(g->n);
  if (kind == L_PRO || kind == L_EPI) return pt_blocks(g->n);
  const FftPass pass = (kind == L_ZF || kind == L_ZI) ? kPassZ : ((kind == L_YF || kind == L_YI) ? kPassY : kPassX);
  return pass_shape(g->prec, pass, g->geom, B).blocks;
}

// One grouped launch over the batch's FFT job table.
static int enqueue_fft_launch(Batch* b, const BLaunch& L) {
  const FftGroup grp{b->dfft, b->dids + L.ids_at, b->doff + L.off_at, L.count};
  switch (L.kind) {
    case L_ZF: HIP_TRY(launch_z_r2c_group(L.prec, L.B, grp, L.blocks, L.lds, b->st)); break;
    case L_YF: HIP_TRY(launch_y_group(L.prec, L.B, grp, L.blocks, L.lds, -1, b->st)); break;
    case L_X: HIP_TRY(launch_x_green_group(L.prec, L.B, L.E, grp, L.blocks, L.lds, b->st)); break;
    case L_YI: HIP_TRY(launch_y_group(L.prec, L.B, grp, L.blocks, L.lds, +1, b->st)); break;
    case L_ZI: HIP_TRY(launch_z_c2r_group(L.prec, L.B, grp, L.blocks, L.lds, b->st)); break;
  }
  return VPFFT_OK;
}

// The axis plan an FFT pass of job J reads (z: az, y: ay, x: ax).
static AxisPlan* job_plan(FftJob* J, FftPass pass) { return pass == kPassZ ? &J->az : (pass == kPassY ? &J->ay : &J->ax); }
static int pass_axis(FftPass pass) { return pass == kPassZ ? 2 : (pass == kPassY ? 1 : 0); }

// The schedule of an FFT pass for grid g (job `job` of the step's table):
// the default one, or the fastest of the candidates when tuning is on,
// measured once per (precision, pass, grid shape) on the batch's stream with
// the data of the step's grids of that shape, all in one launch (the batched
// passes are the generic Stockham ones, so every axis has a radix plan).
// Radix orders are tried only at tune level 2.  The job's plan for the
// pass's axis is left at the choice.
// The default schedule of a batched pass: the register form where the axis
// has one, else the generic pass's default block shape.
static Sched default_sched(const Grid* g, FftPass pass) {
  const PassShape rf = pass_shape(g->prec, pass, g->geom, kRegForm);
  return Sched{rf.ok ? rf.B : pass_shape(g->prec, pass, g->geom, 0).B, 0};
}

// same: the step's active jobs with g's precision and shape (job among
// them); a schedule is timed on all of them in one launch, as the step will
// run them, not on one grid alone (a small grid alone leaves most of the
// GPU idle and favours the schedules with the most blocks).
static int choose_sched(Batch* b, Grid* g, FftPass pass, int job, const std::vector<int>& same, Sched* out) {
  AxisPlan* jp = job_plan(b->hfft + job, pass);
  const int axis = pass_axis(pass);
  const int n = axis == 0 ? g->nx : (axis == 1 ? g->ny : g->nz);
  auto set_plan = [&](int order) {
    if (order == 0) *jp = g->plan[axis];
    else make_axis_plan_ordered(n, kRadixOrders[order], jp);
  };
  Sched best = default_sched(g, pass);
  // level 0, and level 1 for a register-form default (its arithmetic has no
  // block shape to vary): the default
  if (!b->tune || (b->tune == 1 && best.B < 0)) {
    set_plan(0);
    *out = best;
    return VPFFT_OK;
  }
  const auto key = std::make_tuple(g->prec, (int)pass, g->nx, g->ny, g->nz);
  auto it = b->tuned.find(key);
  if (it != b->tuned.end()) {
    set_plan(it->second.order);
    *out = it->second;
    return VPFFT_OK;
  }
  int cand[9];
  int nc = pass_candidates(pass, cand, 8);
  // level 2 also weighs the register form against the generic pass
  if (b->tune >= 2 && pass_shape(g->prec, pass, g->geom, kRegForm).ok) cand[nc++] = kRegForm;
  // distinct plans of the axis (orders giving the same stage sequence once)
  std::vector<int> orders{0};
  if (b->tune >= 2) {
    std::vector<unsigned long long> codes{g->plan[axis].code};
    for (int o = 1; o < kNumRadixOrders; o++) {
      AxisPlan ap;
      if (!make_axis_plan_ordered(n, kRadixOrders[o], &ap)) continue;
      if (std::find(codes.begin(), codes.end(), ap.code) != codes.end()) continue;
      codes.push_back(ap.code);
      orders.push_back(o);
    }
  }
  float best_ms = 1e30f;
  // the group at the end of the pools (the step's launches use at most
  // their first 14 entries per grid, from the front)
  const int m = (int)same.size();
  const int at = b->pool_cap - m - 2;
  if (m < 1 || at < b->pool_cap / 2) return set_error(VPFFT_E_INVALID, "batch tuning group of %d grids", m);
  for (int i = 0; i < m; i++) b->hids[at + i] = same[i];
  for (int order : orders) {
    for (int j : same) {
      AxisPlan* p = job_plan(b->hfft + j, pass);
      if (order == 0) *p = g->plan[axis];
      else make_axis_plan_ordered(n, kRadixOrders[order], p);
      HIP_TRY(hipMemcpyAsync(b->dfft + j, b->hfft + j, sizeof(FftJob), hipMemcpyHostToDevice, b->st));
    }
    for (int c = 0; c < nc; c++) {
      const PassShape ps = pass_shape(g->prec, pass, g->geom, cand[c]);
      if (!ps.ok || (ps.B < 0 && order != orders[0])) continue;  // the register form has no radix plan to vary
      for (int i = 0; i <= m; i++) b->hoff[at + i] = (long)i * ps.blocks;
      HIP_TRY(hipMemcpyAsync(b->dids + at, b->hids + at, sizeof(int) * m, hipMemcpyHostToDevice, b->st));
      HIP_TRY(hipMemcpyAsync(b->doff + at, b->hoff + at, sizeof(long) * (m + 1), hipMemcpyHostToDevice, b->st));
      BLaunch L{pass == kPassZ ? (int)L_ZF : (pass == kPassY ? (int)L_YF : (int)L_X), g->prec, ps.B, ps.E, 0, 0,
                (long)m * ps.blocks, ps.lds, at, at, m};
      // one warm-up launch, then the best of three timings of 12 launches:
      // a small grid's pass takes a few microseconds, so fewer launches
      // would rank candidates by timer noise
      if (int rc = enqueue_fft_launch(b, L)) return rc;
      float ms = 1e30f;
      for (int trial = 0; trial < 3; trial++) {
        HIP_TRY(hipEventRecord(b->ev[0], b->st));
        for (int rep = 0; rep < 12; rep++) {
          int rc = enqueue_fft_launch(b, L);
          if (rc) return rc;
        }
        HIP_TRY(hipEventRecord(b->ev[1], b->st));
        HIP_TRY(hipEventSynchronize(b->ev[1]));
        float t = 0;
        (void)hipEventElapsedTime(&t, b->ev[0], b->ev[1]);
        ms = std::min(ms, t);
      }
      // the default schedule is kept unless a candidate beats it by 3%
      const bool is_default = ps.B == default_sched(g, pass).B && order == 0;
      if (is_default) ms *= 0.97f;
      if (ms < best_ms) {
        best_ms = ms;
        best = Sched{ps.B, order};
      }
    }
  }
  set_plan(best.order);
  b->tuned[key] = best;
  *out = best;
  return VPFFT_OK;
}

// Group the grids of a step into launches and upload the tables.
static int batch_plan(Batch* b, Grid** gs, int count, const double* dt, const std::vector<char>& active) {
  b->launches.clear();
  int ids_at = 0, off_at = 0;
  auto add_group = [&](int kind, int prec, int B, int E, int nexp, int uniform, const std::vector<int>& members) {
    BLaunch L{kind, prec, B, E, nexp, uniform, 0, 0, ids_at, off_at, (int)members.size()};
    b->hoff[off_at] = 0;
    for (size_t m = 0; m < members.size(); m++) {
      Grid* g = gs[members[m]];
      b->hids[ids_at + m] = members[m];
      L.blocks += pass_blocks(g, kind, B);
      b->hoff[off_at + m + 1] = L.blocks;
      if (kind != L_PRO && kind != L_NR && kind != L_EPI) {
        const FftPass pass = (kind == L_ZF || kind == L_ZI) ? kPassZ : ((kind == L_YF || kind == L_YI) ? kPassY : kPassX);
        L.lds = std::max(L.lds, pass_shape(g->prec, pass, g->geom, B).lds);
      }
    }
    ids_at += (int)members.size();
    off_at += (int)members.size() + 1;
    b->launches.push_back(L);
  };
  // NR / epilogue groups (the prologue ran before the plan)
  std::map<std::tuple<int, int, int>, std::vector<int>> nr;
  std::vector<int> epi;
  for (int j = 0; j < count; j++) {
    if (!active[j]) continue;
    nr[std::make_tuple(gs[j]->prec, gs[j]->nexp_uniform, gs[j]->uniform_slips() ? 1 : 0)].push_back(j);
    if (dt[j] > 0) epi.push_back(j);
  }
  // the active grids of each precision and shape (a schedule is tuned on all of them)
  std::map<std::tuple<int, int, int, int>, std::vector<int>> shapes;
  for (int j = 0; j < count; j++)
    if (active[j]) shapes[std::make_tuple(gs[j]->prec, gs[j]->nx, gs[j]->ny, gs[j]->nz)].push_back(j);
  // FFT passes: groups by (precision, block shape)
  const int kinds[5] = {L_ZF, L_YF, L_X, L_YI, L_ZI};
  for (int kind : kinds) {
    const FftPass pass = (kind == L_ZF || kind == L_ZI) ? kPassZ : ((kind == L_YF || kind == L_YI) ? kPassY : kPassX);
    std::map<std::tuple<int, int, int>, std::vector<int>> fg;
    for (int j = 0; j < count; j++) {
      if (!active[j]) continue;
      Sched sc;
      if (kind == L_YI || kind == L_ZI) {  // the inverse passes take the forward pass's schedule
        const auto key = std::make_tuple(gs[j]->prec, (int)pass, gs[j]->nx, gs[j]->ny, gs[j]->nz);
        auto it = b->tuned.find(key);
        sc = (b->tune && it != b->tuned.end()) ? it->second : default_sched(gs[j], pass);
      } else {
        int rc = choose_sched(b, gs[j], pass, j, shapes[std::make_tuple(gs[j]->prec, gs[j]->nx, gs[j]->ny, gs[j]->nz)],
                              &sc);
        if (rc) return rc;
      }
      const PassShape ps = pass_shape(gs[j]->prec, pass, gs[j]->geom, sc.B);
      fg[std::make_tuple(gs[j]->prec, ps.B, ps.E)].push_back(j);
    }
    for (auto& kv : fg) add_group(kind, std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first), 0, 0,
                                  kv.second);
  }
  for (auto& kv : nr)
    add_group(L_NR, std::get<0>(kv.first), 0, 0, std::get<1>(kv.first), std::get<2>(kv.first), kv.second);
  // epilogue for the grids that move in time (MaterialGrid.cpp:1074-1079)
  if (!epi.empty()) add_group(L_EPI, 0, 0, 0, 0, 0, epi);
  // the one-block-per-grid kernels (the convergence test) take the active grids
  int nact = 0;
  for (int j = 0; j < count; j++)
    if (active[j]) b->hids[ids_at + nact++] = j;
  b->launches.push_back(BLaunch{-1, 0, 0, 0, 0, 0, 0, 0, ids_at, off_at, nact});
  ids_at += nact;
  if (ids_at > b->pool_cap - 8 || off_at > b->pool_cap - 8) return set_error(VPFFT_E_INVALID, "batch plan overflow");
  // the job table with each grid's chosen radix plans
  HIP_TRY(hipMemcpyAsync(b->dfft, b->hfft, sizeof(FftJob) * count, hipMemcpyHostToDevice, b->st));
  HIP_TRY(hipMemcpyAsync(b->dids, b->hids, sizeof(int) * ids_at, hipMemcpyHostToDevice, b->st));
  HIP_TRY(hipMemcpyAsync(b->doff, b->hoff, sizeof(long) * std::max(off_at, 1), hipMemcpyHostToDevice, b->st));
  return VPFFT_OK;
}

static VoxGroup vox_group(Batch* b, const BLaunch& L) {
  return VoxGroup{b->dvox, b->dids + L.ids_at, b->doff + L.off_at, L.count};
}

static int batch_iteration(Batch* b, double tol, int max_iter) {
  const BLaunch& all = b->launches.back();
  const int order[5] = {L_ZF, L_YF, L_X, L_YI, L_ZI};
  for (int kind : order)
    for (const BLaunch& L : b->launches)
      if (L.kind == kind) {
        int rc = enqueue_fft_launch(b, L);
        if (rc) return rc;
      }
  for (const BLaunch& L : b->launches)
    if (L.kind == L_NR) HIP_TRY(launch_nr_group(L.prec, L.nexp, L.uniform != 0, vox_group(b, L), L.blocks, b->st));
  HIP_TRY(launch_iter_finalize_group(vox_group(b, all), tol, max_iter, b->st));
  return VPFFT_OK;
}

// Fill infos[j] for a grid that did not take part in the step.
static void info_failed(vpfft_step_info* infos, double* avg, int j, int status, int prec) {
  if (avg)
    for (int i = 0; i < 5; i++) avg[5 * j + i] = std::nan("");
  if (infos) {
    std::memset(infos + j, 0, sizeof(vpfft_step_info));
    infos[j].status = status;
    infos[j].precision = prec;
  }
}

// One strain step of count grids.  A grid whose input is bad (non-finite or
// zero strain rate, non-finite time step, not uploaded) or whose prologue
// fails (degenerate voxels, singular secant modulus, non-finite values) gets
// its status in infos[j] (avg NaN) and sits out the rest of the step; the
// others run as usual.  Returns the first non-zero per-grid status.
static int batch_step(Batch* b, Grid** gs, int count, const double* edot, int max_iter, const double* dt,
                      double tol, double* avg, vpfft_step_info* infos) {
  if (count < 1 || !gs || !edot || !dt) return set_error(VPFFT_E_INVALID, "empty batch");
  auto fail_all = [&](int code) {
    for (int j = 0; j < count; j++) info_failed(infos, avg, j, code, gs[j] ? gs[j]->prec : 0);
    return code;
  };
  if (max_iter < 1) return fail_all(set_error(VPFFT_E_INVALID, "max_iter must be >= 1"));
  if (!(tol >= 0)) return fail_all(set_error(VPFFT_E_INVALID, "bad tolerance"));
  for (int j = 0; j < count; j++) {
    if (!gs[j]) return fail_all(set_error(VPFFT_E_INVALID, "null grid %d", j));
    for (int k = 0; k < j; k++)
      if (gs[k] == gs[j]) return fail_all(set_error(VPFFT_E_INVALID, "grid %d appears twice", j));
  }
  std::vector<int> status(count, VPFFT_OK);
  int first = VPFFT_OK;
  auto grid_fail = [&](int j, int code) {
    status[j] = code;
    if (!first) first = code;
  };
  for (int j = 0; j < count; j++) {
    if (!gs[j]->uploaded) {
      grid_fail(j, set_error(VPFFT_E_STATE, "grid %d: run before upload", j));
      continue;
    }
    if (!std::isfinite(dt[j])) {
      grid_fail(j, set_error(VPFFT_E_INVALID, "grid %d: bad time step", j));
      continue;
    }
    double ee = 0;
    bool fin = true;
    for (int i = 0; i < 5; i++) {
      fin = fin && std::isfinite(edot[5 * j + i]);
      ee += edot[5 * j + i] * edot[5 * j + i];
    }
    if (!fin) grid_fail(j, set_error(VPFFT_E_NONFINITE, "grid %d: non-finite strain rate", j));
    else if (!(ee > 0)) grid_fail(j, set_error(VPFFT_E_DEGENERATE, "grid %d: zero strain rate", j));
  }
  std::vector<char> active(count);
  int nact = 0;
  for (int j = 0; j < count; j++) nact += (active[j] = status[j] == VPFFT_OK);
  auto finish_failed = [&]() {
    for (int j = 0; j < count; j++)
      if (status[j]) info_failed(infos, avg, j, status[j], gs[j]->prec);
  };
  if (nact == 0) {
    finish_failed();
    return first;
  }
  int rc = batch_reserve(b, count);
  if (rc) return fail_all(rc);
  for (int j = 0; j < count; j++) {
    Grid* g = gs[j];
    if ((rc = ensure_records(g, max_iter))) return fail_all(rc);
    fill_host_const(g, edot + 5 * j, dt[j]);
    if (active[j]) {
      HIP_TRY(hipMemcpyAsync(g->dC, g->hC, sizeof(StepConst), hipMemcpyHostToDevice, b->st));
      HIP_TRY(launch_health_reset(g->dhealth, b->st));
    }
    fill_vox_job(g, b->dls + j, b->hvox + j);
    fill_fft_job(g, &b->dls[j].done, b->hfft + j);
  }
  HIP_TRY(hipMemcpyAsync(b->dvox, b->hvox, sizeof(VoxJob) * count, hipMemcpyHostToDevice, b->st));
  HIP_TRY(hipMemcpyAsync(b->dfft, b->hfft, sizeof(FftJob) * count, hipMemcpyHostToDevice, b->st));
  // prologue first (the tuner times the FFT passes on valid Green constants)
  {
    std::map<int, std::vector<int>> pro;
    for (int j = 0; j < count; j++)
      if (active[j]) pro[gs[j]->prec].push_back(j);
    int ids_at = 0, off_at = 0;
    std::vector<BLaunch> pl;
    for (auto& kv : pro) {
      BLaunch L{L_PRO, kv.first, 0, 0, 0, 0, 0, 0, ids_at, off_at, (int)kv.second.size()};
      b->hoff[off_at] = 0;
      for (size_t m = 0; m < kv.second.size(); m++) {
        b->hids[ids_at + m] = kv.second[m];
        L.blocks += pt_blocks(gs[kv.second[m]]->n);
        b->hoff[off_at + m + 1] = L.blocks;
      }
      ids_at += L.count;
      off_at += L.count + 1;
      pl.push_back(L);
    }
    int k = 0;
    for (int j = 0; j < count; j++)
      if (active[j]) b->hids[ids_at + k++] = j;
    HIP_TRY(hipMemcpyAsync(b->dids, b->hids, sizeof(int) * (ids_at + nact), hipMemcpyHostToDevice, b->st));
    HIP_TRY(hipMemcpyAsync(b->doff, b->hoff, sizeof(long) * off_at, hipMemcpyHostToDevice, b->st));
    HIP_TRY(hipEventRecord(b->ev[0], b->st));
    for (const BLaunch& L : pl) HIP_TRY(launch_prologue_group(L.prec, vox_group(b, L), L.blocks, b->st));
    const VoxGroup all{b->dvox, b->dids + ids_at, nullptr, nact};
    HIP_TRY(launch_L0_finalize_group(all, b->st));
    HIP_TRY(launch_loop_reset_group(all, b->st));
    for (int j = 0; j < count; j++)
      if (active[j]) {
        HIP_TRY(hipMemcpyAsync(gs[j]->hprol, gs[j]->dprol, sizeof(PrologueRecord), hipMemcpyDeviceToHost, b->st));
        HIP_TRY(hipMemcpyAsync(gs[j]->hhealth, gs[j]->dhealth, sizeof(HealthRecord), hipMemcpyDeviceToHost, b->st));
      }
    HIP_TRY(hipEventRecord(b->ev[1], b->st));
    HIP_TRY(hipStreamSynchronize(b->st));
    for (int j = 0; j < count; j++) {
      if (!active[j]) continue;
      Grid* g = gs[j];
      if (g->hprol->degenerate > 0)
        grid_fail(j, set_error(VPFFT_E_DEGENERATE, "grid %d: %.0f voxels with max resolved shear stress < 1e-10", j,
                               g->hprol->degenerate));
      else if (g->hhealth->flag)
        grid_fail(j, health_error(g, *g->hhealth, "initialization"));
      else if (!spd5(g->hprol->L0) || !spd5(g->hprol->L0inv))
        grid_fail(j, set_error(VPFFT_E_NONFINITE, "grid %d: reference medium L0 is not symmetric positive definite",
                               j));
      if (status[j]) {
        active[j] = 0;
        nact--;
      }
    }
  }
  if (nact == 0) {
    finish_failed();
    return first;
  }
  if ((rc = batch_plan(b, gs, count, dt, active))) return fail_all(rc);
  // the loop, replayed as a graph of graph_iters iterations
  std::vector<unsigned long> key_grids;
  for (int j = 0; j < count; j++) key_grids.push_back(active[j] ? gs[j]->serial : 0);
  std::vector<int> key_vals{max_iter, count};
  for (const BLaunch& L : b->launches) {
    key_vals.push_back(L.kind);
    key_vals.push_back(L.B);
    key_vals.push_back((int)L.blocks);
  }
  for (int j = 0; j < count; j++) {
    key_vals.push_back(gs[j]->prec);
    key_vals.push_back(gs[j]->uniform_slips());
  }
  if (b->graph_iters > 0) {
    if (!b->gexec || key_grids != b->gkey_grids || key_vals != b->gkey_vals || tol != b->gkey_tol) {
      if (b->gexec) {
        HIP_TRY(hipGraphExecDestroy(b->gexec));
        b->gexec = nullptr;
      }
      hipGraph_t graph;
      HIP_TRY(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
      int crc = VPFFT_OK;
      for (int k = 0; k < b->graph_iters && crc == VPFFT_OK; k++) crc = batch_iteration(b, tol, max_iter);
      hipError_t ce = hipStreamEndCapture(b->st, &graph);
      if (crc) return fail_all(crc);
      HIP_TRY(ce);
      hipError_t ie = hipGraphInstantiate(&b->gexec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      HIP_TRY(ie);
      b->gkey_grids = key_grids;
      b->gkey_vals = key_vals;
      b->gkey_tol = tol;
    }
  }
  for (;;) {
    if (b->graph_iters > 0) {
      HIP_TRY(hipGraphLaunch(b->gexec, b->st));
    } else if ((rc = batch_iteration(b, tol, max_iter))) {
      return fail_all(rc);
    }
    HIP_TRY(hipMemcpyAsync(b->hls, b->dls, sizeof(LoopState) * count, hipMemcpyDeviceToHost, b->st));
    HIP_TRY(hipStreamSynchronize(b->st));
    bool all_done = true;
    for (int j = 0; j < count; j++) all_done = all_done && (!active[j] || b->hls[j].done);
    if (all_done) break;
  }
  std::vector<int> iters(count, 0);
  for (int j = 0; j < count; j++) {
    if (!active[j]) continue;
    iters[j] = b->hls[j].iter;
    HIP_TRY(hipMemcpyAsync(gs[j]->hrec, gs[j]->drec, sizeof(IterRecord) * iters[j], hipMemcpyDeviceToHost, b->st));
  }
  HIP_TRY(hipEventRecord(b->ev[2], b->st));
  // the epilogue kernel returns at entry for a grid whose health flag is set
  for (const BLaunch& L : b->launches)
    if (L.kind == L_EPI) HIP_TRY(launch_epilogue_group(vox_group(b, L), L.blocks, b->st));
  HIP_TRY(hipEventRecord(b->ev[3], b->st));
  // after the epilogue: the record names the loop's failure, or the
  // epilogue's own (the record's kernel field tells which)
  for (int j = 0; j < count; j++)
    if (active[j])
      HIP_TRY(hipMemcpyAsync(gs[j]->hhealth, gs[j]->dhealth, sizeof(HealthRecord), hipMemcpyDeviceToHost, b->st));
  HIP_TRY(hipStreamSynchronize(b->st));
  float ms[3] = {0, 0, 0};
  for (int k = 0; k < 3; k++) (void)hipEventElapsedTime(&ms[k], b->ev[k], b->ev[k + 1]);
  for (int j = 0; j < count; j++) {
    Grid* g = gs[j];
    if (!active[j]) continue;
    g->phi_prec = g->dtil_prec = g->prec;
    const int it = iters[j];
    double nr_total = 0, wave_total = 0;
    int degen_at = 0;
    for (int i = 0; i < it; i++) {
      nr_total += g->hrec[i].nr_steps;
      wave_total += g->hrec[i].nr_wave_steps;
      if (!degen_at && g->hrec[i].degenerate > 0) degen_at = i + 1;
    }
    const IterRecord& last = g->hrec[it - 1];
    bool finite = !g->hhealth->flag;
    for (int i = 0; i < 5; i++) finite = finite && std::isfinite(last.avg_stress[i]);
    finite = finite && std::isfinite(last.err_stress) && std::isfinite(last.err_strain_rate);
    // the epilogue ran (and moved the state on) unless the loop went bad
    if (dt[j] > 0 && finite && !degen_at) {
      g->crss_uniform = g->crss_uniform && g->hardening_uniform;
      double d[3];
      edot_diag(edot + 5 * j, d);
      g->lx *= (1.0 + d[0] * dt[j]);
      g->ly *= (1.0 + d[1] * dt[j]);
      g->lz *= (1.0 + d[2] * dt[j]);
    }
    int st = VPFFT_OK;
    if (degen_at)
      st = set_error(VPFFT_E_DEGENERATE, "grid %d: degenerate voxels (iteration %d)", j, degen_at);
    else if (g->hhealth->flag)
      st = health_error(g, *g->hhealth, "strain step");
    else if (!finite)
      st = set_error(VPFFT_E_NONFINITE, "grid %d: non-finite result", j);
    if (st) grid_fail(j, st);
    if (avg)
      for (int i = 0; i < 5; i++) avg[5 * j + i] = last.avg_stress[i];
    if (infos) {
      vpfft_step_info* info = infos + j;
      std::memset(info, 0, sizeof *info);
      info->iterations = it;
      info->converged = (last.err_stress <= tol && last.err_strain_rate <= tol) ? 1 : 0;
      info->status = st;
      info->precision = g->prec;
      info->err_stress = last.err_stress;
      info->err_strain_rate = last.err_strain_rate;
      for (int i = 0; i < 5; i++) {
        info->avg_stress[i] = last.avg_stress[i];
        info->avg_strain_rate[i] = last.avg_strain_rate[i];
      }
      for (int e = 0; e < 25; e++) info->L0[e] = g->hprol->L0[e];
      info->max_nr_error = last.max_nr_error;
      info->max_nr_steps = last.max_nr_steps;
      info->nr_steps_last = last.nr_steps;
      info->nr_steps_total = nr_total;
      info->nr_wave_steps_total = wave_total;
      info->err_ratio_5 = err_ratio_5(g->hrec, it);
      info->taylor_steps_total = g->hprol->taylor_steps;
      info->ms_prologue = ms[0];
      info->ms_loop = ms[1];
      info->ms_epilogue = ms[2];
    }
  }
  finish_failed();
  return first;
}

static int batch_create(Batch** out) {
  Batch* b = new Batch();
  if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return set_error(VPFFT_E_DEVICE, "hipStreamCreate failed");
  }
  for (auto& e : b->ev)
    if (hipEventCreate(&e) != hipSuccess) {
      batch_destroy(b);
      return set_error(VPFFT_E_DEVICE, "hipEventCreate failed");
    }
  if (const char* k = std::getenv("VPFFT_GRAPH_ITERS")) b->graph_iters = std::max(0, std::min(64, std::atoi(k)));
  *out = b;
  return VPFFT_OK;
}

// ---------------------------------------------------------------------------
// Scratch of vpfft_strain_rate_evolution_noalloc: the per-call constants,
// the twiddles and the half spectrum of one call (and, in the mixed form, the
// FP32 copies of the input and output fields), carved out of one static
// device array so the call allocates nothing.  Every GPU the library runs on
// holds its own copy until the process ends; calls on one GPU share it.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static std::mutex g_scratch_mu;
constexpr int kMaxDevices = 64;
static void* g_scratch_ptr[kMaxDevices];  // the array's device address, per device

// The current device's copy of the array (hipGetSymbolAddress on the first
// call on that device).  Host code uses this address, never the array's name.
static int scratch_address(void** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return set_error(VPFFT_E_DEVICE, "device %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(VPFFT_E_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return VPFFT_OK;
}

// Offsets (256-byte aligned) of one call's parts in the scratch array.
struct NoAllocLayout {
  size_t C, g64, g32, tw64[3], tw32[3], spec, phi32, dtil32, total;
};
static NoAllocLayout noalloc_layout(int nx, int ny, int nz, int prec) {
  auto up = [](size_t x) { return (x + 255) / 256 * 256; };
  NoAllocLayout L{};
  size_t o = 0;
  L.C = o;
  o = up(o + sizeof(StepConst));
  L.g64 = o;
  o = up(o + sizeof(GreenConst<double>));
  L.g32 = o;
  o = up(o + sizeof(GreenConst<float>));
  const int dims[3] = {nx, ny, nz};
  for (int a = 0; a < 3; a++) {
    L.tw64[a] = o;
    o = up(o + sizeof(double2) * dims[a]);
    L.tw32[a] = o;
    o = up(o + sizeof(float2) * dims[a]);
  }
  const size_t n = (size_t)nx * ny * nz, nc = (size_t)nx * ny * padded_half(nz);
  L.spec = o;
  o = up(o + (prec == 0 ? sizeof(double2) : sizeof(float2)) * 5 * nc);
  if (prec == 1) {
    L.phi32 = o;
    o = up(o + sizeof(float) * 5 * n);
    L.dtil32 = o;
    o = up(o + sizeof(float) * 5 * n);
  }
  L.total = o;
  return L;
}

}  // namespace vpfft

using vpfft::Grid;
using vpfft::set_error;

static Grid* as_grid(void* p) { return static_cast<Grid*>(p); }

extern "C" {

const char* vpfft_last_error(void) { return vpfft::g_err.c_str(); }
const char* vpfft_version(void) { return "vpfft-mi355x 0.1 (gfx950)"; }

int vpfft_grid_create(const vpfft_grid_desc* desc, const vpfft_material* mat, void** out_grid) {
  Grid* g = nullptr;
  int rc = vpfft::grid_create(desc, mat, &g);
  if (out_grid) *out_grid = rc == VPFFT_OK ? (void*)g : nullptr;
  return rc;
}

int vpfft_grid_destroy(void* grid) {
  vpfft::grid_free(as_grid(grid));
  return VPFFT_OK;
}

int vpfft_grid_set_voxel_length(void* grid, double lx, double ly, double lz) {
  Grid* g = as_grid(grid);
  if (!g) return set_error(VPFFT_E_INVALID, "null grid");
  if (!(lx > 0 && ly > 0 && lz > 0)) return set_error(VPFFT_E_INVALID, "voxel lengths must be positive");
  g->lx = lx;
  g->ly = ly;
  g->lz = lz;
  return VPFFT_OK;
}

int vpfft_grid_get_voxel_length(void* grid, double out[3]) {
  Grid* g = as_grid(grid);
  if (!g || !out) return set_error(VPFFT_E_INVALID, "null argument");
  out[0] = g->lx;
  out[1] = g->ly;
  out[2] = g->lz;
  return VPFFT_OK;
}

int vpfft_grid_set_precision(void* grid, int precision) {
  Grid* g = as_grid(grid);
  if (!g) return set_error(VPFFT_E_INVALID, "null grid");
  if (precision != VPFFT_FP64 && precision != VPFFT_MIXED) return set_error(VPFFT_E_INVALID, "bad precision");
  g->prec = precision;
  return VPFFT_OK;
}

int vpfft_grid_set_profiling(void* grid, int enable) {
  Grid* g = as_grid(grid);
  if (!g) return set_error(VPFFT_E_INVALID, "null grid");
  g->profile = enable != 0;
  for (int k = 0; k < vpfft::kNumKernels; k++) {
    g->kms[k] = 0;
    g->kcount[k] = 0;
    g->kev_used[k] = false;
  }
  return VPFFT_OK;
}

int vpfft_grid_kernel_times(void* grid, double* total_ms, long* launches, int capacity) {
  Grid* g = as_grid(grid);
  if (!g || !total_ms || !launches) return set_error(VPFFT_E_INVALID, "null argument");
  const int n = std::min(capacity, (int)vpfft::kNumKernels);
  for (int k = 0; k < n; k++) {
    total_ms[k] = g->kms[k];
    launches[k] = g->kcount[k];
  }
  return vpfft::kNumKernels;
}

const char* vpfft_kernel_name(int id) {
  return (id >= 0 && id < vpfft::kNumKernels) ? vpfft::kKernelNames[id] : "";
}

int vpfft_grid_upload(void* grid, const double* orientation, const double* crss, double crss_value,
                      const double* stress) {
  Grid* g = as_grid(grid);
  if (!g || !orientation) return set_error(VPFFT_E_INVALID, "null grid or orientation");
  int rc;
  if ((rc = vpfft::upload_aos(g, orientation, g->R, 9))) return rc;
  if (crss) {
    if ((rc = vpfft::upload_aos(g, crss, g->crss, 12))) return rc;
    g->crss_uniform = vpfft::crss_rows_uniform(crss, g->n);
  } else {
    g->crss_uniform = true;
    if (!(crss_value > 0)) return set_error(VPFFT_E_INVALID, "crss_value must be positive");
    hipLaunchKernelGGL(vpfft::k_fill, dim3(vpfft::fill_blocks(12 * g->n)), dim3(256), 0, g->st, g->crss, crss_value,
                       12 * g->n);
  }
  if (stress) {
    if ((rc = vpfft::upload_aos(g, stress, g->sigma, 5))) return rc;
  } else {
    hipLaunchKernelGGL(vpfft::k_fill, dim3(vpfft::fill_blocks(5 * g->n)), dim3(256), 0, g->st, g->sigma, 0.0,
                       5 * g->n);
  }
  hipLaunchKernelGGL(vpfft::k_fill, dim3(vpfft::fill_blocks(g->n)), dim3(256), 0, g->st, g->acc, 0.0, g->n);
  hipError_t e = hipStreamSynchronize(g->st);
  if (e != hipSuccess) return set_error(VPFFT_E_DEVICE, "upload: %s", hipGetErrorString(e));
  g->uploaded = true;
  return VPFFT_OK;
}

int vpfft_grid_upload_field(void* grid, int field, const double* host) {
  Grid* g = as_grid(grid);
  if (!g || !host) return set_error(VPFFT_E_INVALID, "null argument");
  switch (field) {
    case VPFFT_FIELD_ORIENTATION: return vpfft::upload_aos(g, host, g->R, 9);
    case VPFFT_FIELD_CRSS:
      g->crss_uniform = vpfft::crss_rows_uniform(host, g->n);
      return vpfft::upload_aos(g, host, g->crss, 12);
    case VPFFT_FIELD_STRESS: return vpfft::upload_aos(g, host, g->sigma, 5);
    case VPFFT_FIELD_LAGRANGE: return vpfft::upload_aos(g, host, g->lambda, 5);
    case VPFFT_FIELD_POLARIZATION: g->phi_prec = g->prec; return vpfft::upload_field5(g, host, g->phi, g->prec);
    case VPFFT_FIELD_ACC_SHEAR: {
      hipError_t e = hipMemcpy(g->acc, host, sizeof(double) * g->n, hipMemcpyHostToDevice);
      return e == hipSuccess ? VPFFT_OK : set_error(VPFFT_E_DEVICE, "upload acc shear: %s", hipGetErrorString(e));
    }
    default: return set_error(VPFFT_E_INVALID, "field %d cannot be uploaded", field);
  }
}

int vpfft_grid_download(void* grid, int field, double* host) {
  Grid* g = as_grid(grid);
  if (!g || !host) return set_error(VPFFT_E_INVALID, "null argument");
  switch (field) {
    case VPFFT_FIELD_ORIENTATION: return vpfft::download_aos(g, g->R, host, 9);
    case VPFFT_FIELD_CRSS: return vpfft::download_aos(g, g->crss, host, 12);
    case VPFFT_FIELD_STRESS: return vpfft::download_aos(g, g->sigma, host, 5);
    case VPFFT_FIELD_LAGRANGE: return vpfft::download_aos(g, g->lambda, host, 5);
    case VPFFT_FIELD_POLARIZATION: return vpfft::download_field5(g, g->phi, host, g->phi_prec);
    case VPFFT_FIELD_STRAIN_RATE_VARIATION: return vpfft::download_field5(g, g->dtil, host, g->dtil_prec);
    case VPFFT_FIELD_ACC_SHEAR: {
      hipError_t e = hipMemcpy(host, g->acc, sizeof(double) * g->n, hipMemcpyDeviceToHost);
      return e == hipSuccess ? VPFFT_OK : set_error(VPFFT_E_DEVICE, "download acc shear: %s", hipGetErrorString(e));
    }
    case VPFFT_FIELD_SCHMID: {
      const long chunk = g->staging_elems / 60;
      for (long off = 0; off < g->n; off += chunk) {
        const long cnt = std::min(chunk, g->n - off);
        hipLaunchKernelGGL(vpfft::k_schmid_aos, dim3(vpfft::fill_blocks(cnt)), dim3(256), 0, g->st, g->R, g->n, off,
                           cnt, g->staging);
        hipError_t e = hipMemcpyAsync(host + 60 * off, g->staging, sizeof(double) * 60 * cnt, hipMemcpyDeviceToHost,
                                      g->st);
        if (e == hipSuccess) e = hipStreamSynchronize(g->st);
        if (e != hipSuccess) return set_error(VPFFT_E_DEVICE, "download schmid: %s", hipGetErrorString(e));
      }
      return VPFFT_OK;
    }
    case VPFFT_FIELD_STRAIN_RATE: {
      vpfft::RateLaw law;
      for (int k = 0; k < vpfft::kSlips; k++) {
        law.gdot0[k] = g->mat.gamma_dot_base[k];
        law.nexp[k] = g->mat.rate_sensitivity[k];
      }
      const long chunk = g->staging_elems / 5;
      for (long off = 0; off < g->n; off += chunk) {
        const long cnt = std::min(chunk, g->n - off);
        hipLaunchKernelGGL(vpfft::k_strain_rate_aos, dim3(vpfft::fill_blocks(cnt)), dim3(256), 0, g->st, g->R,
                           g->crss, g->sigma, g->n, off, cnt, law, g->staging);
        hipError_t e = hipMemcpyAsync(host + 5 * off, g->staging, sizeof(double) * 5 * cnt, hipMemcpyDeviceToHost,
                                      g->st);
        if (e == hipSuccess) e = hipStreamSynchronize(g->st);
        if (e != hipSuccess) return set_error(VPFFT_E_DEVICE, "download strain rate: %s", hipGetErrorString(e));
      }
      return VPFFT_OK;
    }
    default: return set_error(VPFFT_E_INVALID, "unknown field %d", field);
  }
}

int vpfft_run_strain_step(void* grid, const double edot[5], int max_iter, double dt, double tol,
                          double avg_stress[5], vpfft_step_info* info) {
  Grid* g = as_grid(grid);
  if (!g || !edot) return set_error(VPFFT_E_INVALID, "null argument");
  return vpfft::strain_step(g, edot, max_iter, dt, tol, avg_stress, info);
}

int vpfft_run_vpfft(void* grid, const double edot[5], int max_iter, int num_steps, double dt, double tol,
                    double* avg_stress_out, vpfft_step_info* infos, const char* stress_strain_path) {
  Grid* g = as_grid(grid);
  if (!g || !edot || num_steps < 0) return set_error(VPFFT_E_INVALID, "bad argument");
  FILE* os = nullptr;
  if (stress_strain_path) {
    os = std::fopen(stress_strain_path, "w");
    if (!os) return set_error(VPFFT_E_INVALID, "cannot open %s", stress_strain_path);
  }
  double E[5] = {0, 0, 0, 0, 0};
  for (int s = 0; s < num_steps; s++) {
    double avg[5];
    vpfft_step_info info;
    int rc = vpfft::strain_step(g, edot, max_iter, dt, tol, avg, &info);
    if (rc) {
      if (os) std::fclose(os);
      return rc;
    }
    if (avg_stress_out)
      for (int i = 0; i < 5; i++) avg_stress_out[5 * s + i] = avg[i];
    if (infos) infos[s] = info;
    for (int i = 0; i < 5; i++) E[i] += edot[i] * dt;  // MaterialGrid.cpp:1217
    if (os) {  // MaterialGrid.cpp:1221-1225 (ostream default format == %g)
      double nE = 0, nS = 0;
      for (int i = 0; i < 5; i++) { nE += E[i] * E[i]; nS += avg[i] * avg[i]; }
      std::fprintf(os, "%d %g %g %g %g %g  %g %g %g %g %g  %g %g \n", s, E[0], E[1], E[2], E[3], E[4], avg[0], avg[1],
                   avg[2], avg[3], avg[4], std::sqrt(nE), std::sqrt(nS));
    }
  }
  if (os) std::fclose(os);
  return VPFFT_OK;
}

int vpfft_strain_rate_evolution(void* grid, const double L0[25], const double* phi_in, double* dtilde_out) {
  Grid* g = as_grid(grid);
  if (!g || !L0 || !phi_in || !dtilde_out) return set_error(VPFFT_E_INVALID, "null argument");
  const double e0[5] = {1, 0, 0, 0, 0};
  vpfft::fill_host_const(g, e0, 0.0);
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 5; j++) g->hC->L0[j * 5 + i] = 0.5 * (L0[j * 5 + i] + L0[i * 5 + j]);
  hipError_t e = hipMemcpyAsync(g->dC, g->hC, sizeof(vpfft::StepConst), hipMemcpyHostToDevice, g->st);
  if (e != hipSuccess) return set_error(VPFFT_E_DEVICE, "%s", hipGetErrorString(e));
  e = vpfft::launch_derive_from_L0(g->dC, g->dG64, g->dG32, g->st);
  if (e != hipSuccess) return set_error(VPFFT_E_DEVICE, "%s", hipGetErrorString(e));
  int rc = vpfft::upload_field5(g, phi_in, g->phi, g->prec);
  if (rc) return rc;
  g->phi_prec = g->prec;
  if ((rc = vpfft::run_fft_green(g, g->prec, nullptr))) return rc;
  e = hipStreamSynchronize(g->st);
  if (e != hipSuccess) return set_error(VPFFT_E_DEVICE, "%s", hipGetErrorString(e));
  return vpfft::download_field5(g, g->dtil, dtilde_out, g->dtil_prec);
}

int vpfft_strain_rate_evolution_noalloc(int nx, int ny, int nz, const double voxel_length[3], int precision,
                                        const double L0[25], const double* d_phi, double* d_dtilde, void* stream) {
  using namespace vpfft;
  if (!voxel_length || !L0 || !d_phi || !d_dtilde) return set_error(VPFFT_E_INVALID, "null argument");
  if (nx < 2 || ny < 2 || nz < 2 || nx > 4096 || ny > 4096 || nz > 4096)
    return set_error(VPFFT_E_INVALID, "grid edges must be in [2, 4096]");
  if (precision != VPFFT_FP64 && precision != VPFFT_MIXED)
    return set_error(VPFFT_E_INVALID, "precision must be VPFFT_FP64 or VPFFT_MIXED");
  for (int a = 0; a < 3; a++)
    if (!(voxel_length[a] > 0) || !std::isfinite(voxel_length[a]))
      return set_error(VPFFT_E_INVALID, "voxel lengths must be positive");
  const int dims[3] = {nx, ny, nz};
  AxisPlan plan[3];
  for (int a = 0; a < 3; a++)
    if (!make_axis_plan(dims[a], &plan[a]))
      return set_error(VPFFT_E_UNSUPPORTED, "grid edge %d has a prime factor other than 2, 3, 5, 7", dims[a]);
  if ((long)4 * nz * 16 * 2 > 160 * 1024 || (long)ny * 16 * 2 > 160 * 1024 || (long)nx * 16 * 8 * 2 > 160 * 1024)
    return set_error(VPFFT_E_UNSUPPORTED, "grid edge too long for the LDS-resident FFT");
  const NoAllocLayout lay = noalloc_layout(nx, ny, nz, precision);
  if (lay.total > kNoAllocScratchBytes)
    return set_error(VPFFT_E_INVALID,
                     "a %d x %d x %d grid needs %zu bytes of scratch; the no-allocation entry point holds %zu "
                     "(a 64^3 grid fits in either precision)", nx, ny, nz, lay.total, kNoAllocScratchBytes);
  void* base = nullptr;
  int rc = scratch_address(&base);
  if (rc) return rc;
  unsigned char* b = static_cast<unsigned char*>(base);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int prec = precision;
  const long n = (long)nx * ny * nz;
  // the constants vpfft_strain_rate_evolution derives, from the same inputs
  StepConst hc;
  std::memset(&hc, 0, sizeof hc);
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 5; j++) hc.L0[j * 5 + i] = 0.5 * (L0[j * 5 + i] + L0[i * 5 + j]);
  for (int a = 0; a < 3; a++) hc.kscale[a] = 1.0 / ((double)dims[a] * voxel_length[a]);
  hc.inv_n = 1.0 / (double)n;
  hc.nx = nx;
  hc.ny = ny;
  hc.nz = nz;
  hc.nzh = nz / 2 + 1;
  auto* dC = reinterpret_cast<StepConst*>(b + lay.C);
  auto* g64 = reinterpret_cast<GreenConst<double>*>(b + lay.g64);
  auto* g32 = reinterpret_cast<GreenConst<float>*>(b + lay.g32);
  HIP_TRY(hipMemcpyAsync(dC, &hc, sizeof hc, hipMemcpyHostToDevice, st));
  void* tw[3];
  std::vector<double2> t64s[3];  // kept until the stream is synchronised below
  std::vector<float2> t32s[3];
  for (int a = 0; a < 3; a++) {  // the twiddles vpfft_grid_create makes
    std::vector<double2>& t64 = t64s[a];
    std::vector<float2>& t32 = t32s[a];
    t64.resize(dims[a]);
    t32.resize(dims[a]);
    for (int t = 0; t < dims[a]; t++) {
      const long double ang = -2.0L * 3.14159265358979323846264338327950288L * (long double)t / (long double)dims[a];
      t64[t].x = (double)cosl(ang);
      t64[t].y = (double)sinl(ang);
      t32[t].x = (float)t64[t].x;
      t32[t].y = (float)t64[t].y;
    }
    HIP_TRY(hipMemcpyAsync(b + lay.tw64[a], t64.data(), sizeof(double2) * dims[a], hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(b + lay.tw32[a], t32.data(), sizeof(float2) * dims[a], hipMemcpyHostToDevice, st));
    tw[a] = prec == 0 ? (void*)(b + lay.tw64[a]) : (void*)(b + lay.tw32[a]);
  }
  HIP_TRY(launch_derive_from_L0(dC, g64, g32, st));
  FftGeom geom{};
  geom.n = n;
  geom.nc = (long)nx * ny * padded_half(nz);
  geom.nrows = (long)nx * ny;
  geom.nx = nx;
  geom.ny = ny;
  geom.nz = nz;
  geom.nzh = nz / 2 + 1;
  geom.nzp = padded_half(nz);
  void* spec = b + lay.spec;
  const void* field = d_phi;
  void* out = d_dtilde;
  if (prec == 1) {
    HIP_TRY(launch_convert(1, b + lay.phi32, 0, d_phi, 5 * n, st));
    field = b + lay.phi32;
    out = b + lay.dtil32;
  }
  const void* gc = prec == 0 ? (const void*)g64 : (const void*)g32;
  HIP_TRY(launch_z_r2c(prec, field, spec, geom, plan[2], tw[2], 0, nullptr, st));
  HIP_TRY(launch_y(prec, spec, geom, plan[1], tw[1], -1, 0, nullptr, st));
  HIP_TRY(launch_x_green(prec, spec, geom, plan[0], tw[0], gc, 0, nullptr, st));
  HIP_TRY(launch_y(prec, spec, geom, plan[1], tw[1], +1, 0, nullptr, st));
  HIP_TRY(launch_z_c2r(prec, spec, out, geom, plan[2], tw[2], 1.0 / (double)n, 0, nullptr, st));
  if (prec == 1) HIP_TRY(launch_convert(0, d_dtilde, 1, b + lay.dtil32, 5 * n, st));
  // the host sources of the copies above (hc, the twiddles) live until here
  HIP_TRY(hipStreamSynchronize(st));
  return VPFFT_OK;
}

// ---- per-voxel solver batches --------------------------------------------
namespace {
struct BatchBufs {
  std::vector<void*> ptrs;
  ~BatchBufs() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  void* get(size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    ptrs.push_back(p);
    return p;
  }
};

int batch_const(const vpfft_material* mat, const double L0[25], const double edot[5], vpfft::StepConst* hc) {
  std::memset(hc, 0, sizeof *hc);
  int nmin = mat->rate_sensitivity[0];
  hc->uniform_n = 1;
  for (int k = 0; k < vpfft::kSlips; k++) {
    hc->gdot0[k] = mat->gamma_dot_base[k];
    hc->nexp[k] = mat->rate_sensitivity[k];
    if (hc->nexp[k] != hc->nexp[0]) hc->uniform_n = 0;
    nmin = std::min(nmin, hc->nexp[k]);
  }
  double ee = 0;
  for (int i = 0; i < 5; i++) {
    hc->edot[i] = edot[i];
    ee += edot[i] * edot[i];
  }
  hc->imposed_max_rss = std::max(2.0, 2.0 * std::pow(ee / hc->gdot0[0], 1.0 / (double)(float)nmin));
  if (L0)
    for (int i = 0; i < 5; i++)
      for (int j = 0; j < 5; j++) hc->L0[j * 5 + i] = 0.5 * (L0[j * 5 + i] + L0[i * 5 + j]);
  else
    for (int i = 0; i < 5; i++) hc->L0[i * 5 + i] = 1.0;
  hc->kscale[0] = hc->kscale[1] = hc->kscale[2] = 1.0;
  hc->nx = hc->ny = hc->nz = 2;
  return VPFFT_OK;
}
}  // namespace

int vpfft_solve_constrained_batch(long n, const vpfft_material* mat, const double* orientation, const double* crss,
                                  const double L0[25], const double edot[5], const double* stress_in,
                                  const double* lambda, const double* dtilde, double* stress_out, double* nr_error,
                                  int* nr_steps) {
  if (n < 1 || !mat || !orientation || !crss || !L0 || !edot || !stress_in || !lambda || !dtilde || !stress_out ||
      !nr_error)
    return set_error(VPFFT_E_INVALID, "bad argument");
  int rc = vpfft::check_material(mat);
  if (rc) return rc;
  vpfft::StepConst hc;
  batch_const(mat, L0, edot, &hc);
  BatchBufs B;
  auto* dC = (vpfft::StepConst*)B.get(sizeof hc);
  auto* g64 = (vpfft::GreenConst<double>*)B.get(sizeof(vpfft::GreenConst<double>));
  auto* g32 = (vpfft::GreenConst<float>*)B.get(sizeof(vpfft::GreenConst<float>));
  auto* dR = (double*)B.get(sizeof(double) * 9 * n);
  auto* dT = (double*)B.get(sizeof(double) * 12 * n);
  auto* dS = (double*)B.get(sizeof(double) * 5 * n);
  auto* dL = (double*)B.get(sizeof(double) * 5 * n);
  auto* dD = (double*)B.get(sizeof(double) * 5 * n);
  auto* dO = (double*)B.get(sizeof(double) * 5 * n);
  auto* dE = (double*)B.get(sizeof(double) * n);
  auto* dN = (int*)B.get(sizeof(int) * n);
  if (!dC || !g64 || !g32 || !dR || !dT || !dS || !dL || !dD || !dO || !dE || !dN)
    return set_error(VPFFT_E_NOMEM, "batch allocation failed");
  hipError_t e = hipSuccess;
  auto H2D = [&](void* d, const void* h, size_t b) {
    if (e == hipSuccess) e = hipMemcpy(d, h, b, hipMemcpyHostToDevice);
  };
  H2D(dC, &hc, sizeof hc);
  H2D(dR, orientation, sizeof(double) * 9 * n);
  H2D(dT, crss, sizeof(double) * 12 * n);
  H2D(dS, stress_in, sizeof(double) * 5 * n);
  H2D(dL, lambda, sizeof(double) * 5 * n);
  H2D(dD, dtilde, sizeof(double) * 5 * n);
  H2D(dE, nr_error, sizeof(double) * n);
  if (e == hipSuccess) e = vpfft::launch_derive_from_L0(dC, g64, g32, nullptr);
  if (e == hipSuccess) e = vpfft::launch_constrained_batch(n, dR, dT, dC, dS, dL, dD, dO, dE, dN, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(stress_out, dO, sizeof(double) * 5 * n, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(nr_error, dE, sizeof(double) * n, hipMemcpyDeviceToHost);
  if (e == hipSuccess && nr_steps) e = hipMemcpy(nr_steps, dN, sizeof(int) * n, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return set_error(VPFFT_E_DEVICE, "constrained batch: %s", hipGetErrorString(e));
  return VPFFT_OK;
}

int vpfft_solve_taylor_batch(long n, const vpfft_material* mat, const double* orientation, const double* crss,
                             const double edot[5], const double* stress_in, double* stress_out) {
  if (n < 1 || !mat || !orientation || !crss || !edot || !stress_in || !stress_out)
    return set_error(VPFFT_E_INVALID, "bad argument");
  int rc = vpfft::check_material(mat);
  if (rc) return rc;
  vpfft::StepConst hc;
  batch_const(mat, nullptr, edot, &hc);
  BatchBufs B;
  auto* dC = (vpfft::StepConst*)B.get(sizeof hc);
  auto* dR = (double*)B.get(sizeof(double) * 9 * n);
  auto* dT = (double*)B.get(sizeof(double) * 12 * n);
  auto* dS = (double*)B.get(sizeof(double) * 5 * n);
  auto* dO = (double*)B.get(sizeof(double) * 5 * n);
  if (!dC || !dR || !dT || !dS || !dO) return set_error(VPFFT_E_NOMEM, "batch allocation failed");
  hipError_t e = hipMemcpy(dC, &hc, sizeof hc, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dR, orientation, sizeof(double) * 9 * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dT, crss, sizeof(double) * 12 * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dS, stress_in, sizeof(double) * 5 * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = vpfft::launch_taylor_batch(n, dR, dT, dC, dS, dO, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(stress_out, dO, sizeof(double) * 5 * n, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return set_error(VPFFT_E_DEVICE, "taylor batch: %s", hipGetErrorString(e));
  return VPFFT_OK;
}

namespace {
// Constants of one Solvers call: the material, the strain rate, the
// admissibility bound and L0 as the caller gives them (not symmetrised).
void solver_const(const vpfft_material* mat, const double* L0, const double* edot, double max_rss,
                  vpfft::StepConst* hc) {
  std::memset(hc, 0, sizeof *hc);
  if (mat) {
    hc->uniform_n = 1;
    for (int k = 0; k < vpfft::kSlips; k++) {
      hc->gdot0[k] = mat->gamma_dot_base[k];
      hc->nexp[k] = mat->rate_sensitivity[k];
      if (hc->nexp[k] != hc->nexp[0]) hc->uniform_n = 0;
      hc->tau0[k] = mat->tau0[k];
      hc->tau1[k] = mat->tau1[k];
      hc->theta0[k] = mat->theta0[k];
      hc->theta1[k] = mat->theta1[k];
    }
    for (int e = 0; e < vpfft::kSlips * vpfft::kSlips; e++) hc->hmat[e] = mat->hardening[e];
  }
  if (edot)
    for (int i = 0; i < 5; i++) hc->edot[i] = edot[i];
  if (L0)
    for (int e = 0; e < 25; e++) hc->L0[e] = L0[e];
  hc->imposed_max_rss = max_rss;
}

// One host array of a Solvers call: n x width doubles, copied in (in),
// copied back (out), or both; width 0 marks an array the operation does not
// take.
struct SolverArg {
  const double* in;
  double* out;
  int width;
  double* dev;
};

int solvers_run(const char* what, int op, long n, const vpfft_material* mat, const double* L0, const double* edot,
                double eps, double max_rss, int max_it, double dt, SolverArg* args, int nargs) {
  if (n < 0) return set_error(VPFFT_E_INVALID, "%s: n = %ld", what, n);
  if (n == 0) return VPFFT_OK;
  for (int a = 0; a < nargs; a++)
    if (args[a].width && !args[a].in && !args[a].out)
      return set_error(VPFFT_E_INVALID, "%s: null array argument", what);
  if (mat) {
    int rc = vpfft::check_material(mat);
    if (rc) return rc;
  }
  vpfft::StepConst hc;
  solver_const(mat, L0, edot, max_rss, &hc);
  BatchBufs B;
  auto* dC = (vpfft::StepConst*)B.get(sizeof hc);
  if (!dC) return set_error(VPFFT_E_NOMEM, "%s: allocation failed", what);
  hipError_t e = hipMemcpy(dC, &hc, sizeof hc, hipMemcpyHostToDevice);
  for (int a = 0; a < nargs && e == hipSuccess; a++) {
    if (!args[a].width) continue;
    const size_t bytes = sizeof(double) * args[a].width * (size_t)n;
    args[a].dev = (double*)B.get(bytes);
    if (!args[a].dev) return set_error(VPFFT_E_NOMEM, "%s: allocation of %zu bytes failed", what, bytes);
    if (args[a].in) e = hipMemcpy(args[a].dev, args[a].in, bytes, hipMemcpyHostToDevice);
  }
  // slots, in the order of launch_solvers: schmid, anti, crss, a5, b5, c5, acc, err, out
  double* d[9] = {nullptr};
  for (int a = 0; a < nargs && a < 9; a++) d[a] = args[a].dev;
  if (e == hipSuccess)
    e = vpfft::launch_solvers(op, n, d[0], d[1], d[2], dC, d[3], d[4], d[5], eps, max_rss, max_it, dt, d[6], d[7],
                              d[8], nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  for (int a = 0; a < nargs && e == hipSuccess; a++)
    if (args[a].width && args[a].out) e = hipMemcpy(args[a].out, args[a].dev, sizeof(double) * args[a].width * (size_t)n,
                                   hipMemcpyDeviceToHost);
  if (e != hipSuccess) return set_error(VPFFT_E_DEVICE, "%s: %s", what, hipGetErrorString(e));
  return VPFFT_OK;
}

constexpr SolverArg kUnused = {nullptr, nullptr, 0, nullptr};
SolverArg in_arg(const double* p, int w) { return {p, nullptr, w, nullptr}; }
SolverArg out_arg(double* p, int w) { return {nullptr, p, w, nullptr}; }
SolverArg inout_arg(double* p, int w) { return {p, p, w, nullptr}; }
}  // namespace

int vpfft_solvers_taylor(long n, const vpfft_material* mat, const double* schmid, const double* crss,
                         const double edot[5], const double* stress_in, double eps, double max_rss, int max_it,
                         double* stress_out) {
  if (!mat || !edot) return set_error(VPFFT_E_INVALID, "vpfft_solvers_taylor: null argument");
  SolverArg a[9] = {in_arg(schmid, 60), kUnused, in_arg(crss, 12), in_arg(stress_in, 5), kUnused,
                    kUnused, kUnused, kUnused, out_arg(stress_out, 5)};
  return solvers_run("vpfft_solvers_taylor", vpfft::SOLVER_TAYLOR, n, mat, nullptr, edot, eps, max_rss, max_it, 0.0,
                     a, 9);
}

int vpfft_solvers_constrained(long n, const vpfft_material* mat, const double* schmid, const double* crss,
                              const double edot[5], const double* lambda, const double* dtilde, const double L0[25],
                              const double* stress_in, double eps, double max_rss, int max_it, double* stress_out,
                              double* nr_error) {
  if (!mat || !edot || !L0) return set_error(VPFFT_E_INVALID, "vpfft_solvers_constrained: null argument");
  SolverArg a[9] = {in_arg(schmid, 60), kUnused, in_arg(crss, 12), in_arg(stress_in, 5), in_arg(lambda, 5),
                    in_arg(dtilde, 5), kUnused, inout_arg(nr_error, 1), out_arg(stress_out, 5)};
  return solvers_run("vpfft_solvers_constrained", vpfft::SOLVER_CONSTRAINED, n, mat, L0, edot, eps, max_rss, max_it,
                     0.0, a, 9);
}

int vpfft_solvers_sachs(long n, const double* schmid, const double* crss, const double edot[5], double* out) {
  if (!edot) return set_error(VPFFT_E_INVALID, "vpfft_solvers_sachs: null argument");
  SolverArg a[9] = {in_arg(schmid, 60), kUnused, in_arg(crss, 12), kUnused, kUnused,
                    kUnused, kUnused, kUnused, out_arg(out, 5)};
  return solvers_run("vpfft_solvers_sachs", vpfft::SOLVER_SACHS, n, nullptr, nullptr, edot, 0.0, 0.0, 0, 0.0, a, 9);
}

int vpfft_solvers_apply(long n, const vpfft_material* mat, const double* schmid, const double* crss,
                        const double* stress, double* strain_rate_out) {
  if (!mat) return set_error(VPFFT_E_INVALID, "vpfft_solvers_apply: null argument");
  SolverArg a[9] = {in_arg(schmid, 60), kUnused, in_arg(crss, 12), in_arg(stress, 5), kUnused,
                    kUnused, kUnused, kUnused, out_arg(strain_rate_out, 5)};
  return solvers_run("vpfft_solvers_apply", vpfft::SOLVER_APPLY, n, mat, nullptr, nullptr, 0.0, 0.0, 0, 0.0, a, 9);
}

int vpfft_solvers_shear_rate(long n, const vpfft_material* mat, const double* schmid, const double* anti,
                             const double* crss, const double* stress, double* shear_rate_out) {
  if (!mat) return set_error(VPFFT_E_INVALID, "vpfft_solvers_shear_rate: null argument");
  SolverArg a[9] = {in_arg(schmid, 60), in_arg(anti, 108), in_arg(crss, 12), in_arg(stress, 5), kUnused,
                    kUnused, kUnused, kUnused, out_arg(shear_rate_out, 9)};
  return solvers_run("vpfft_solvers_shear_rate", vpfft::SOLVER_SHEAR_RATE, n, mat, nullptr, nullptr, 0.0, 0.0, 0,
                     0.0, a, 9);
}

int vpfft_solvers_hardening(long n, const vpfft_material* mat, const double* schmid, const double* crss,
                            const double* stress, double dt, double* accumulated_shear, double* crss_out) {
  if (!mat) return set_error(VPFFT_E_INVALID, "vpfft_solvers_hardening: null argument");
  SolverArg a[9] = {in_arg(schmid, 60), kUnused, in_arg(crss, 12), in_arg(stress, 5), kUnused,
                    kUnused, inout_arg(accumulated_shear, 1), kUnused, out_arg(crss_out, 12)};
  return solvers_run("vpfft_solvers_hardening", vpfft::SOLVER_HARDENING, n, mat, nullptr, nullptr, 0.0, 0.0, 0, dt,
                     a, 9);
}

int vpfft_batch_create(void** out_batch) {
  if (!out_batch) return set_error(VPFFT_E_INVALID, "null argument");
  vpfft::Batch* b = nullptr;
  int rc = vpfft::batch_create(&b);
  *out_batch = rc == VPFFT_OK ? (void*)b : nullptr;
  return rc;
}

int vpfft_batch_destroy(void* batch) {
  vpfft::batch_destroy(static_cast<vpfft::Batch*>(batch));
  return VPFFT_OK;
}

int vpfft_batch_set_tuning(void* batch, int level) {
  if (!batch) return set_error(VPFFT_E_INVALID, "null batch");
  if (level < 0 || level > 2) return set_error(VPFFT_E_INVALID, "tuning level must be 0, 1 or 2");
  vpfft::Batch* b = static_cast<vpfft::Batch*>(batch);
  if (level != b->tune) b->tuned.clear();
  b->tune = level;
  return VPFFT_OK;
}

int vpfft_batch_block_shape(void* batch, const void* grid, int pass, int* block_out) {
  vpfft::Batch* b = static_cast<vpfft::Batch*>(batch);
  const Grid* g = static_cast<const Grid*>(grid);
  if (!b || !g || !block_out || pass < 0 || pass > 2) return set_error(VPFFT_E_INVALID, "bad argument");
  const auto key = std::make_tuple(g->prec, pass, g->nx, g->ny, g->nz);
  auto it = b->tuned.find(key);
  *block_out = (b->tune && it != b->tuned.end()) ? it->second.B : vpfft::default_sched(g, (vpfft::FftPass)pass).B;
  return VPFFT_OK;
}

int vpfft_batch_radix_plan(void* batch, const void* grid, int pass, int* radices, int* nstages) {
  vpfft::Batch* b = static_cast<vpfft::Batch*>(batch);
  const Grid* g = static_cast<const Grid*>(grid);
  if (!b || !g || !radices || !nstages || pass < 0 || pass > 2) return set_error(VPFFT_E_INVALID, "bad argument");
  const auto key = std::make_tuple(g->prec, pass, g->nx, g->ny, g->nz);
  auto it = b->tuned.find(key);
  const vpfft::Sched sc = (b->tune && it != b->tuned.end()) ? it->second : vpfft::default_sched(g, (vpfft::FftPass)pass);
  const int axis = pass == 0 ? 2 : (pass == 1 ? 1 : 0);  // pass 0 z, 1 y, 2 x
  const int n = axis == 0 ? g->nx : (axis == 1 ? g->ny : g->nz);
  if (sc.B < 0) {  // the register form: its compile-time radix chain
    *nstages = vpfft::reg_form_radices(n, radices);
    return VPFFT_OK;
  }
  vpfft::AxisPlan ap;
  if (!vpfft::make_axis_plan_ordered(n, vpfft::kRadixOrders[sc.order], &ap)) return set_error(VPFFT_E_STATE, "no plan");
  for (int s = 0; s < ap.nstages; s++) radices[s] = ap.radix[s];
  *nstages = ap.nstages;
  return VPFFT_OK;
}

int vpfft_batch_run_strain_step(void* batch, void* const* grids, int count, const double* edot, int max_iter,
                                const double* dt, double tol, double* avg_stress_out, vpfft_step_info* infos) {
  vpfft::Batch* b = static_cast<vpfft::Batch*>(batch);
  if (!b) return set_error(VPFFT_E_INVALID, "null batch");
  std::vector<Grid*> gs(count > 0 ? count : 0);
  for (int j = 0; j < count; j++) gs[j] = static_cast<Grid*>(grids[j]);
  return vpfft::batch_step(b, gs.data(), count, edot, max_iter, dt, tol, avg_stress_out, infos);
}

int vpfft_grid_bytes(const vpfft_grid_desc* desc, uint64_t* bytes_out) {
  if (!desc || !bytes_out || desc->nx < 2 || desc->ny < 2 || desc->nz < 2)
    return set_error(VPFFT_E_INVALID, "bad argument");
  *bytes_out = (uint64_t)vpfft::grid_bytes(desc);
  return VPFFT_OK;
}

int vpfft_grid_save_state(void* grid) {
  Grid* g = as_grid(grid);
  if (!g) return set_error(VPFFT_E_INVALID, "null grid");
  if (!g->uploaded) return set_error(VPFFT_E_STATE, "save_state before upload");
  const long n = g->n;
  if (!g->saved) {
    void* p = nullptr;
    int rc = vpfft::dalloc(g, &p, sizeof(double) * 32 * n);
    if (rc) return rc;
    g->saved = (double*)p;
  }
  double* d = g->saved;
  const std::pair<double*, long> parts[5] = {{g->R, 9}, {g->crss, 12}, {g->acc, 1}, {g->sigma, 5}, {g->lambda, 5}};
  for (const auto& pr : parts) {
    if (hipMemcpyAsync(d, pr.first, sizeof(double) * pr.second * n, hipMemcpyDeviceToDevice, g->st) != hipSuccess)
      return set_error(VPFFT_E_DEVICE, "save_state copy failed");
    d += pr.second * n;
  }
  if (hipStreamSynchronize(g->st) != hipSuccess) return set_error(VPFFT_E_DEVICE, "save_state sync failed");
  g->saved_len[0] = g->lx;
  g->saved_len[1] = g->ly;
  g->saved_len[2] = g->lz;
  g->saved_crss_uniform = g->crss_uniform;
  g->has_saved = true;
  return VPFFT_OK;
}

int vpfft_grid_restore_state(void* grid) {
  Grid* g = as_grid(grid);
  if (!g) return set_error(VPFFT_E_INVALID, "null grid");
  if (!g->has_saved) return set_error(VPFFT_E_STATE, "restore_state without a saved state");
  const long n = g->n;
  const double* d = g->saved;
  const std::pair<double*, long> parts[5] = {{g->R, 9}, {g->crss, 12}, {g->acc, 1}, {g->sigma, 5}, {g->lambda, 5}};
  for (const auto& pr : parts) {
    if (hipMemcpyAsync(pr.first, d, sizeof(double) * pr.second * n, hipMemcpyDeviceToDevice, g->st) != hipSuccess)
      return set_error(VPFFT_E_DEVICE, "restore_state copy failed");
    d += pr.second * n;
  }
  if (hipStreamSynchronize(g->st) != hipSuccess) return set_error(VPFFT_E_DEVICE, "restore_state sync failed");
  g->lx = g->saved_len[0];
  g->ly = g->saved_len[1];
  g->lz = g->saved_len[2];
  g->crss_uniform = g->saved_crss_uniform;
  return VPFFT_OK;
}

int vpfft_grid_set_block_shape(void* grid, int precision, int pass, int block) {
  Grid* g = as_grid(grid);
  if (!g || pass < 0 || pass > 2 || (precision != VPFFT_FP64 && precision != VPFFT_MIXED) || block < 0)
    return set_error(VPFFT_E_INVALID, "bad argument");
  if (block > 0) {
    int cand[8];
    const int nc = vpfft::pass_candidates((vpfft::FftPass)pass, cand, 8);
    bool listed = false;
    for (int c = 0; c < nc; c++) listed = listed || cand[c] == block;
    if (!listed || !vpfft::pass_shape(precision, (vpfft::FftPass)pass, g->geom, block).ok)
      return set_error(VPFFT_E_INVALID, "block shape %d not available for pass %d of this grid", block, pass);
  }
  g->fftB[precision][pass] = block;
  g->gkey.prec = -1;  // recapture the loop graph
  return VPFFT_OK;
}

}  // extern "C"



