// dugma_ctx.hip -- context lifetime: init (init.cpp / init.cu), fr_m
// (fr_m.cpp / fr_m.cu), error reporting and the energy-path point records.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <limits>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "dugma_internal.h"
#include "dugma_pairs.h"

namespace dugma {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg) {
  set_error(msg);
  return code;
}

int hip_fail(hipError_t e, const char *what) {
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  set_error(m);
  return e == hipErrorOutOfMemory ? DUGMA_ERR_ALLOC : DUGMA_ERR_DEVICE;
}

int DevBuf::ensure(size_t need) {
  if (need <= bytes) return DUGMA_OK;
  if (ptr) {
    hipError_t e = hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
    if (e != hipSuccess) return hip_fail(e, "hipFree(scratch)");
  }
  size_t want = need + need / 4 + 4096;
  hipError_t e = hipMalloc(&ptr, want);
  if (e != hipSuccess) {
    ptr = nullptr;
    return hip_fail(e, "hipMalloc(scratch)");
  }
  bytes = want;
  return DUGMA_OK;
}

PairTimer::PairTimer(dugma_ctx *ctx) : c(ctx) {
  if (c->timing) (void)hipEventRecord(c->ev0, c->stream);
}

void PairTimer::stop() {
  if (c->timing) (void)hipEventRecord(c->ev1, c->stream);
}

void PairTimer::finish() {
  if (!c->timing) return;
  (void)hipEventSynchronize(c->ev1);
  float ms = 0.0f;
  if (hipEventElapsedTime(&ms, c->ev0, c->ev1) == hipSuccess) {
    c->pair_ms += ms;
    c->pair_launches += 1;
  }
}

template <class F>
static void preset_nan(F *slots, int n) {
  volatile F *v = slots;
  for (int a = 0; a < n; ++a) v[a] = std::numeric_limits<F>::quiet_NaN();
  std::atomic_thread_fence(std::memory_order_seq_cst);
}
void preset_slots(double *slots, int n) { preset_nan(slots, n); }
void preset_slots(float *slots, int n) { preset_nan(slots, n); }

template <class F>
static int wait_nan(dugma_ctx *c, const F *slots, int n) {
  if (!c->timing) {
    const volatile F *v = slots;
    const auto t0 = std::chrono::steady_clock::now();
    for (unsigned it = 1;; ++it) {
      int set = 0;
      for (int a = 0; a < n; ++a) set += v[a] == v[a] ? 1 : 0;  // not NaN
      if (set == n) {
        std::atomic_thread_fence(std::memory_order_acquire);
        return DUGMA_OK;
      }
      if ((it & 1023u) == 0 &&
          std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(kSpinUs))
        break;
      __builtin_ia32_pause();
    }
  }
  DUGMA_HIP(hipStreamSynchronize(c->stream));
  return DUGMA_OK;
}
int wait_slots(dugma_ctx *c, const double *slots, int n) { return wait_nan(c, slots, n); }
int wait_slots(dugma_ctx *c, const float *slots, int n) { return wait_nan(c, slots, n); }

void DevBuf::release() {
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
}

int PinBuf::ensure(size_t need) {
  if (need <= bytes) return DUGMA_OK;
  if (ptr) {
    hipError_t e = hipHostFree(ptr);
    ptr = nullptr;
    bytes = 0;
    if (e != hipSuccess) return hip_fail(e, "hipHostFree(staging)");
  }
  size_t want = need + need / 4 + 4096;
  hipError_t e = hipHostMalloc(&ptr, want, hipHostMallocDefault);
  if (e != hipSuccess) {
    ptr = nullptr;
    return hip_fail(e, "hipHostMalloc(staging)");
  }
  bytes = want;
  return DUGMA_OK;
}

void PinBuf::release() {
  if (ptr) (void)hipHostFree(ptr);
  ptr = nullptr;
  bytes = 0;
}

// ---------------------------------------------------------------------------
// init.cu:22-234 restated: FP64 -> FP32 points and covariances, then the
// adjugate inverse and g = sqrtf(1/det) in FP32, operation by operation in
// the reference's order (this file is compiled without FP contraction, and
// HIP's FP32 division and sqrtf are correctly rounded).
// ---------------------------------------------------------------------------
template <int D>
__global__ void k_prep_compat(const double *__restrict__ P,
                              const double *__restrict__ C, int n,
                              float *__restrict__ Pf,
                              float *__restrict__ Cinv,
                              float *__restrict__ g) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  for (int a = 0; a < D; ++a) Pf[(size_t)k * D + a] = (float)P[(size_t)k * D + a];
  const double *c = C + (size_t)k * D * D;
  float a32[D * D];
  for (int q = 0; q < D * D; ++q) a32[q] = (float)c[q];
  compat_inverse<D>(a32, Cinv + (size_t)k * D * D, g + k);
}

// Validity rule of the convolved energy: finite coordinates, finite
// covariance with positive leading principal minors (FP64).
static bool valid_point(int D, const double *p, const double *S) {
  for (int a = 0; a < D; ++a)
    if (!std::isfinite(p[a])) return false;
  for (int a = 0; a < D * D; ++a)
    if (!std::isfinite(S[a])) return false;
  if (!(S[0] > 0)) return false;
  double m2 = S[0] * S[D + 1] - S[1] * S[D];
  if (!(m2 > 0)) return false;
  if (D == 3) {
    double det = S[0] * (S[4] * S[8] - S[7] * S[5]) -
                 S[3] * (S[1] * S[8] - S[7] * S[2]) +
                 S[6] * (S[1] * S[5] - S[4] * S[2]);
    if (!(det > 0)) return false;
  }
  return true;
}

static bool isotropic(int D, const double *S) {
  if (D == 2) return S[1] == 0 && S[2] == 0 && S[0] == S[3];
  return S[1] == 0 && S[2] == 0 && S[3] == 0 && S[5] == 0 && S[6] == 0 &&
         S[7] == 0 && S[0] == S[4] && S[0] == S[8];
}

static double sym(const double *S, int D, int a, int b) {
  return 0.5 * (S[a + D * b] + S[b + D * a]);
}

// symmetric part of the FP32-rounded covariance
static double sym32(const double *S, int D, int a, int b) {
  return 0.5 * ((double)(float)S[a + D * b] + (double)(float)S[b + D * a]);
}

static double f32(double v) { return (double)(float)v; }

// Host-built target records (the target cloud does not move).  Inputs are
// rounded to FP32 first (as init.cu:192-234 does), then scaled by a power of
// two, which is exact.
static void make_target_records(int D, const std::vector<int> &keep,
                                const double *X, const double *XC, double sc,
                                std::vector<char> &full, std::vector<char> &iso,
                                bool want_iso) {
  size_t n = keep.size();
  const double sc2 = sc * sc;
  if (D == 3) {
    full.assign(n * sizeof(Rec3F), 0);
    Rec3F *r = reinterpret_cast<Rec3F *>(full.data());
    Rec3I *ri = nullptr;
    if (want_iso) {
      iso.assign(n * sizeof(Rec3I), 0);
      ri = reinterpret_cast<Rec3I *>(iso.data());
    }
    for (size_t k = 0; k < n; ++k) {
      const double *p = X + (size_t)keep[k] * 3;
      const double *S = XC + (size_t)keep[k] * 9;
      r[k].hx = (float)(f32(p[0]) * sc);
      r[k].hy = (float)(f32(p[1]) * sc);
      r[k].hz = (float)(f32(p[2]) * sc);
      r[k].s11 = (float)(f32(S[0]) * sc2);
      r[k].s22 = (float)(f32(S[4]) * sc2);
      r[k].s33 = (float)(f32(S[8]) * sc2);
      r[k].s12 = (float)(sym32(S, 3, 0, 1) * sc2);
      r[k].s13 = (float)(sym32(S, 3, 0, 2) * sc2);
      r[k].s23 = (float)(sym32(S, 3, 1, 2) * sc2);
      if (ri) {
        ri[k].hx = r[k].hx; ri[k].hy = r[k].hy; ri[k].hz = r[k].hz;
        ri[k].s = (float)((double)r[k].s11 * kLn4);  // isotropic records: ln 4 times the variance
      }
    }
  } else {
    full.assign(n * sizeof(Rec2F), 0);
    Rec2F *r = reinterpret_cast<Rec2F *>(full.data());
    Rec2I *ri = nullptr;
    if (want_iso) {
      iso.assign(n * sizeof(Rec2I), 0);
      ri = reinterpret_cast<Rec2I *>(iso.data());
    }
    for (size_t k = 0; k < n; ++k) {
      const double *p = X + (size_t)keep[k] * 2;
      const double *S = XC + (size_t)keep[k] * 4;
      r[k].hx = (float)(f32(p[0]) * sc);
      r[k].hy = (float)(f32(p[1]) * sc);
      r[k].s11 = (float)(f32(S[0]) * sc2);
      r[k].s22 = (float)(f32(S[3]) * sc2);
      r[k].s12 = (float)(sym32(S, 2, 0, 1) * sc2);
      if (ri) {
        ri[k].hx = r[k].hx; ri[k].hy = r[k].hy;
        ri[k].s = (float)((double)r[k].s11 * kLn4);  // isotropic records: ln 4 times the variance
      }
    }
  }
}

int refresh_host_sources(dugma_ctx *c) {
  if (!c->hSsrc_stale) return DUGMA_OK;
  if (!c->hSsrc.empty())
    DUGMA_HIP(hipMemcpyAsync(c->hSsrc.data(), c->dSsrc, sizeof(double) * c->hSsrc.size(),
                             hipMemcpyDeviceToHost, c->stream));
  DUGMA_HIP(hipStreamSynchronize(c->stream));
  c->hSsrc_stale = false;
  return DUGMA_OK;
}

}  // namespace dugma

using namespace dugma;

static void free_ctx(dugma_ctx *c) {
  if (!c) return;
  void *ptrs[] = {c->dX, c->dXCinv, c->dY, c->dYCinv, c->dgX, c->dgY,
                  c->dTidx, c->dSidx, c->dTrecF, c->dTrecI, c->dSsrc,
                  c->dSrecF, c->dSrecI};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  c->part.release();
  c->scratch2.release();
  c->sigbuf.release();
  c->hstage.release();
  if (c->hpinned) (void)hipHostFree(c->hpinned);
  if (c->hparts) (void)hipHostFree(c->hparts);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->evpts) (void)hipEventDestroy(c->evpts);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

template <class T>
static int dmalloc(T **p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void **)p, count * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return hip_fail(e, "hipMalloc");
  }
  return DUGMA_OK;
}

extern "C" int dugma_init(const double *X, const double *X_C, const double *Y,
                          const double *Y_C, int N, int M, int D,
                          dugma_ctx **out, double *ok) {
  if (out) *out = nullptr;
  if (ok) *ok = 0.0;
  if (!X || !X_C || !Y || !Y_C || !out)
    return fail(DUGMA_ERR_ARG, "dugma_init: null pointer argument");
  if (D != 2 && D != 3) return fail(DUGMA_ERR_ARG, "dugma_init: D must be 2 or 3");
  if (N < 1 || M < 1) return fail(DUGMA_ERR_ARG, "dugma_init: N and M must be >= 1");
  if (N > DUGMA_MAX_POINTS || M > DUGMA_MAX_POINTS)
    return fail(DUGMA_ERR_SIZE, "dugma_init: the number of points exceeds the maximum");

  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(DUGMA_ERR_DEVICE, "dugma_init: no HIP device available");

  dugma_ctx *c = new dugma_ctx();
  c->N = N;
  c->M = M;
  c->D = D;
  int rc = DUGMA_OK;
#define TRY(x)                      \
  do {                              \
    rc = (x);                       \
    if (rc != DUGMA_OK) goto fail_; \
  } while (0)
  {
    hipError_t e = hipGetDevice(&c->device);
    if (e != hipSuccess) TRY(hip_fail(e, "hipGetDevice"));
    TRY(noalloc_setup(c->device));
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) TRY(hip_fail(e, "hipStreamCreateWithFlags"));
    // coherent: result slots are polled by the host while kernels run
    e = hipHostMalloc((void **)&c->hpinned, sizeof(double) * dugma_ctx::kHostStage,
                      hipHostMallocCoherent);
    if (e != hipSuccess) {
      c->hpinned = nullptr;
      TRY(hip_fail(e, "hipHostMalloc"));
    }
    c->hres = c->hpinned + dugma_ctx::kHostStage / 2;
    e = hipHostGetDevicePointer((void **)&c->dres, c->hres, 0);
    if (e != hipSuccess) TRY(hip_fail(e, "hipHostGetDevicePointer"));
    e = hipHostMalloc((void **)&c->hparts, sizeof(double) * kPosePart * dugma_ctx::kHostParts,
                      hipHostMallocDefault);
    if (e != hipSuccess) {
      c->hparts = nullptr;
      TRY(hip_fail(e, "hipHostMalloc(parts)"));
    }
    e = hipHostGetDevicePointer((void **)&c->dparts, c->hparts, 0);
    if (e != hipSuccess) TRY(hip_fail(e, "hipHostGetDevicePointer(parts)"));


    // ---- reference state (init.cu:264-331) ----
    const int DD = D * D;
    TRY(dmalloc(&c->dX, (size_t)N * D));
    TRY(dmalloc(&c->dXCinv, (size_t)N * DD));
    TRY(dmalloc(&c->dY, (size_t)M * D));
    TRY(dmalloc(&c->dYCinv, (size_t)M * DD));
    TRY(dmalloc(&c->dgX, (size_t)N));
    TRY(dmalloc(&c->dgY, (size_t)M));
    {
      double *st = nullptr;
      size_t nst = (size_t)N * (D + DD) + (size_t)M * (D + DD);
      TRY(dmalloc(&st, nst));
      double *sX = st, *sXC = sX + (size_t)N * D, *sY = sXC + (size_t)N * DD,
             *sYC = sY + (size_t)M * D;
      hipError_t e2 = hipMemcpyAsync(sX, X, sizeof(double) * N * D, hipMemcpyHostToDevice, c->stream);
      if (e2 == hipSuccess) e2 = hipMemcpyAsync(sXC, X_C, sizeof(double) * N * DD, hipMemcpyHostToDevice, c->stream);
      if (e2 == hipSuccess) e2 = hipMemcpyAsync(sY, Y, sizeof(double) * M * D, hipMemcpyHostToDevice, c->stream);
      if (e2 == hipSuccess) e2 = hipMemcpyAsync(sYC, Y_C, sizeof(double) * M * DD, hipMemcpyHostToDevice, c->stream);
      if (e2 == hipSuccess) {
        const int B = 256;
        if (D == 3) {
          hipLaunchKernelGGL(k_prep_compat<3>, dim3((N + B - 1) / B), dim3(B), 0, c->stream, sX, sXC, N, c->dX, c->dXCinv, c->dgX);
          hipLaunchKernelGGL(k_prep_compat<3>, dim3((M + B - 1) / B), dim3(B), 0, c->stream, sY, sYC, M, c->dY, c->dYCinv, c->dgY);
        } else {
          hipLaunchKernelGGL(k_prep_compat<2>, dim3((N + B - 1) / B), dim3(B), 0, c->stream, sX, sXC, N, c->dX, c->dXCinv, c->dgX);
          hipLaunchKernelGGL(k_prep_compat<2>, dim3((M + B - 1) / B), dim3(B), 0, c->stream, sY, sYC, M, c->dY, c->dYCinv, c->dgY);
        }
        e2 = hipGetLastError();
      }
      if (e2 == hipSuccess) e2 = hipStreamSynchronize(c->stream);
      (void)hipFree(st);
      if (e2 != hipSuccess) TRY(hip_fail(e2, "dugma_init: upload/prep"));
    }

    // ---- convolved-energy state ----
    // The energy path works on the FP32 points and covariances (as the
    // reference's device state does), so validity is judged on those.
    std::vector<int> keepT, keepS;
    bool isoT = true, isoS = true;
    double logv = 0.0;
    long nlog = 0;
    auto consider = [&](const double *P, const double *C, int n, std::vector<int> &keep, bool &iso) {
      double p32[3], c32[9];
      for (int i = 0; i < n; ++i) {
        for (int a = 0; a < D; ++a) p32[a] = f32(P[(size_t)i * D + a]);
        for (int a = 0; a < DD; ++a) c32[a] = f32(C[(size_t)i * DD + a]);
        if (!valid_point(D, p32, c32)) continue;
        keep.push_back(i);
        iso = iso && isotropic(D, c32);
        double tr = 0;
        for (int a = 0; a < D; ++a) tr += c32[a * (D + 1)];
        logv += std::log2(tr / D);
        ++nlog;
      }
    };
    consider(X, X_C, N, keepT, isoT);
    consider(Y, Y_C, M, keepS, isoS);
    c->nT = (int)keepT.size();
    c->nS = (int)keepS.size();
    c->isoT = isoT;
    c->isoS = isoS;
    c->iso = isoT && isoS && c->nT > 0 && c->nS > 0;
    // s = 2^k with s^2 * (geometric-mean variance) ~ 1
    int k = nlog ? (int)std::lround(-0.5 * logv / (double)nlog) : 0;
    k = std::max(-60, std::min(60, k));
    c->lscale = k;
    c->scale = std::ldexp(1.0, k);
    const double sc = c->scale, sc2 = sc * sc;
    std::vector<char> recF, recI;
    make_target_records(D, keepT, X, X_C, sc, recF, recI, c->iso);
    const size_t recsz = (D == 3) ? sizeof(Rec3F) : sizeof(Rec2F);
    const size_t recszI = (D == 3) ? sizeof(Rec3I) : sizeof(Rec2I);
    const size_t nTa = (size_t)std::max(c->nT, 1), nSa = (size_t)std::max(c->nS, 1);
    TRY(dmalloc(&c->dTidx, nTa));
    TRY(dmalloc(&c->dSidx, nSa));
    TRY(dmalloc((char **)&c->dTrecF, nTa * recsz));
    TRY(dmalloc((char **)&c->dSrecF, nSa * recsz));
    if (c->iso) {
      TRY(dmalloc((char **)&c->dTrecI, nTa * recszI));
      TRY(dmalloc((char **)&c->dSrecI, nSa * recszI));
    }
    // source data: FP32-rounded point (D) then symmetric covariance (D*D)
    // per kept point, scaled, held in FP64 for the per-call transform
    std::vector<double> src((size_t)c->nS * (D + DD));
    for (int q = 0; q < c->nS; ++q) {
      const double *p = Y + (size_t)keepS[q] * D;
      const double *S = Y_C + (size_t)keepS[q] * DD;
      double *o = src.data() + (size_t)q * (D + DD);
      for (int a = 0; a < D; ++a) o[a] = f32(p[a]) * sc;
      for (int a = 0; a < D; ++a)
        for (int b = 0; b < D; ++b) o[D + a + D * b] = sym32(S, D, a, b) * sc2;
    }
    TRY(dmalloc(&c->dSsrc, std::max<size_t>(src.size(), 1)));
    hipError_t e3 = hipSuccess;
    if (c->nT > 0) {
      e3 = hipMemcpy(c->dTidx, keepT.data(), sizeof(int) * c->nT, hipMemcpyHostToDevice);
      if (e3 == hipSuccess) e3 = hipMemcpy(c->dTrecF, recF.data(), recF.size(), hipMemcpyHostToDevice);
      if (e3 == hipSuccess && c->iso) e3 = hipMemcpy(c->dTrecI, recI.data(), recI.size(), hipMemcpyHostToDevice);
    }
    if (e3 == hipSuccess && c->nS > 0) {
      e3 = hipMemcpy(c->dSidx, keepS.data(), sizeof(int) * c->nS, hipMemcpyHostToDevice);
      if (e3 == hipSuccess) e3 = hipMemcpy(c->dSsrc, src.data(), sizeof(double) * src.size(), hipMemcpyHostToDevice);
    }
    if (e3 != hipSuccess) TRY(hip_fail(e3, "dugma_init: energy-state upload"));
    c->hTidx = std::move(keepT);
    c->hSidx = std::move(keepS);
    c->hSsrc = std::move(src);
  }
#undef TRY
  *out = c;
  if (ok) *ok = 1.0;
  return DUGMA_OK;
fail_:
  free_ctx(c);
  return rc;
}

extern "C" int dugma_fr_m(dugma_ctx *ctx, double *ok) {
  if (ok) *ok = 0.0;
  if (!ctx) return fail(DUGMA_ERR_ARG, "dugma_fr_m: null context");
  (void)hipStreamSynchronize(ctx->stream);
  free_ctx(ctx);
  if (ok) *ok = 1.0;
  return DUGMA_OK;
}

extern "C" int dugma_ctx_dims(const dugma_ctx *ctx, int *N, int *M, int *D) {
  if (!ctx) return fail(DUGMA_ERR_ARG, "dugma_ctx_dims: null context");
  if (N) *N = ctx->N;
  if (M) *M = ctx->M;
  if (D) *D = ctx->D;
  return DUGMA_OK;
}

extern "C" int dugma_ctx_energy_info(const dugma_ctx *ctx, int *n_target,
                                     int *n_source, int *isotropic) {
  if (!ctx) return fail(DUGMA_ERR_ARG, "dugma_ctx_energy_info: null context");
  if (n_target) *n_target = ctx->nT;
  if (n_source) *n_source = ctx->nS;
  if (isotropic) *isotropic = ctx->iso ? 1 : 0;
  return DUGMA_OK;
}

extern "C" void *dugma_ctx_stream(const dugma_ctx *ctx) {
  return ctx ? (void *)ctx->stream : nullptr;
}

extern "C" int dugma_ctx_handles(const dugma_ctx *ctx, void **ptrs) {
  if (!ctx || !ptrs) return fail(DUGMA_ERR_ARG, "dugma_ctx_handles: null argument");
  ptrs[0] = ctx->dX;
  ptrs[1] = ctx->dXCinv;
  ptrs[2] = ctx->dY;
  ptrs[3] = ctx->dYCinv;
  ptrs[4] = ctx->dgX;
  ptrs[5] = ctx->dgY;
  return DUGMA_OK;
}

extern "C" int dugma_read_handle(dugma_ctx *ctx, int which, float *out) {
  if (!ctx || !out) return fail(DUGMA_ERR_ARG, "dugma_read_handle: null argument");
  const size_t N = ctx->N, M = ctx->M, D = ctx->D;
  const float *src[6] = {ctx->dX, ctx->dXCinv, ctx->dY, ctx->dYCinv, ctx->dgX, ctx->dgY};
  const size_t cnt[6] = {N * D, N * D * D, M * D, M * D * D, N, M};
  if (which < 0 || which > 5) return fail(DUGMA_ERR_ARG, "dugma_read_handle: which must be 0..5");
  DUGMA_HIP(hipMemcpyAsync(out, src[which], cnt[which] * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  DUGMA_HIP(hipStreamSynchronize(ctx->stream));
  return DUGMA_OK;
}

extern "C" int dugma_ctx_set_timing(dugma_ctx *ctx, int on) {
  if (!ctx) return fail(DUGMA_ERR_ARG, "dugma_ctx_set_timing: null context");
  if (on && !ctx->ev0) {
    DUGMA_HIP(hipEventCreate(&ctx->ev0));
    DUGMA_HIP(hipEventCreate(&ctx->ev1));
  }
  ctx->timing = on != 0;
  ctx->pair_ms = 0.0;
  ctx->pair_launches = 0;
  return DUGMA_OK;
}

extern "C" int dugma_ctx_pair_time(const dugma_ctx *ctx, double *ms,
                                   long long *launches) {
  if (!ctx) return fail(DUGMA_ERR_ARG, "dugma_ctx_pair_time: null context");
  if (ms) *ms = ctx->pair_ms;
  if (launches) *launches = ctx->pair_launches;
  return DUGMA_OK;
}

extern "C" const char *dugma_last_error(void) { return g_last_error.c_str(); }

extern "C" const char *dugma_version(void) { return "dugma-gfx950 0.1"; }
