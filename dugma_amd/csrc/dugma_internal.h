// dugma_internal.h -- shared definitions of the gfx950 DUGMA library.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstddef>
#include <cstdint>
#include <string>
#include <thread>
#include <vector>

#include "../../include/dugma.h"

namespace dugma {

// ---------------------------------------------------------------------------
// Point records.  A point's position is kept as an FP32 pair p = hi + lo
// (hi = fl32(p), lo = fl32(p - hi)), so pair differences of nearby points
// keep ~48 bits although every per-pair operation is FP32.  The records are
// read wave-uniformly by the pair kernels (scalar loads into SGPRs), so each
// is a power-of-two size and 16-byte aligned.
// ---------------------------------------------------------------------------
struct alignas(16) Rec3F {  // 3-D, full symmetric covariance, 64 B
  float hx, hy, hz, s11;
  float lx, ly, lz, s22;
  float s12, s13, s23, s33;
  float pad0, pad1, pad2, pad3;
};
struct alignas(16) Rec3I {  // 3-D, isotropic covariance s*I, 32 B
  float hx, hy, hz, s;
  float lx, ly, lz, pad0;
};
struct alignas(16) Rec2F {  // 2-D, full symmetric covariance, 32 B
  float hx, hy, lx, ly;
  float s11, s12, s22, pad0;
};
struct alignas(16) Rec2I {  // 2-D, isotropic covariance s*I, 32 B
  float hx, hy, lx, ly;
  float s, pad0, pad1, pad2;
};

// Per-stationary-point accumulator layout of the pair kernels (FP64):
//  [0]      energy sum_j e_ij                    (without the (2pi)^-D/2 factor)
//  [1..D]   EW = sum_j e_ij w_ij,  w = S^-1 d,  d = p - q
//  full covariance:  B = sum_j e (w w^T - S^-1)  (D(D+1)/2 entries, upper
//                    triangle row by row)
//  isotropic:        G = sum_j (e/s)(|d|^2/s - D)
enum { kAcc3F = 10, kAcc3I = 5, kAcc2F = 6, kAcc2I = 4 };

// Pose-gradient partial per workgroup: [E, dofs...], padded to 8 doubles.
enum { kPosePart = 8 };

// init.cu:22-190 for one point (detX_gpu / detY_gpu): the FP32 column-major
// covariance a -> its inverse o (adjugate / det, in the reference's
// operation order; callers are compiled without FP contraction, and HIP's
// FP32 division and sqrtf are correctly rounded) and g = sqrtf(1/det);
// det <= 0 or NaN: inverse 100000 in every entry, g = 0 (init.cu:48-57,86-100).
template <int D>
__device__ __forceinline__ void compat_inverse(const float *a, float *o, float *g) {
  if (D == 2) {
    const float a1 = a[0], a2 = a[1], a3 = a[2], a4 = a[3];
    const float d = a1 * a4 - a2 * a3;
    if (d > 0) {
      o[0] = a4 / d;
      o[1] = (-a2) / d;
      o[2] = (-a3) / d;
      o[3] = a1 / d;
      *g = sqrtf(1.0f / d);
    } else {
      for (int q = 0; q < 4; ++q) o[q] = 100000.0f;
      *g = 0.0f;
    }
  } else {
    const float a11 = a[0], a21 = a[1], a31 = a[2];
    const float a12 = a[3], a22 = a[4], a32 = a[5];
    const float a13 = a[6], a23 = a[7], a33 = a[8];
    const float d = a11 * a22 * a33 - a11 * a23 * a32 - a12 * a21 * a33 +
                    a12 * a23 * a31 + a13 * a21 * a32 - a13 * a22 * a31;
    if (d > 0) {
      o[0] = (a22 * a33 - a23 * a32) / d;
      o[1] = (-(a21 * a33 - a23 * a31)) / d;
      o[2] = (a21 * a32 - a22 * a31) / d;
      o[3] = (-(a12 * a33 - a13 * a32)) / d;
      o[4] = (a11 * a33 - a13 * a31) / d;
      o[5] = (-(a11 * a32 - a12 * a31)) / d;
      o[6] = (a12 * a23 - a13 * a22) / d;
      o[7] = (-(a11 * a23 - a13 * a21)) / d;
      o[8] = (a11 * a22 - a12 * a21) / d;
      *g = sqrtf(1.0f / d);
    } else {
      for (int q = 0; q < 9; ++q) o[q] = 100000.0f;
      *g = 0.0f;
    }
  }
}

// ---------------------------------------------------------------------------
// Error plumbing
// ---------------------------------------------------------------------------
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);

#define DUGMA_HIP(call)                                   \
  do {                                                    \
    hipError_t e_ = (call);                               \
    if (e_ != hipSuccess) return ::dugma::hip_fail(e_, #call); \
  } while (0)

// Timing bracket around a pair-kernel launch (no-op unless timing is on);
// finish() waits for the stop event, then adds the elapsed time.
struct PairTimer {
  dugma_ctx *c;
  explicit PairTimer(dugma_ctx *ctx);
  void stop();
  void finish();
};

// Waiting for results that a call's last kernel stores to the context's
// pinned result slots (coherent host memory, written without the GPU's L2):
// the host presets the slots to a quiet NaN before the launch and spins until
// none is NaN, for up to kSpinUs, then falls back to hipStreamSynchronize
// (which also reports a failed launch, and covers results that are NaN).
// Each slot is one aligned 8-byte (or 4-byte) store, so a slot that is no
// longer NaN holds its final value.  The spin saves the wake-up latency of a
// blocking synchronisation on short calls.
constexpr long kSpinUs = 2000;
void preset_slots(double *slots, int n);
void preset_slots(float *slots, int n);
int wait_slots(dugma_ctx *c, const double *slots, int n);
int wait_slots(dugma_ctx *c, const float *slots, int n);

// Host work over n items on up to kHostThreads threads (the caller's thread
// included); items are claimed in increasing order from a shared counter.
constexpr unsigned kHostThreads = 8;
template <class F>
inline void parallel_for(size_t n, F &&f) {
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const unsigned nt = (unsigned)std::min<size_t>(std::min(hw, kHostThreads), n);
  std::atomic<size_t> next{0};
  auto work = [&]() {
    for (size_t q; (q = next.fetch_add(1)) < n;) f(q);
  };
  std::vector<std::thread> pool;
  for (unsigned k = 1; k < nt; ++k) {
    try {
      pool.emplace_back(work);
    } catch (...) {
      break;  // no more threads: the ones running take the rest
    }
  }
  work();
  for (auto &th : pool) th.join();
}

// Grow-only device scratch buffer owned by a context.
struct DevBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
  int ensure(size_t need);
  void release();
};

// Grow-only pinned host buffer owned by a context (device-to-host copies of
// bulk results at DMA speed).
struct PinBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
  int ensure(size_t need);
  void release();
};

}  // namespace dugma

struct dugma_ctx;
namespace dugma {
// Bring the host copy of the source data (hSsrc) up to date after a sigma
// update rewrote it on the device.
int refresh_host_sources(dugma_ctx *c);
// Device address of dugma_energy_grad_noalloc's scratch array on `device`
// (the current device), looked up once per device.
int noalloc_setup(int device);
}  // namespace dugma

// ---------------------------------------------------------------------------
// Context (opaque to callers).  The first six device arrays are the state the
// reference keeps behind its six MEX handles (init.cu:326-331).
// ---------------------------------------------------------------------------
struct dugma_ctx {
  int N = 0, M = 0, D = 0;
  int device = 0;
  hipStream_t stream = nullptr;

  // reference state (FP32, as after init.cu)
  float *dX = nullptr;      // N*D
  float *dXCinv = nullptr;  // N*D*D  inverse covariances, column-major
  float *dY = nullptr;      // M*D
  float *dYCinv = nullptr;  // M*D*D
  float *dgX = nullptr;     // N      sqrtf(1/det)
  float *dgY = nullptr;     // M

  // convolved-energy state: valid points only
  int nT = 0, nS = 0;        // kept targets / sources
  bool iso = false;          // every kept covariance is s^2 I (isoT && isoS)
  bool isoT = false, isoS = false;  // the same, per cloud
  int *dTidx = nullptr;      // nT: original target index of each kept point
  int *dSidx = nullptr;      // nS
  void *dTrecF = nullptr;    // nT full-covariance target records
  void *dTrecI = nullptr;    // nT isotropic target records (iso only)
  double *dSsrc = nullptr;   // nS * (D + D*D) FP64 source points + covariances
  void *dSrecF = nullptr;    // nS per-call transformed source records
  void *dSrecI = nullptr;

  // Positions are held times scale = 2^lscale and covariances times scale^2,
  // so that covariances are O(1) in FP32 whatever the caller's units.
  int lscale = 0;
  double scale = 1.0;
  // pinned host slots kernels write their final results to: [0, 64) the
  // pose-pass sums (8 per reducing workgroup), [kResMoments, kResMoments +
  // 400) the compi moments, [kResSig] the update_sig value (as a float)
  double *hres = nullptr;
  double *dres = nullptr;        // device address of hres
  static constexpr int kResMoments = 64, kResSig = 512;

  std::vector<int> hTidx, hSidx;  // host copies of the kept-point lists
  std::vector<double> hSsrc;      // host copy of dSsrc
  bool hSsrc_stale = false;       // dSsrc changed on the device (sigma update)

  // launch shapes chosen by dugma_autotune: [energy / pose][full / iso]
  struct Tuned {
    int T = 2;
    int target_waves = 0;  // 0: not tuned
  };
  Tuned tuned[2][2];

  // optional kernel timing (HIP events on `stream`)
  bool timing = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t evpts = nullptr;  // per-point gradients: the target sums have reached the host
  double pair_ms = 0.0;
  long long pair_launches = 0;

  dugma::DevBuf part;        // pair-kernel partials
  dugma::DevBuf scratch2;    // reductions / k-NN merge
  dugma::DevBuf sigbuf;      // sigma update: gathered points, sigma, counts
  dugma::PinBuf hstage;      // per-point sums on their way to the host
  double *hpinned = nullptr; // pinned host staging, kHostStage doubles
  static constexpr int kHostStage = 4096;
  // pinned host buffer the pair kernel writes its per-workgroup partials to
  // when there are at most kHostParts of them (summed on the host in a fixed
  // order: no reduction launch).  Host reads of device-written pinned memory
  // cost more than the reduction launch beyond a few hundred partials.
  double *hparts = nullptr;
  double *dparts = nullptr;  // device address of hparts
  static constexpr int kHostParts = 256;
};

