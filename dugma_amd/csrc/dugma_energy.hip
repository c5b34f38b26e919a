// dugma_energy.hip -- convolved-uncertainty energy, fused pose adjoint,
// per-point adjoint and the finite-difference ("plain") gradient.
//
// Kernel shape: a workgroup is 4 waves; each lane owns T stationary points
// in VGPRs (the transformed sources in the pose and energy passes; targets,
// then sources, in the two per-point passes) and walks a chunk of the other
// cloud.  The streamed records are read with wave-uniform
// addresses, so they come through the scalar cache into SGPRs and feed the
// VALU directly: no LDS, no barriers in the pair loop.  (On gfx950 an
// instruction with an SGPR operand issues at about half the rate of an
// all-VGPR one, but staging the records in LDS and reading them into VGPRs
// was slower: the wave-wide LDS reads become the bound.)  Per-pair terms are
// FP32, summed in FP32 over sub-tiles of kSub streamed points and then in
// FP64 per lane; workgroup partials are reduced in a fixed order, so results
// are deterministic run to run.  Nothing of size N x M is stored.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "dugma_internal.h"
#include "dugma_pairs.h"

namespace dugma {

constexpr int kBlock = 256;  // 4 waves of 64
constexpr int kSub = 64;     // FP32 sub-tile length before the FP64 flush
constexpr int kMinChunk = 32;  // fewest streamed points per workgroup

// kModeHess: the pose pass plus the sums of an approximate pose Hessian
// (3-D, transformed sources stationary)
enum { kModeEnergy = 0, kModePose = 1, kModePoint = 2, kModeHess = 3 };
constexpr int kHessPart = 32;  // per workgroup: the kPosePart pose values, then 21 Hessian entries

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Entry (i <= j) of the Hessian of E for moving a stationary point p by
// p - t -> p - t + theta x (p - t) + delta, from its sums a = [E, EW (3), B
// (6, upper triangle)] with d = p - q: with z = p - t and J = [-[z]x | I],
// H = J^T B J plus -1/2 (EW z^T + z EW^T) + (EW . z) I in the theta block
// (the second-order turn of z).  The turn of p's own covariance enters only
// the gradient.  FP32: the Hessian only shapes the search direction.
__device__ __forceinline__ float point_hessian_entry(int i, int j, const float *z, const float *a) {
  auto Bm = [&](int u, int v) -> float {
    const int lo = u < v ? u : v, hi = u < v ? v : u;
    return a[4 + (lo == 0 ? hi : lo == 1 ? 2 + hi : 5)];
  };
  auto Mm = [&](int u, int v) -> float {  // [z]x
    if (u == v) return 0.0f;
    const int w = 3 - u - v;
    return ((u + 1) % 3 == v) ? -z[w] : z[w];
  };
  if (i >= 3) return Bm(i - 3, j - 3);
  if (j >= 3) {
    const int c = j - 3;
    return Mm(i, 0) * Bm(0, c) + Mm(i, 1) * Bm(1, c) + Mm(i, 2) * Bm(2, c);
  }
  float v = 0.0f;
  for (int u = 0; u < 3; ++u)
    for (int w = 0; w < 3; ++w) v += Mm(u, i) * Bm(u, w) * Mm(w, j);
  const float ez = a[1] * z[0] + a[2] * z[1] + a[3] * z[2];
  return v - 0.5f * (a[1 + i] * z[j] + z[i] * a[1 + j]) + (i == j ? ez : 0.0f);
}

struct PoseArg {
  double R[9];  // column-major
  double t[3];
};

// ---------------------------------------------------------------------------
// Pair kernel
// ---------------------------------------------------------------------------
// With one stationary point per lane every variant fits in 64 VGPRs (no
// scratch), which allows 8 waves per SIMD instead of 6; two per lane need
// more registers and keep the compiler's default.
//
// Tiles are (stationary block, streamed chunk) pairs.  With xcd_map set
// they lie on a 1-D grid mapped XCD-aware: workgroups are dispatched to the 8
// XCDs round-robin, so XCD x runs workgroups x, x + 8, ...; each XCD takes a
// contiguous eighth of the stationary blocks (bpx of them) and walks every
// streamed chunk across them, chunk by chunk.  A chunk is then fetched into
// an XCD's L2 once and used by its bpx blocks, and the XCD's stationary
// records (an eighth of them) stay in that L2 for every chunk.  The grid is
// padded to 8 x bpx x qchunks workgroups; the padding exits at once.
// Without xcd_map (clouds that fit one L2 anyway) the grid is 2-D:
// block blockIdx.x of chunk blockIdx.y.
constexpr int kXcds = 8;
constexpr size_t kXcdMapBytes = size_t(4) << 20;  // both clouds' records beyond one XCD's L2 (4 MiB)
__device__ __forceinline__ bool xcd_tile(int pblocks, int qchunks, int xcd_map, int *sb, int *qc) {
  if (!xcd_map) {
    *sb = blockIdx.x;
    *qc = blockIdx.y;
    return true;
  }
  const int g = blockIdx.x;
  const int bpx = (pblocks + kXcds - 1) / kXcds;
  const int xcd = g % kXcds, local = g / kXcds;
  *sb = xcd * bpx + local % bpx;
  *qc = local / bpx;
  return *sb < pblocks && *qc < qchunks;
}
static inline dim3 tile_grid(int pblocks, int qchunks, int xcd_map) {
  return xcd_map ? dim3(kXcds * ((pblocks + kXcds - 1) / kXcds) * qchunks) : dim3(pblocks, qchunks);
}

template <class Rec, int MODE, int T, int COMP>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(T == 1 && MODE != kModeHess ? 8 : 1, 8))) void k_pairs(
    const Rec *__restrict__ P, int nP, const Rec *__restrict__ Q, int nQ,
    int qchunk, int pblocks, int qchunks, int xcd_map, double *__restrict__ out, double t0, double t1,
    double t2) {
  using Tr = RecTraits<Rec>;
  constexpr bool GRAD = MODE != kModeEnergy;
  // the Hessian pass needs B: isotropic records then hold it in two parts
  constexpr int NA = MODE == kModeHess ? (Tr::FULL ? Tr::NACC : 11) : (GRAD ? Tr::NACC : 1);
  constexpr int PART = MODE == kModeHess ? kHessPart : kPosePart;
  int sb, qc;
  if (!xcd_tile(pblocks, qchunks, xcd_map, &sb, &qc)) return;  // grid padding (whole workgroup)
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int pbase = sb * (kBlock * T) + wave * (64 * T) + lane;
  const int q0 = qc * qchunk;
  const int q1 = min(q0 + qchunk, nQ);

  Rec p[T];
#pragma unroll
  for (int k = 0; k < T; ++k) p[k] = P[min(pbase + 64 * k, nP - 1)];

  double dacc[T][NA];
#pragma unroll
  for (int k = 0; k < T; ++k)
#pragma unroll
    for (int a = 0; a < NA; ++a) dacc[k][a] = 0.0;

  // the streamed records are software-pipelined two deep: record j + 2 is
  // requested while record j is used (q0 < q1 for every workgroup; indices
  // clamped to the chunk)
  Rec qa = Q[q0], qb = Q[min(q0 + 1, q1 - 1)];
  for (int j0 = q0; j0 < q1; j0 += kSub) {
    float acc[T][NA];
#pragma unroll
    for (int k = 0; k < T; ++k)
#pragma unroll
      for (int a = 0; a < NA; ++a) acc[k][a] = 0.0f;
    const int j1 = min(j0 + kSub, q1);
#pragma unroll 2
    for (int j = j0; j < j1; ++j) {
      const Rec q = qa;
      qa = qb;
      qb = Q[min(j + 2, q1 - 1)];
#pragma unroll
      for (int k = 0; k < T; ++k) pair_eval<MODE, COMP>(p[k], q, acc[k]);
    }
#pragma unroll
    for (int k = 0; k < T; ++k)
#pragma unroll
      for (int a = 0; a < NA; ++a) dacc[k][a] += (double)acc[k][a];
  }

  // Isotropic records: variances were times ln 4 (ESCALE, GSCALE).  A pose
  // pass with COMP 2 has the transformed sources stationary: d = x' - y
  // flips the sign of EW (and of nothing else: B holds d d^T), and the
  // gradient is then taken about the sources, whose rotation also turns
  // their covariances.  With EW and B both negated, pose_contrib's
  // (p - t) x EW + vee(S_p B - B S_p) is exactly that gradient.
  constexpr bool kSwapped = (MODE == kModePose || MODE == kModeHess) && COMP == 2;
  if (!RecTraits<Rec>::FULL || kSwapped) {
    constexpr double gs = kSwapped ? -RecTraits<Rec>::GSCALE : RecTraits<Rec>::GSCALE;
#pragma unroll
    for (int k = 0; k < T; ++k) {
      dacc[k][0] *= RecTraits<Rec>::ESCALE;
#pragma unroll
      for (int a = 1; a < NA; ++a) dacc[k][a] *= gs;
    }
  }

  if (MODE == kModePoint) {
#pragma unroll
    for (int k = 0; k < T; ++k) {
      const int pi = pbase + 64 * k;
      if (pi < nP) {
        double *o = out + ((size_t)qc * nP + pi) * NA;
#pragma unroll
        for (int a = 0; a < NA; ++a) o[a] = dacc[k][a];
      }
    }
    return;
  }

  double r[kPosePart];
#pragma unroll
  for (int a = 0; a < kPosePart; ++a) r[a] = 0.0;
  const double tt[3] = {t0, t1, t2};
#pragma unroll
  for (int k = 0; k < T; ++k) {
    if (pbase + 64 * k < nP) {
      if (GRAD)
        pose_contrib(p[k], dacc[k], tt, r);
      else
        r[0] += dacc[k][0];
    }
  }
  double *o = out + ((size_t)qc * pblocks + sb) * PART;
  if (GRAD) {
    // Workgroup sums of the kPosePart values in a fixed order: every lane's
    // values go to LDS; thread (a, k) = (tid / 32, tid % 32) adds entries
    // k, k + 32, ..., k + 224 of value a (consecutive lanes read consecutive
    // doubles: no bank conflict); each half wave then sums its 32 partials
    // (DPP row rotations, then the two row sums), value 2 * wave in the low
    // half and 2 * wave + 1 in the high half.
    static_assert(kPosePart * 32 == kBlock, "one half wave per pose value");
    __shared__ double stage[kPosePart][kBlock];
#pragma unroll
    for (int a = 0; a < kPosePart; ++a) stage[a][threadIdx.x] = r[a];
    __syncthreads();
    const int a = threadIdx.x >> 5, k = threadIdx.x & 31;
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < kBlock / 32; ++i) s += stage[a][k + 32 * i];
    double lo, hi;
    half_wave_sums(s, lo, hi);
    if (lane == 0) {
      o[2 * wave] = lo;
      o[2 * wave + 1] = hi;
    }
    if constexpr (MODE == kModeHess && Tr::D == 3) {
      // the Hessian entries, one at a time: lane sums over its T points, a
      // wave sum, then the waves' sums in order.  The block above negated
      // EW and B (sources stationary); the entries need them as summed.
      constexpr float sg = kSwapped ? -1.0f : 1.0f;
      __shared__ float hred[kBlock / 64][21];
      int q = 0;
#pragma unroll
      for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = i; j < 6; ++j) {
          float v = 0.0f;
#pragma unroll
          for (int kk = 0; kk < T; ++kk) {
            if (pbase + 64 * kk >= nP) continue;
            float av[10];
            av[0] = 0.0f;
#pragma unroll
            for (int u = 1; u < 4; ++u) av[u] = sg * (float)dacc[kk][u];
            if (Tr::FULL) {
#pragma unroll
              for (int u = 4; u < 10; ++u) av[u] = sg * (float)dacc[kk][u];
            } else {  // B = P - Q I
              const float qd = (float)dacc[kk][NA - 1];
#pragma unroll
              for (int u = 4; u < 10; ++u)
                av[u] = sg * ((float)dacc[kk][u] - ((u == 4 || u == 7 || u == 9) ? qd : 0.0f));
            }
            const float z[3] = {(float)((double)p[kk].hx + (double)p[kk].lx - t0),
                                (float)((double)p[kk].hy + (double)p[kk].ly - t1),
                                (float)((double)p[kk].hz + (double)p[kk].lz - t2)};
            v += point_hessian_entry(i, j, z, av);
          }
          v = wave_sum_f(v);
          if (lane == 0) hred[wave][q] = v;
          ++q;
        }
      __syncthreads();
      if (threadIdx.x < 21) {
        double s2 = 0.0;
#pragma unroll
        for (int w = 0; w < kBlock / 64; ++w) s2 += (double)hred[w][threadIdx.x];
        o[kPosePart + threadIdx.x] = s2;
      }
    }
  } else {
    r[0] = wave_sum(r[0]);
    __shared__ double red[kBlock / 64];
    if (lane == 0) red[wave] = r[0];
    __syncthreads();
    if (threadIdx.x < kPosePart) {
      double s = 0.0;
      if (threadIdx.x == 0) {
        s = red[0];
#pragma unroll
        for (int w = 1; w < kBlock / 64; ++w) s += red[w];
      }
      o[threadIdx.x] = s;
    }
  }
}

// Fixed-order sum of the per-workgroup partials over kRedGroups workgroups
// at most: workgroup g sums the contiguous range [g len, (g + 1) len) --
// thread t adds partials t, t + 256, ... of it, then the 256 lanes' values
// are summed through LDS and half-wave sums as in the pair kernel -- and
// writes its kPosePart sums to res[kPosePart g, ...), the context's pinned
// host slots; the host adds the workgroups' sums in order.
constexpr int kRedGroups = 8;
__global__ __launch_bounds__(kBlock) void k_reduce_parts(
    const double *__restrict__ parts, int nparts, double *__restrict__ res) {
  const int len = (nparts + (int)gridDim.x - 1) / (int)gridDim.x;
  const int q0 = blockIdx.x * len, q1 = min(q0 + len, nparts);
  double s[kPosePart];
#pragma unroll
  for (int a = 0; a < kPosePart; ++a) s[a] = 0.0;
  for (int q = q0 + (int)threadIdx.x; q < q1; q += kBlock) {
    const double4 *p = reinterpret_cast<const double4 *>(parts + (size_t)q * kPosePart);
    const double4 u = p[0], v = p[1];
    s[0] += u.x; s[1] += u.y; s[2] += u.z; s[3] += u.w;
    s[4] += v.x; s[5] += v.y; s[6] += v.z; s[7] += v.w;
  }
  __shared__ double stage[kPosePart][kBlock];
#pragma unroll
  for (int a = 0; a < kPosePart; ++a) stage[a][threadIdx.x] = s[a];
  __syncthreads();
  const int a = threadIdx.x >> 5, k = threadIdx.x & 31;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double t = 0.0;
#pragma unroll
  for (int i = 0; i < kBlock / 32; ++i) t += stage[a][k + 32 * i];
  double lo, hi;
  half_wave_sums(t, lo, hi);
  if (lane == 0) {
    res[kPosePart * blockIdx.x + 2 * wave] = lo;
    res[kPosePart * blockIdx.x + 2 * wave + 1] = hi;
  }
}

// The Hessian pass's partials (kHessPart doubles per workgroup of the pair
// kernel, 29 used): workgroup g sums its contiguous range, each thread over
// whole records, then the 256 threads' values value by value through LDS in
// a fixed tree order; the host adds the workgroups' sums in order.
constexpr int kHessGroups = 64;
constexpr int kHessUsed = kPosePart + 21;
__global__ __launch_bounds__(kBlock) void k_reduce_hess(const double *__restrict__ parts, int nparts,
                                                        double *__restrict__ res) {
  const int len = (nparts + (int)gridDim.x - 1) / (int)gridDim.x;
  const int q0 = blockIdx.x * len, q1 = min(q0 + len, nparts);
  double s[kHessUsed];
#pragma unroll
  for (int a = 0; a < kHessUsed; ++a) s[a] = 0.0;
  for (int q = q0 + (int)threadIdx.x; q < q1; q += kBlock) {
    const double *row = parts + (size_t)q * kHessPart;
#pragma unroll
    for (int a = 0; a < kHessUsed; ++a) s[a] += row[a];
  }
  __shared__ double stage[kBlock];
#pragma unroll 1
  for (int a = 0; a < kHessUsed; ++a) {
    stage[threadIdx.x] = s[a];
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) stage[threadIdx.x] += stage[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) res[kHessPart * blockIdx.x + a] = stage[0];
    __syncthreads();
  }
}

// Per-point sums over streamed chunks, in chunk order.
__global__ __launch_bounds__(kBlock) void k_reduce_points(
    const double *__restrict__ parts, int nP, int nchunks, int nacc,
    double *__restrict__ res) {
  const size_t idx = (size_t)blockIdx.x * kBlock + threadIdx.x;
  const size_t total = (size_t)nP * nacc;
  if (idx >= total) return;
  double s = 0.0;
  for (int c = 0; c < nchunks; ++c) s += parts[(size_t)c * total + idx];
  res[idx] = s;
}

// ---------------------------------------------------------------------------
// Per-call source records: x' = R x + t and S' = R S R^T in FP64, then
// split to FP32 hi/lo (positions) and rounded (covariance).
// ---------------------------------------------------------------------------
template <int D, bool ISO>
__global__ __launch_bounds__(kBlock) void k_transform_src(
    const double *__restrict__ src, int n, PoseArg ps, void *__restrict__ outv) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= n) return;
  const double *s = src + (size_t)k * (D + D * D);
  double xp[3];
#pragma unroll
  for (int a = 0; a < D; ++a) {
    double v = ps.t[a];
#pragma unroll
    for (int b = 0; b < D; ++b) v = fma(ps.R[a + D * b], s[b], v);
    xp[a] = v;
  }
  float hi[3], lo[3];
#pragma unroll
  for (int a = 0; a < D; ++a) {
    hi[a] = (float)xp[a];
    lo[a] = (float)(xp[a] - (double)hi[a]);
  }
  const double *S = s + D;
  if (ISO) {
    if (D == 3) {
      Rec3I r;
      r.hx = hi[0]; r.hy = hi[1]; r.hz = hi[2];
      r.lx = lo[0]; r.ly = lo[1]; r.lz = lo[2];
      r.s = (float)(S[0] * kLn4);
      r.pad0 = 0.0f;
      reinterpret_cast<Rec3I *>(outv)[k] = r;
    } else {
      Rec2I r;
      r.hx = hi[0]; r.hy = hi[1]; r.lx = lo[0]; r.ly = lo[1];
      r.s = (float)(S[0] * kLn4);
      r.pad0 = r.pad1 = r.pad2 = 0.0f;
      reinterpret_cast<Rec2I *>(outv)[k] = r;
    }
    return;
  }
  // S' = R S R^T
  double RS[9], Sp[9];
#pragma unroll
  for (int a = 0; a < D; ++a)
#pragma unroll
    for (int c = 0; c < D; ++c) {
      double v = 0.0;
#pragma unroll
      for (int b = 0; b < D; ++b) v = fma(ps.R[a + D * b], S[b + D * c], v);
      RS[a + D * c] = v;
    }
#pragma unroll
  for (int a = 0; a < D; ++a)
#pragma unroll
    for (int c = 0; c < D; ++c) {
      double v = 0.0;
#pragma unroll
      for (int b = 0; b < D; ++b) v = fma(RS[a + D * b], ps.R[c + D * b], v);
      Sp[a + D * c] = v;
    }
  if (D == 3) {
    Rec3F r;
    r.hx = hi[0]; r.hy = hi[1]; r.hz = hi[2];
    r.lx = lo[0]; r.ly = lo[1]; r.lz = lo[2];
    r.s11 = (float)Sp[0];
    r.s22 = (float)Sp[4];
    r.s33 = (float)Sp[8];
    r.s12 = (float)(0.5 * (Sp[3] + Sp[1]));
    r.s13 = (float)(0.5 * (Sp[6] + Sp[2]));
    r.s23 = (float)(0.5 * (Sp[7] + Sp[5]));
    r.pad0 = r.pad1 = r.pad2 = r.pad3 = 0.0f;
    reinterpret_cast<Rec3F *>(outv)[k] = r;
  } else {
    Rec2F r;
    r.hx = hi[0]; r.hy = hi[1]; r.lx = lo[0]; r.ly = lo[1];
    r.s11 = (float)Sp[0];
    r.s22 = (float)Sp[3];
    r.s12 = (float)(0.5 * (Sp[2] + Sp[1]));
    r.pad0 = 0.0f;
    reinterpret_cast<Rec2F *>(outv)[k] = r;
  }
}

// ---------------------------------------------------------------------------
// Launch planning
// ---------------------------------------------------------------------------
struct Plan {
  int T = 1;
  int pblocks = 1, qchunks = 1, qchunk = 1;
};

static int env_int(const char *name, int dflt) {
  const char *v = getenv(name);
  return (v && *v) ? atoi(v) : dflt;
}

// Launch shape: T stationary points per lane and a target number of waves
// (streamed chunks are sized to reach it).  A context may carry tuned values
// (dugma_autotune); otherwise: T = 2 and 4096 waves, 16384 for very large
// problems, so the grid covers the 256 CUs several times.
static Plan make_plan(int nP, int nQ, int mode, int T, int target_waves, int max_chunks = 64) {
  Plan pl;
  pl.T = T;
  pl.pblocks = (nP + kBlock * T - 1) / (kBlock * T);
  const int waves_p = pl.pblocks * (kBlock / 64);
  int qc = (target_waves + waves_p - 1) / waves_p;
  int maxc = (nQ + kMinChunk - 1) / kMinChunk;
  if (mode == kModePoint) maxc = std::min(maxc, max_chunks);
  // grid limits: 65,535 chunks on the 2-D grid, 2^31 - 1 threads on the
  // padded 1-D (XCD-mapped) one
  maxc = std::min({maxc, 65535, (int)(INT32_MAX / kBlock / (pl.pblocks + kXcds))});
  qc = std::max(1, std::min(qc, maxc));
  pl.qchunk = (nQ + qc - 1) / qc;
  pl.qchunks = (nQ + pl.qchunk - 1) / pl.qchunk;
  return pl;
}

static Plan ctx_plan(const dugma_ctx *c, int nP, int nQ, int mode, bool iso) {
  int T = 2, tw = ((double)nP * nQ >= 1e9) ? 16384 : 4096;
  const dugma_ctx::Tuned &tu = c->tuned[mode == kModeEnergy ? 0 : 1][iso ? 1 : 0];
  if (tu.target_waves > 0) {
    T = tu.T;
    tw = tu.target_waves;
  }
  T = env_int("DUGMA_LANE_POINTS", T) == 1 ? 1 : 2;
  tw = env_int("DUGMA_TARGET_WAVES", tw);
  return make_plan(nP, nQ, mode, T, tw);
}

template <class Rec, int MODE, int COMP>
static void launch_pairs_c(const Plan &pl, const Rec *P, int nP, const Rec *Q, int nQ,
                           double *out, const double *t, hipStream_t s) {
  // the XCD tile order for clouds beyond one L2, except in the per-point
  // passes (measured 2% slower with it there); DUGMA_XCD_ORDER=0 in the
  // environment keeps the 2-D order (the sums are the same, bit for bit)
  const int xcd_map = MODE != kModePoint && (size_t)(nP + nQ) * sizeof(Rec) > kXcdMapBytes &&
                              env_int("DUGMA_XCD_ORDER", 1) != 0
                          ? 1
                          : 0;
  const dim3 grid = tile_grid(pl.pblocks, pl.qchunks, xcd_map), block(kBlock);
  const double t0 = t[0], t1 = t[1], t2 = (RecTraits<Rec>::D == 3) ? t[2] : 0.0;
  if (pl.T == 2)
    hipLaunchKernelGGL((k_pairs<Rec, MODE, 2, COMP>), grid, block, 0, s, P, nP, Q, nQ, pl.qchunk,
                       pl.pblocks, pl.qchunks, xcd_map, out, t0, t1, t2);
  else
    hipLaunchKernelGGL((k_pairs<Rec, MODE, 1, COMP>), grid, block, 0, s, P, nP, Q, nQ, pl.qchunk,
                       pl.pblocks, pl.qchunks, xcd_map, out, t0, t1, t2);
}

// comp: 0 none, 1 streamed side carries the lo part, 2 stationary side does.
template <class Rec>
static void launch_pairs(int mode, int comp, const Plan &pl, const Rec *P,
                         int nP, const Rec *Q, int nQ, double *out,
                         const double *t, hipStream_t s) {
#define DUGMA_L(M, C) launch_pairs_c<Rec, M, C>(pl, P, nP, Q, nQ, out, t, s)
  if (mode == kModeEnergy) {
    // comp 2 in these passes: the transformed sources are stationary (see pose_pass)
    if (comp == 2) DUGMA_L(kModeEnergy, 2); else if (comp == 1) DUGMA_L(kModeEnergy, 1); else DUGMA_L(kModeEnergy, 0);
  } else if (mode == kModePose) {
    if (comp == 2) DUGMA_L(kModePose, 2); else if (comp == 1) DUGMA_L(kModePose, 1); else DUGMA_L(kModePose, 0);
  } else {
    if (comp == 1) DUGMA_L(kModePoint, 1);
    else if (comp == 2) DUGMA_L(kModePoint, 2);
    else DUGMA_L(kModePoint, 0);
  }
#undef DUGMA_L
}

// Device data are in scaled units: positions times s = 2^k, covariances
// times s^2 (chosen at init so covariances are O(1)); exact in binary FP.
static PoseArg make_pose(const dugma_ctx *c, const double *R, const double *t) {
  PoseArg ps;
  memset(&ps, 0, sizeof(ps));
  for (int a = 0; a < c->D * c->D; ++a) ps.R[a] = R[a];
  for (int a = 0; a < c->D; ++a) ps.t[a] = t[a] * c->scale;
  return ps;
}

static int transform_sources(dugma_ctx *c, const PoseArg &ps, bool iso) {
  if (c->nS == 0) return DUGMA_OK;
  dim3 grid((c->nS + kBlock - 1) / kBlock), block(kBlock);
  void *out = iso ? c->dSrecI : c->dSrecF;
  if (c->D == 3) {
    if (iso)
      hipLaunchKernelGGL((k_transform_src<3, true>), grid, block, 0, c->stream, c->dSsrc, c->nS, ps, out);
    else
      hipLaunchKernelGGL((k_transform_src<3, false>), grid, block, 0, c->stream, c->dSsrc, c->nS, ps, out);
  } else {
    if (iso)
      hipLaunchKernelGGL((k_transform_src<2, true>), grid, block, 0, c->stream, c->dSsrc, c->nS, ps, out);
    else
      hipLaunchKernelGGL((k_transform_src<2, false>), grid, block, 0, c->stream, c->dSsrc, c->nS, ps, out);
  }
  DUGMA_HIP(hipGetLastError());
  return DUGMA_OK;
}

static int resolve_iso(dugma_ctx *c, int cov_mode, bool *iso) {
  if (cov_mode == DUGMA_COV_FULL) {
    *iso = false;
  } else if (cov_mode == DUGMA_COV_ISO) {
    if (!c->iso) return fail(DUGMA_ERR_ARG, "DUGMA_COV_ISO: covariances are not all isotropic");
    *iso = true;
  } else if (cov_mode == DUGMA_COV_AUTO) {
    *iso = c->iso;
  } else {
    return fail(DUGMA_ERR_ARG, "unknown cov_mode");
  }
  return DUGMA_OK;
}

static double gauss_norm(int D) {
  const double two_pi = 6.283185307179586476925;
  return D == 3 ? 1.0 / (two_pi * std::sqrt(two_pi)) : 1.0 / two_pi;
}

static bool comp_enabled() { return env_int("DUGMA_COMPENSATED", 1) != 0; }
// the pose and energy passes keep the transformed sources stationary unless
// DUGMA_STATIONARY_SOURCES=0 (then the targets, with the lo correction on
// the streamed side); both give the same sums to FP32 rounding
static bool sources_stationary() { return comp_enabled() && env_int("DUGMA_STATIONARY_SOURCES", 1) != 0; }

// Energy-only or fused pose pass; res receives [E, rotation, translation]
// in caller units.
static int pose_pass(dugma_ctx *c, const double *R, const double *t, bool iso,
                     int mode, double *res, double *fixed_parts = nullptr, size_t fixed_records = 0) {
  const int D = c->D;
  const int width = mode == kModeEnergy ? 1 : (D == 3 ? 7 : 4);
  for (int a = 0; a < width; ++a) res[a] = 0.0;
  if (c->nT == 0 || c->nS == 0) return DUGMA_OK;
  // everything the launches need is settled first, so that the three
  // launches leave the host back to back (the transform takes ~3 us)
  PoseArg ps = make_pose(c, R, t);
  // The passes keep the transformed sources (FP32 hi + lo) stationary and
  // stream the targets (exact in FP32): the lo correction is then an
  // operation on VGPRs only (COMP 2) instead of a second one on the streamed
  // SGPRs (COMP 1).  E is symmetric in the two roles; the gradient sums
  // change sign (kernel epilogue).  Without compensation the targets stay
  // stationary (COMP 0).
  const bool swap = sources_stationary();
  const int comp = swap ? 2 : (comp_enabled() ? 1 : 0);
  Plan pl = swap ? ctx_plan(c, c->nS, c->nT, mode, iso) : ctx_plan(c, c->nT, c->nS, mode, iso);
  const size_t nparts = (size_t)pl.pblocks * pl.qchunks;
  // few partials: the pair kernel writes them to pinned host memory and the
  // host sums them after the sync; many: a device reduction launch
  const bool host_sum = nparts <= (size_t)dugma_ctx::kHostParts;
  double *parts = c->dparts;
  int rc;
  if (!host_sum && fixed_parts) {
    if (nparts > fixed_records)
      return fail(DUGMA_ERR_SIZE, "dugma_energy_grad_noalloc: the launch needs more partial records than the "
                                  "scratch array holds (lower DUGMA_TARGET_WAVES or the autotuned wave count)");
    parts = fixed_parts;
  } else if (!host_sum) {
    rc = c->part.ensure(nparts * kPosePart * sizeof(double));
    if (rc) return rc;
    parts = (double *)c->part.ptr;
  }
  const int groups = (int)std::min<size_t>(kRedGroups, (nparts + kBlock - 1) / kBlock);
  if (!host_sum) preset_slots(c->hres, kPosePart * groups);
  rc = transform_sources(c, ps, iso);
  if (rc) return rc;
  PairTimer timer(c);
  const void *tgt = iso ? c->dTrecI : c->dTrecF, *src = iso ? c->dSrecI : c->dSrecF;
  const void *P = swap ? src : tgt, *Q = swap ? tgt : src;
  const int nP = swap ? c->nS : c->nT, nQ = swap ? c->nT : c->nS;
  if (D == 3) {
    if (iso)
      launch_pairs<Rec3I>(mode, comp, pl, (const Rec3I *)P, nP, (const Rec3I *)Q, nQ, parts, ps.t, c->stream);
    else
      launch_pairs<Rec3F>(mode, comp, pl, (const Rec3F *)P, nP, (const Rec3F *)Q, nQ, parts, ps.t, c->stream);
  } else {
    if (iso)
      launch_pairs<Rec2I>(mode, comp, pl, (const Rec2I *)P, nP, (const Rec2I *)Q, nQ, parts, ps.t, c->stream);
    else
      launch_pairs<Rec2F>(mode, comp, pl, (const Rec2F *)P, nP, (const Rec2F *)Q, nQ, parts, ps.t, c->stream);
  }
  timer.stop();
  DUGMA_HIP(hipGetLastError());
  if (!host_sum) {
    hipLaunchKernelGGL(k_reduce_parts, dim3(groups), dim3(kBlock), 0, c->stream, parts, (int)nparts, c->dres);
  }
  DUGMA_HIP(hipGetLastError());
  if (host_sum) {
    DUGMA_HIP(hipStreamSynchronize(c->stream));
  } else {
    rc = wait_slots(c, c->hres, kPosePart * groups);
    if (rc) return rc;
  }
  timer.finish();
  double sum[kPosePart] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (host_sum) {
    const volatile double *hp = c->hparts;
    for (size_t q = 0; q < nparts; ++q)
      for (int a = 0; a < width; ++a) sum[a] += hp[q * kPosePart + a];
  } else {
    const volatile double *hr = c->hres;
    for (int g = 0; g < groups; ++g)
      for (int a = 0; a < width; ++a) sum[a] += hr[kPosePart * g + a];
  }
  // unscale: E ~ s^D, rotation gradient ~ s^D, translation gradient ~ s^(D+1)
  const double cn = gauss_norm(D);
  const double sD = std::pow(c->scale, D);
  const int nrot = D == 3 ? 3 : 1;
  res[0] = cn * sD * sum[0];
  for (int a = 1; a < width; ++a) res[a] = cn * sD * sum[a] * (a > nrot ? c->scale : 1.0);
  return DUGMA_OK;
}

// The fused pose pass with the sums of an approximate Hessian (3-D, the
// transformed sources stationary): res = [E, rotation (3), translation
// (3)] as pose_pass gives them, H = the upper triangle, row by row, of the
// 6 x 6 Hessian of E in (w, dt) for R <- exp([w]x) R, t <- t + dt, the turn
// of the sources' covariances left out of its second order; caller units.
static int pose_hess_pass(dugma_ctx *c, const double *R, const double *t, bool iso, double *res, double *H) {
  for (int a = 0; a < 7; ++a) res[a] = 0.0;
  for (int a = 0; a < 21; ++a) H[a] = 0.0;
  if (c->D != 3) return fail(DUGMA_ERR_ARG, "the pose Hessian is 3-D only");
  if (c->nT == 0 || c->nS == 0) return DUGMA_OK;
  PoseArg ps = make_pose(c, R, t);
  const Plan pl = ctx_plan(c, c->nS, c->nT, kModePose, iso);
  const size_t nparts = (size_t)pl.pblocks * pl.qchunks;
  const int groups = (int)std::min<size_t>(kHessGroups, (nparts + kBlock - 1) / kBlock);
  int rc = c->part.ensure(nparts * kHessPart * sizeof(double));
  if (!rc) rc = c->scratch2.ensure(sizeof(double) * kHessPart * groups);
  if (!rc) rc = c->hstage.ensure(sizeof(double) * kHessPart * groups);
  if (!rc) rc = transform_sources(c, ps, iso);
  if (rc) return rc;
  double *parts = (double *)c->part.ptr, *gres = (double *)c->scratch2.ptr;
  PairTimer timer(c);
  if (iso)
    launch_pairs_c<Rec3I, kModeHess, 2>(pl, (const Rec3I *)c->dSrecI, c->nS, (const Rec3I *)c->dTrecI, c->nT,
                                          parts, ps.t, c->stream);
  else
    launch_pairs_c<Rec3F, kModeHess, 2>(pl, (const Rec3F *)c->dSrecF, c->nS, (const Rec3F *)c->dTrecF, c->nT,
                                          parts, ps.t, c->stream);
  timer.stop();
  hipLaunchKernelGGL(k_reduce_hess, dim3(groups), dim3(kBlock), 0, c->stream, parts, (int)nparts, gres);
  DUGMA_HIP(hipGetLastError());
  DUGMA_HIP(hipMemcpyAsync(c->hstage.ptr, gres, sizeof(double) * kHessPart * groups, hipMemcpyDeviceToHost,
                           c->stream));
  DUGMA_HIP(hipStreamSynchronize(c->stream));
  timer.finish();
  double sum[kHessUsed];
  for (int a = 0; a < kHessUsed; ++a) sum[a] = 0.0;
  const double *hs = (const double *)c->hstage.ptr;
  for (int g = 0; g < groups; ++g)
    for (int a = 0; a < kHessUsed; ++a) sum[a] += hs[kHessPart * g + a];
  // unscale: E ~ s^3, and a factor s per translation index
  const double cn = gauss_norm(3), sc = c->scale, sD = sc * sc * sc;
  res[0] = cn * sD * sum[0];
  for (int a = 1; a < 7; ++a) res[a] = cn * sD * sum[a] * (a > 3 ? sc : 1.0);
  int q = 0;
  for (int i = 0; i < 6; ++i)
    for (int j = i; j < 6; ++j, ++q) H[q] = cn * sD * sum[kPosePart + q] * (i >= 3 ? sc : 1.0) * (j >= 3 ? sc : 1.0);
  return DUGMA_OK;
}

static int check_pose_args(dugma_ctx *c, const double *R, const double *t) {
  if (!c) return fail(DUGMA_ERR_ARG, "null context");
  if (!R || !t) return fail(DUGMA_ERR_ARG, "null R or t");
  for (int a = 0; a < c->D * c->D; ++a)
    if (!std::isfinite(R[a])) return fail(DUGMA_ERR_ARG, "R is not finite");
  for (int a = 0; a < c->D; ++a)
    if (!std::isfinite(t[a])) return fail(DUGMA_ERR_ARG, "t is not finite");
  return DUGMA_OK;
}

// Left perturbation of the pose: 3-D p = [w, dt], R <- exp([w]x) R;
// 2-D p = [theta, dt], R <- Rot(theta) R.
static void perturb(int D, const double *R, const double *t, const double *p,
                    double *Rn, double *tn) {
  if (D == 3) {
    const double *w = p;
    double th = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    double A = th < 1e-12 ? 1.0 : std::sin(th) / th;
    double B = th < 1e-12 ? 0.5 : (1.0 - std::cos(th)) / (th * th);
    double K[9] = {0, w[2], -w[1], -w[2], 0, w[0], w[1], -w[0], 0};
    double E[9];
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < 3; ++c) {
        double k2 = 0;
        for (int b = 0; b < 3; ++b) k2 += K[a + 3 * b] * K[b + 3 * c];
        E[a + 3 * c] = (a == c ? 1.0 : 0.0) + A * K[a + 3 * c] + B * k2;
      }
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < 3; ++c) {
        double s = 0;
        for (int b = 0; b < 3; ++b) s += E[a + 3 * b] * R[b + 3 * c];
        Rn[a + 3 * c] = s;
      }
    for (int a = 0; a < 3; ++a) tn[a] = t[a] + p[3 + a];
  } else {
    double cs = std::cos(p[0]), sn = std::sin(p[0]);
    double E[4] = {cs, sn, -sn, cs};
    for (int a = 0; a < 2; ++a)
      for (int c = 0; c < 2; ++c)
        Rn[a + 2 * c] = E[a] * R[2 * c] + E[a + 2] * R[1 + 2 * c];
    tn[0] = t[0] + p[1];
    tn[1] = t[1] + p[2];
  }
}

constexpr size_t kHostChunk = 4096;  // points per host task of the per-point conversion

// One per-point pass (stationary set P, streamed set Q), full covariance
// records; writes per-point FP64 sums [nP][NACC] to out (device).
// Launch: 65,536 waves, at most 64 streamed chunks and 1 GiB of per-chunk
// partials.  At 76,800 x 76,800 the two passes take 14.9 ms at 55 chunks
// against 17.6 ms at 7 chunks (8,192 waves).
constexpr size_t kPointPartBytes = size_t(1) << 30;
static int point_pass(dugma_ctx *c, const void *P, int nP, const void *Q,
                      int nQ, int comp, double *out) {
  const int D = c->D;
  const int nacc = D == 3 ? kAcc3F : kAcc2F;
  const int max_chunks =
      (int)std::max<size_t>(1, std::min<size_t>(64, kPointPartBytes / ((size_t)nP * nacc * sizeof(double))));
  Plan pl = make_plan(nP, nQ, kModePoint, 1, env_int("DUGMA_POINT_WAVES", 65536), max_chunks);
  const size_t nparts = (size_t)pl.qchunks * nP * nacc;
  int rc = c->part.ensure(nparts * sizeof(double));
  if (rc) return rc;
  double *parts = (double *)c->part.ptr;
  const double t0[3] = {0, 0, 0};
  if (!comp_enabled()) comp = 0;
  PairTimer timer(c);
  if (D == 3)
    launch_pairs<Rec3F>(kModePoint, comp, pl, (const Rec3F *)P, nP, (const Rec3F *)Q, nQ, parts, t0, c->stream);
  else
    launch_pairs<Rec2F>(kModePoint, comp, pl, (const Rec2F *)P, nP, (const Rec2F *)Q, nQ, parts, t0, c->stream);
  timer.stop();
  DUGMA_HIP(hipGetLastError());
  const size_t total = (size_t)nP * nacc;
  hipLaunchKernelGGL(k_reduce_points, dim3((unsigned)((total + kBlock - 1) / kBlock)), dim3(kBlock), 0, c->stream,
                     parts, nP, pl.qchunks, nacc, out);
  DUGMA_HIP(hipGetLastError());
  timer.finish();
  return DUGMA_OK;
}

// ---------------------------------------------------------------------------
// Scratch of dugma_energy_grad_noalloc: the pair kernel's per-workgroup
// partials (kPosePart doubles each) in a file-scope device array instead of
// the context's grow-only buffer.  Each GPU holds its own copy until the
// process ends; its address on a device comes from hipGetSymbolAddress with
// that device current (dugma_init calls noalloc_setup), and host code only
// ever passes that address.  Every partial record of a launch is written by
// its workgroup before k_reduce_parts reads it, so nothing is cleared.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocRecords = size_t(1) << 18;  // 16 MiB: 2^18 workgroups' partials
static __device__ __align__(256) double scratch[kNoAllocRecords * kPosePart];
constexpr int kNoAllocDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[kNoAllocDevices];

int noalloc_setup(int device) {
  if (device < 0 || device >= kNoAllocDevices) return DUGMA_OK;  // the entry point reports it
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return DUGMA_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
  g_scratch_ptr[device] = static_cast<double *>(p);
  return DUGMA_OK;
}

static double *noalloc_scratch(int device) {
  if (device < 0 || device >= kNoAllocDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[device];
}

}  // namespace dugma

using namespace dugma;

extern "C" int dugma_energy_grad_noalloc(dugma_ctx *ctx, const double *R, const double *t, int cov_mode,
                                         double *E, double *gpose) {
  int rc = check_pose_args(ctx, R, t);
  if (rc) return rc;
  double *sp = noalloc_scratch(ctx->device);
  if (!sp) return fail(DUGMA_ERR_STATE, "dugma_energy_grad_noalloc: no scratch address for this device");
  bool iso;
  rc = resolve_iso(ctx, cov_mode, &iso);
  if (rc) return rc;
  double res[kPosePart];
  rc = pose_pass(ctx, R, t, iso, kModePose, res, sp, kNoAllocRecords);
  if (rc) return rc;
  if (E) *E = res[0];
  const int nd = ctx->D == 3 ? 6 : 3;
  if (gpose)
    for (int a = 0; a < nd; ++a) gpose[a] = res[1 + a];
  return DUGMA_OK;
}

extern "C" int dugma_energy(dugma_ctx *ctx, const double *R, const double *t,
                            int cov_mode, double *E) {
  int rc = check_pose_args(ctx, R, t);
  if (rc) return rc;
  if (!E) return fail(DUGMA_ERR_ARG, "null E");
  bool iso;
  rc = resolve_iso(ctx, cov_mode, &iso);
  if (rc) return rc;
  double res[kPosePart];
  rc = pose_pass(ctx, R, t, iso, kModeEnergy, res);
  if (rc) return rc;
  *E = res[0];
  return DUGMA_OK;
}

extern "C" int dugma_energy_grad(dugma_ctx *ctx, const double *R,
                                 const double *t, int cov_mode, double *E,
                                 double *gpose) {
  int rc = check_pose_args(ctx, R, t);
  if (rc) return rc;
  bool iso;
  rc = resolve_iso(ctx, cov_mode, &iso);
  if (rc) return rc;
  double res[kPosePart];
  rc = pose_pass(ctx, R, t, iso, kModePose, res);
  if (rc) return rc;
  if (E) *E = res[0];
  const int nd = ctx->D == 3 ? 6 : 3;
  if (gpose)
    for (int a = 0; a < nd; ++a) gpose[a] = res[1 + a];
  return DUGMA_OK;
}

extern "C" int dugma_energy_grad_hess(dugma_ctx *ctx, const double *R, const double *t, int cov_mode, double *E,
                                      double *gpose, double *H) {
  int rc = check_pose_args(ctx, R, t);
  if (rc) return rc;
  if (ctx->D != 3) return fail(DUGMA_ERR_ARG, "dugma_energy_grad_hess: 3-D contexts only");
  bool iso;
  rc = resolve_iso(ctx, cov_mode, &iso);
  if (rc) return rc;
  double res[7], h[21];
  rc = pose_hess_pass(ctx, R, t, iso, res, h);
  if (rc) return rc;
  if (E) *E = res[0];
  if (gpose)
    for (int a = 0; a < 6; ++a) gpose[a] = res[1 + a];
  if (H) {
    int q = 0;
    for (int i = 0; i < 6; ++i)
      for (int j = i; j < 6; ++j, ++q) H[6 * i + j] = H[6 * j + i] = h[q];
  }
  return DUGMA_OK;
}

extern "C" int dugma_energy_grad_fd(dugma_ctx *ctx, const double *R,
                                    const double *t, int cov_mode, double h,
                                    double *E, double *gpose) {
  int rc = check_pose_args(ctx, R, t);
  if (rc) return rc;
  if (!(h > 0) || !std::isfinite(h)) return fail(DUGMA_ERR_ARG, "h must be positive");
  bool iso;
  rc = resolve_iso(ctx, cov_mode, &iso);
  if (rc) return rc;
  const int D = ctx->D, nd = D == 3 ? 6 : 3;
  double res[kPosePart];
  rc = pose_pass(ctx, R, t, iso, kModeEnergy, res);
  if (rc) return rc;
  if (E) *E = res[0];
  for (int k = 0; k < nd; ++k) {
    double p[6] = {0, 0, 0, 0, 0, 0}, Rn[9], tn[3], ep, em;
    p[k] = h;
    perturb(D, R, t, p, Rn, tn);
    rc = pose_pass(ctx, Rn, tn, iso, kModeEnergy, res);
    if (rc) return rc;
    ep = res[0];
    p[k] = -h;
    perturb(D, R, t, p, Rn, tn);
    rc = pose_pass(ctx, Rn, tn, iso, kModeEnergy, res);
    if (rc) return rc;
    em = res[0];
    if (gpose) gpose[k] = (ep - em) / (2.0 * h);
  }
  return DUGMA_OK;
}

extern "C" int dugma_energy_grad_points(dugma_ctx *ctx, const double *R,
                                        const double *t, int cov_mode,
                                        double *E, double *gpose, double *gX,
                                        double *gXC, double *gY, double *gYC,
                                        double *gdT) {
  int rc = check_pose_args(ctx, R, t);
  if (rc) return rc;
  if (cov_mode != DUGMA_COV_AUTO && cov_mode != DUGMA_COV_FULL && cov_mode != DUGMA_COV_ISO)
    return fail(DUGMA_ERR_ARG, "unknown cov_mode");
  dugma_ctx *c = ctx;
  const int D = c->D, DD = D * D, N = c->N, M = c->M;
  const int nacc = D == 3 ? kAcc3F : kAcc2F;
  const int nd = D == 3 ? 6 : 3;
  // Points left out of the records (non-finite) keep zero gradients.
  auto clear_points = [&]() {
    if (gX) std::fill(gX, gX + (size_t)N * D, 0.0);
    if (gXC) std::fill(gXC, gXC + (size_t)N * DD, 0.0);
    if (gY) std::fill(gY, gY + (size_t)M * D, 0.0);
    if (gYC) std::fill(gYC, gYC + (size_t)M * DD, 0.0);
  };
  if (gpose) std::fill(gpose, gpose + nd, 0.0);
  if (gdT) std::fill(gdT, gdT + (D == 3 ? 12 : 4), 0.0);
  if (E) *E = 0.0;
  if (c->nT == 0 || c->nS == 0) {
    clear_points();
    return DUGMA_OK;
  }

  // per-point passes use full-covariance records (isotropic points included)
  PoseArg ps = make_pose(c, R, t);
  rc = transform_sources(c, ps, false);
  if (rc) return rc;
  const size_t nT = c->nT, nS = c->nS;
  rc = c->scratch2.ensure(sizeof(double) * (nT + nS) * nacc);
  if (rc) return rc;
  double *acc = (double *)c->scratch2.ptr;
  rc = c->hstage.ensure(sizeof(double) * (nT + nS) * nacc);
  if (rc) return rc;
  const double *hT = (const double *)c->hstage.ptr, *hS = hT + nT * nacc;
  rc = refresh_host_sources(c);  // the host copy of the sources, before the passes (it may sync)
  if (rc) return rc;
  if (!c->evpts) DUGMA_HIP(hipEventCreateWithFlags(&c->evpts, hipEventDisableTiming));
  // target pass, its sums on their way to the host, then the source pass:
  // the host converts the target sums while the source pass runs
  rc = point_pass(c, c->dTrecF, c->nT, c->dSrecF, c->nS, 1, acc);
  hipError_t e = hipSuccess;
  if (!rc) e = hipMemcpyAsync(c->hstage.ptr, acc, sizeof(double) * nT * nacc, hipMemcpyDeviceToHost, c->stream);
  if (!rc && e == hipSuccess) e = hipEventRecord(c->evpts, c->stream);
  if (!rc && e == hipSuccess) rc = point_pass(c, c->dSrecF, c->nS, c->dTrecF, c->nT, 2, acc + nT * nacc);
  if (!rc && e == hipSuccess)
    e = hipMemcpyAsync((double *)c->hstage.ptr + nT * nacc, acc + nT * nacc, sizeof(double) * nS * nacc,
                       hipMemcpyDeviceToHost, c->stream);
  clear_points();  // while the passes run
  if (rc) return rc;
  if (e == hipSuccess) e = hipEventSynchronize(c->evpts);
  if (e != hipSuccess) return hip_fail(e, "dugma_energy_grad_points: download");

  // Everything below is in scaled units until the final unscaling:
  // positions x s, covariances x s^2, energy / s^D.
  const double cn = gauss_norm(D);
  const double sc = c->scale, sD = std::pow(sc, D);
  auto cov_grad = [&](const double *a, double *G) {
    if (D == 3) {
      static const int ij[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
      for (int k = 0; k < 6; ++k) {
        double v = 0.5 * cn * a[4 + k];
        G[ij[k][0] + 3 * ij[k][1]] = v;
        G[ij[k][1] + 3 * ij[k][0]] = v;
      }
    } else {
      static const int ij[3][2] = {{0, 0}, {0, 1}, {1, 1}};
      for (int k = 0; k < 3; ++k) {
        double v = 0.5 * cn * a[3 + k];
        G[ij[k][0] + 2 * ij[k][1]] = v;
        G[ij[k][1] + 2 * ij[k][0]] = v;
      }
    }
  };
  const double gpos = sD * sc, gcov = sD * sc * sc;  // unscaling of point / covariance gradients
  // Per-point results go to the caller's arrays from worker threads, over
  // fixed chunks of points; the energy and pose sums of each chunk are then
  // added in chunk order, so results do not depend on the thread count.
  struct PoseSums {
    double gw[3] = {0, 0, 0}, gt[3] = {0, 0, 0}, gth = 0.0, G[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  };
  auto target_chunk = [&](size_t k0, size_t k1, double &Epart) {
    double e = 0.0;
    for (size_t k = k0; k < k1; ++k) {
      const double *a = hT + k * nacc;
      e += a[0];
      const int i = c->hTidx[k];
      if (gX)
        for (int d = 0; d < D; ++d) gX[(size_t)i * D + d] = -cn * a[1 + d] * gpos;
      if (gXC) {
        double G[9];
        cov_grad(a, G);
        for (int q = 0; q < DD; ++q) gXC[(size_t)i * DD + q] = G[q] * gcov;
      }
    }
    Epart = e;
  };
  auto source_chunk = [&](size_t k0, size_t k1, PoseSums &P) {
    for (size_t k = k0; k < k1; ++k) {
      const double *a = hS + k * nacc;
      const double *s = c->hSsrc.data() + k * (D + DD);
      double g[3], H[9], p[3], Sp[9];
      for (int d = 0; d < D; ++d) g[d] = -cn * a[1 + d];  // dE/dx'_j
      cov_grad(a, H);                                     // dE/dS'_j
      for (int d = 0; d < D; ++d) {
        double v = 0;
        for (int b = 0; b < D; ++b) v += R[d + D * b] * s[b];
        p[d] = v;  // R x_j  (= x'_j - t)
      }
      double RS[9];
      for (int a1 = 0; a1 < D; ++a1)
        for (int c1 = 0; c1 < D; ++c1) {
          double v = 0;
          for (int b = 0; b < D; ++b) v += R[a1 + D * b] * s[D + b + D * c1];
          RS[a1 + D * c1] = v;
        }
      for (int a1 = 0; a1 < D; ++a1)
        for (int c1 = 0; c1 < D; ++c1) {
          double v = 0;
          for (int b = 0; b < D; ++b) v += RS[a1 + D * b] * R[c1 + D * b];
          Sp[a1 + D * c1] = v;
        }
      for (int d = 0; d < D; ++d) P.gt[d] += g[d];
      double HS[9], SH[9];
      for (int a1 = 0; a1 < D; ++a1)
        for (int c1 = 0; c1 < D; ++c1) {
          double v = 0, w = 0;
          for (int b = 0; b < D; ++b) {
            v += H[a1 + D * b] * Sp[b + D * c1];
            w += Sp[a1 + D * b] * H[b + D * c1];
          }
          HS[a1 + D * c1] = v;
          SH[a1 + D * c1] = w;
        }
      // dE/d dR at identity: g p^T + 2 H S'
      for (int a1 = 0; a1 < D; ++a1)
        for (int c1 = 0; c1 < D; ++c1) P.G[a1 + D * c1] += g[a1] * p[c1] + 2.0 * HS[a1 + D * c1];
      if (D == 3) {
        P.gw[0] += p[1] * g[2] - p[2] * g[1];
        P.gw[1] += p[2] * g[0] - p[0] * g[2];
        P.gw[2] += p[0] * g[1] - p[1] * g[0];
        // -2 vee(S'H - HS')
        P.gw[0] -= 2.0 * (SH[2 + 3 * 1] - HS[2 + 3 * 1]);
        P.gw[1] -= 2.0 * (SH[0 + 3 * 2] - HS[0 + 3 * 2]);
        P.gw[2] -= 2.0 * (SH[1 + 3 * 0] - HS[1 + 3 * 0]);
      } else {
        P.gth += g[1] * p[0] - g[0] * p[1];
        // tr(H (J S' - S' J)) = tr((S'H - HS') J) = 2 (S'H - HS')_12
        P.gth += 2.0 * (SH[0 + 2 * 1] - HS[0 + 2 * 1]);
      }
      const int j = c->hSidx[k];
      if (gY)
        for (int b = 0; b < D; ++b) {
          double v = 0;
          for (int a1 = 0; a1 < D; ++a1) v += R[a1 + D * b] * g[a1];
          gY[(size_t)j * D + b] = v * gpos;
        }
      if (gYC) {
        double *o = gYC + (size_t)j * DD;
        for (int b = 0; b < D; ++b)
          for (int c1 = 0; c1 < D; ++c1) {
            double v = 0;
            for (int a1 = 0; a1 < D; ++a1)
              for (int e = 0; e < D; ++e) v += R[a1 + D * b] * H[a1 + D * e] * R[e + D * c1];
            o[b + D * c1] = v * gcov;
          }
      }
    }
  };
  const size_t ncT = (nT + kHostChunk - 1) / kHostChunk, ncS = (nS + kHostChunk - 1) / kHostChunk;
  std::vector<double> Epart(ncT);
  std::vector<PoseSums> Ppart(ncS);
  parallel_for(ncT, [&](size_t q) { target_chunk(q * kHostChunk, std::min(nT, (q + 1) * kHostChunk), Epart[q]); });
  e = hipStreamSynchronize(c->stream);  // the source pass and its sums
  if (e != hipSuccess) return hip_fail(e, "dugma_energy_grad_points: download");
  parallel_for(ncS, [&](size_t q) {
    source_chunk(q * kHostChunk, std::min(nS, (q + 1) * kHostChunk), Ppart[q]);
  });
  double Etot = 0.0;
  for (double e : Epart) Etot += e;
  if (E) *E = cn * Etot * sD;
  double gw[3] = {0, 0, 0}, gt[3] = {0, 0, 0}, gth = 0.0, G[9] = {0};
  for (const PoseSums &P : Ppart) {
    for (int a = 0; a < 3; ++a) {
      gw[a] += P.gw[a];
      gt[a] += P.gt[a];
    }
    gth += P.gth;
    for (int q = 0; q < 9; ++q) G[q] += P.G[q];
  }
  // unscale: rotation / dR parts ~ s^D, translation parts ~ s^(D+1)
  if (gpose) {
    if (D == 3) {
      for (int a = 0; a < 3; ++a) {
        gpose[a] = gw[a] * sD;
        gpose[3 + a] = gt[a] * gpos;
      }
    } else {
      gpose[0] = gth * sD;
      gpose[1] = gt[0] * gpos;
      gpose[2] = gt[1] * gpos;
    }
  }
  if (gdT) {
    if (D == 3) {
      // energy3D.m variable order: t1 t2 t3 r11 r12 r13 r21 ... r33 (dR row-major)
      for (int a = 0; a < 3; ++a) gdT[a] = gt[a] * gpos;
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) gdT[3 + 3 * a + b] = G[a + 3 * b] * sD;
    } else {
      // energy2D.m: t1 t2 r11 r21 with dR = [[r11, -r21], [r21, r11]]
      gdT[0] = gt[0] * gpos;
      gdT[1] = gt[1] * gpos;
      gdT[2] = (G[0] + G[3]) * sD;
      gdT[3] = (G[1] - G[2]) * sD;
    }
  }
  return DUGMA_OK;
}

extern "C" int dugma_autotune(dugma_ctx *ctx, int cov_mode) {
  if (!ctx) return fail(DUGMA_ERR_ARG, "dugma_autotune: null context");
  bool iso;
  int rc = resolve_iso(ctx, cov_mode, &iso);
  if (rc) return rc;
  if (ctx->nT == 0 || ctx->nS == 0) return DUGMA_OK;
  const int D = ctx->D;
  double R[9] = {0}, t[3] = {0, 0, 0};
  for (int a = 0; a < D; ++a) R[a * (D + 1)] = 1.0;
  hipEvent_t e0, e1;
  DUGMA_HIP(hipEventCreate(&e0));
  DUGMA_HIP(hipEventCreate(&e1));
  static const int kT[2] = {1, 2};
  // The largest targets serve large problems: at 200K x 200K the kernel
  // takes 17.75 ms at 32,768 waves (11 source chunks), 16.89 ms at 131,072
  // (42) and 16.70-16.76 ms at 524,288 (168).
  static const int kW[] = {2048, 4096, 8192, 16384, 32768, 65536, 131072, 262144, 524288};
  constexpr int kNW = (int)(sizeof(kW) / sizeof(kW[0]));
  const double pairs = (double)ctx->nT * ctx->nS;
  const int reps = pairs > 4e9 ? 1 : (pairs > 1e8 ? 3 : 10);
  for (int m = 0; m < 2; ++m) {
    const int mode = m == 0 ? kModeEnergy : kModePose;
    dugma_ctx::Tuned best;
    float best_ms = 1e30f;
    // the stationary and streamed sets of pose_pass
    const int nP = sources_stationary() ? ctx->nS : ctx->nT, nQ = sources_stationary() ? ctx->nT : ctx->nS;
    for (int ti = 0; ti < 2; ++ti) {
      int last_chunks = -1;
      for (int wi = 0; wi < kNW; ++wi) {
        // targets that give the launch of the previous one are not timed again
        const int chunks = make_plan(nP, nQ, mode, kT[ti], kW[wi]).qchunks;
        if (chunks == last_chunks) continue;
        last_chunks = chunks;
        dugma_ctx::Tuned &tu = ctx->tuned[m][iso ? 1 : 0];
        tu.T = kT[ti];
        tu.target_waves = kW[wi];
        double res[kPosePart];
        rc = pose_pass(ctx, R, t, iso, mode, res);  // warm
        if (rc) break;
        (void)hipEventRecord(e0, ctx->stream);
        for (int r = 0; r < reps && !rc; ++r) rc = pose_pass(ctx, R, t, iso, mode, res);
        (void)hipEventRecord(e1, ctx->stream);
        (void)hipEventSynchronize(e1);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (!rc && ms < best_ms) {
          best_ms = ms;
          best = tu;
        }
      }
    }
    ctx->tuned[m][iso ? 1 : 0] = best;
    if (rc) break;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

extern "C" int dugma_ctx_plan(const dugma_ctx *ctx, int cov_mode, int *lane_points,
                              int *target_waves) {
  if (!ctx) return fail(DUGMA_ERR_ARG, "dugma_ctx_plan: null context");
  const bool iso = cov_mode == DUGMA_COV_ISO || (cov_mode == DUGMA_COV_AUTO && ctx->iso);
  const dugma_ctx::Tuned &tu = ctx->tuned[1][iso ? 1 : 0];
  if (lane_points) *lane_points = tu.T;
  if (target_waves) *target_waves = tu.target_waves;
  return DUGMA_OK;
}
