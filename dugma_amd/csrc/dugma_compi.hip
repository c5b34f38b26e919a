// dugma_compi.hip -- compi (compi.cpp:8-74, compi.cu:511-628): coefficients
// of the surrogate polynomial
//   F(dR, dt) = sum_ij P_ij [ u^T A'_j u + d^T B_i d ],
//   d = n_i - dR m'_j - dt,  u = dR^T d,
// with P_ij = gX_i gY_j (exp(-d0^T C_j d0 / 2) + exp(-d0^T B_i d0 / 2)),
// d0 = n_i - R y_j - t, C_j = R A_j^T R^T, m'_j = fl32(R y_j + t),
// A'_j = fl32(C_j), A_j = fl32(inv(CY_j) Isig), B_i = fl32(inv(CX_i) Isig)
// (compi.cu:642-791).  Written in energy3D.m:15 / energy2D.m:7 order.
//
// Instead of expanding 637 terms per pair (and 637 FP64 atomics, as
// compi.cu:800-3403 does), every coefficient is a sum over pairs of P_ij
// times (a feature of target i) times (a feature of source j).  So:
//  1. k_compi_w: for every source j, W_j = sum_i P_ij G_i over the 20 (3-D)
//     or 12 (2-D) target features G = [1, n, n n^T, n^T B n, (B+B^T) n,
//     sym(B)]; sources sit in lanes (64 per workgroup), the workgroup's
//     target chunk is split over its 4 waves and streams through SGPRs; the
//     waves' sums meet in LDS, one partial per (chunk, source).  P's
//     exponents are FP64, as in compi.cu:747-772: the pair difference from
//     the FP64 transform, both Mahalanobis forms and log2 of the target's
//     factor gX fl32(Isig^3/2) in FP64; each lane scales its sums by a
//     running power of two (the largest exponent seen, raised in steps of at
//     least 2^16), so that exp2 runs in FP32 on arguments <= 16 and pairs far
//     below FP32's range keep their FP64 weight; the scale is applied to the
//     lane's FP64 sums at the end.
//  2. k_compi_mom: per block of 32 sources, the chunk partials are summed
//     into the features, staged in LDS, and the 325 (85) moments
//     sum_j W'_j,f * m'_j.. * A'_j,.. accumulated; k_compi_msum adds the
//     block partials, one wave per moment.
//  3. host: a fixed sparse map (built once per D from the expansion of each
//     moment) turns the moments into the polynomial coefficients in FP64.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

#include "dugma_internal.h"
#include "dugma_monomials.inc"
#include "dugma_pairs.h"

namespace dugma {

constexpr int kCBlock = 256;
constexpr int kCWaves = kCBlock / 64;  // waves per workgroup of k_compi_w
constexpr int kCSub = 64;
constexpr int kMomSrc = 32;            // sources per workgroup of k_compi_mom
constexpr int kFeat = 32;  // per-source feature stride (doubles) for the moments
constexpr double kNegHalfLog2eD = -0.72134752044448170368;  // -log2(e) / 2
constexpr double kRescale = 16.0;  // a lane's scale rises when an exponent exceeds it by this (log2)

struct CompiPose {
  double R[9];  // column-major
  double t[3];
  float Isig;   // FP32 1/sig (compi.cu:646 round-to-nearest, :110 round-down in 2-D)
  float gsc;    // sqrtf(Isig^3) in 3-D (compi.cu:662), Isig in 2-D (compi.cu:126)
};

// Streamed target record: the exponent's inputs in FP64 (the quadratic
// form of B = fl32(inv(CX) Isig) as [B11 B22 B33 2B12 2B13 2B23] (2-D
// [B11 B22 2B12]) and lg = log2(fl32(gX gsc))), then n, sym(B) and the
// features n n^T, n^T B n, (B+B^T) n in FP32.  3-D 144 B, 2-D 80 B.
struct alignas(16) TgtRec3 {
  double q[6], lg;
  float n[3], b[6];
  float nn[6], s, v[3];
};
struct alignas(16) TgtRec2 {
  double q[3], lg;
  float n[2], b[3];
  float nn[3], s, v[2];
};
static_assert(sizeof(TgtRec3) == 144 && sizeof(TgtRec2) == 80, "target record layout");

// Per-lane source record: C = R A^T R^T (2-D R A R^T) as a quadratic form
// in FP64 ([C11 C22 C33 2C12 2C13 2C23]), m' = R y + t as FP32 hi + lo, and
// gY fl32(gsc).  3-D 80 B, 2-D 48 B.
struct alignas(16) SrcRec3 {
  double q[6];
  float mh[3], ml[3], g, pad;
};
struct alignas(16) SrcRec2 {
  double q[3];
  float mh[2], ml[2], g, pad;
};
static_assert(sizeof(SrcRec3) == 80 && sizeof(SrcRec2) == 48, "source record layout");

// sym index of (a,b), a <= b: 3-D 00 01 02 11 12 22 ; 2-D 00 01 11
__host__ __device__ constexpr int sidx3(int a, int b) {
  return a <= b ? (a == 0 ? b : (a == 1 ? 2 + b : 5)) : sidx3(b, a);
}
__host__ __device__ constexpr int sidx2(int a, int b) { return a <= b ? a + b : a + b; }
template <int D>
__host__ __device__ constexpr int sidx(int a, int b) { return D == 3 ? sidx3(a, b) : sidx2(a, b); }

// Per-call target records.  B = fl32(inv(CX) * Isig) (compi.cu:655-658).
template <int D>
__device__ void compi_tgt(int i, const float *__restrict__ X, const float *__restrict__ XCinv,
                          const float *__restrict__ gX, const CompiPose &ps, void *__restrict__ outv) {
  float B[9];
#pragma unroll
  for (int q = 0; q < D * D; ++q) B[q] = XCinv[(size_t)i * D * D + q] * ps.Isig;
  double n[3];
#pragma unroll
  for (int a = 0; a < D; ++a) n[a] = (double)X[(size_t)i * D + a];
  double s = 0.0, v[3];
#pragma unroll
  for (int a = 0; a < D; ++a) {
    double va = 0.0;
#pragma unroll
    for (int b = 0; b < D; ++b) {
      s += n[a] * (double)B[a + D * b] * n[b];
      va += ((double)B[a + D * b] + (double)B[b + D * a]) * n[b];
    }
    v[a] = va;
  }
  const float g = gX[i] * ps.gsc;
  const double lg = log2((double)g);  // -inf for the fallback g = 0
  if (D == 3) {
    TgtRec3 r;
    for (int a = 0; a < 3; ++a) {
      r.n[a] = (float)n[a];
      r.v[a] = (float)v[a];
    }
    for (int a = 0; a < 3; ++a)
      for (int b = a; b < 3; ++b) {
        const double sab = 0.5 * ((double)B[a + 3 * b] + (double)B[b + 3 * a]);
        r.b[sidx3(a, b)] = (float)sab;
        r.nn[sidx3(a, b)] = (float)(n[a] * n[b]);
      }
    r.q[0] = B[0];
    r.q[1] = B[4];
    r.q[2] = B[8];
    r.q[3] = (double)B[3] + (double)B[1];
    r.q[4] = (double)B[6] + (double)B[2];
    r.q[5] = (double)B[7] + (double)B[5];
    r.lg = lg;
    r.s = (float)s;
    reinterpret_cast<TgtRec3 *>(outv)[i] = r;
  } else {
    TgtRec2 r;
    for (int a = 0; a < 2; ++a) {
      r.n[a] = (float)n[a];
      r.v[a] = (float)v[a];
    }
    for (int a = 0; a < 2; ++a)
      for (int b = a; b < 2; ++b) {
        r.b[sidx2(a, b)] = (float)(0.5 * ((double)B[a + 2 * b] + (double)B[b + 2 * a]));
        r.nn[sidx2(a, b)] = (float)(n[a] * n[b]);
      }
    r.q[0] = B[0];
    r.q[1] = B[3];
    r.q[2] = (double)B[2] + (double)B[1];
    r.lg = lg;
    r.s = (float)s;
    reinterpret_cast<TgtRec2 *>(outv)[i] = r;
  }
}

// Per-call source state: m' = R y + t (FP64; hi = fl32(m') is the reference's
// m', lo keeps the rest for the pair difference), C = R A^T R^T (FP64, 3-D;
// R A R^T in 2-D as compi.cu:180-183 writes it), gY.
// feat (doubles per source, kFeat stride): [.. W set later .., m' at FM, A' at FA]
template <int D>
__device__ void compi_src(int j, const float *__restrict__ Y, const float *__restrict__ YCinv,
                          const float *__restrict__ gY, const CompiPose &ps, void *__restrict__ srcv,
                          double *__restrict__ feat) {
  constexpr int FM = D == 3 ? 20 : 12;
  constexpr int FA = FM + D;
  float A[9];
#pragma unroll
  for (int q = 0; q < D * D; ++q) A[q] = YCinv[(size_t)j * D * D + q] * ps.Isig;
  double C[9];
#pragma unroll
  for (int p = 0; p < D; ++p)
#pragma unroll
    for (int q = 0; q < D; ++q) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < D; ++k) {
        double inner = 0.0;
#pragma unroll
        for (int l = 0; l < D; ++l) {
          // 3-D: C = R A^T R^T  (a_lk r_ql); 2-D: C = R A R^T (a_kl r_ql)
          const double a = (D == 3) ? (double)A[l + D * k] : (double)A[k + D * l];
          inner += a * ps.R[q + D * l];
        }
        acc += ps.R[p + D * k] * inner;
      }
      C[p + D * q] = acc;
    }
  double *f = feat + (size_t)j * kFeat;
  float mh[3], ml[3];
#pragma unroll
  for (int a = 0; a < D; ++a) {
    double v = ps.R[a] * (double)Y[(size_t)j * D];
#pragma unroll
    for (int b = 1; b < D; ++b) v = v + ps.R[a + D * b] * (double)Y[(size_t)j * D + b];
    v = v + ps.t[a];
    mh[a] = (float)v;
    ml[a] = (float)(v - (double)mh[a]);
    f[FM + a] = (double)mh[a];
  }
  const float g = gY[j] * ps.gsc;
  if (D == 3) {
    SrcRec3 r;
    r.q[0] = C[0];
    r.q[1] = C[4];
    r.q[2] = C[8];
    r.q[3] = C[3] + C[1];
    r.q[4] = C[6] + C[2];
    r.q[5] = C[7] + C[5];
    for (int a = 0; a < 3; ++a) {
      r.mh[a] = mh[a];
      r.ml[a] = ml[a];
    }
    r.g = g;
    r.pad = 0.0f;
    reinterpret_cast<SrcRec3 *>(srcv)[j] = r;
  } else {
    SrcRec2 r;
    r.q[0] = C[0];
    r.q[1] = C[3];
    r.q[2] = C[2] + C[1];
    for (int a = 0; a < 2; ++a) {
      r.mh[a] = mh[a];
      r.ml[a] = ml[a];
    }
    r.g = g;
    r.pad = 0.0f;
    reinterpret_cast<SrcRec2 *>(srcv)[j] = r;
  }
#pragma unroll
  for (int q = 0; q < D * D; ++q) f[FA + q] = (double)(float)C[q];
}

// Both per-call preparations in one launch: thread i prepares target i
// (i < N) and source i (i < M).
template <int D>
__global__ __launch_bounds__(kCBlock) void k_compi_prep(
    const float *__restrict__ X, const float *__restrict__ XCinv, const float *__restrict__ gX, int N,
    const float *__restrict__ Y, const float *__restrict__ YCinv, const float *__restrict__ gY, int M,
    CompiPose ps, void *__restrict__ tgt, void *__restrict__ src, double *__restrict__ feat) {
  const int i = blockIdx.x * kCBlock + threadIdx.x;
  if (i < N) compi_tgt<D>(i, X, XCinv, gX, ps, tgt);
  if (i < M) compi_src<D>(i, Y, YCinv, gY, ps, src, feat);
}

// W_j = sum_i P_ij G_i.  Lane = source j (stationary, 64 per workgroup); the
// workgroup's target chunk is split into kCWaves contiguous parts, one per
// wave, streamed through SGPRs one record ahead (index clamped to the part).
// Per pair: d = n_i - m'_j in FP64, e1 = -log2(e)/2 d^T C_j d + lg_i,
// e2 = -log2(e)/2 d^T B_i d + lg_i (FP64), P 2^-o = exp2(e1 - o) + exp2(e2 - o)
// in FP32 with the lane's running scale o.  The sums are FP32 over sub-tiles
// of kCSub targets, then FP64; at the end each wave's sums are multiplied by
// 2^o in FP64 and the waves' sums meet in LDS in a fixed order:
// (w0 + w2) + (w1 + w3).
template <int D>
__global__ __launch_bounds__(kCBlock) void k_compi_w(
    const void *__restrict__ srcv, int M, const void *__restrict__ tgtv, int N,
    int chunk, double *__restrict__ part) {
  constexpr int NW = D == 3 ? 20 : 12;
  using Tgt = typename std::conditional<D == 3, TgtRec3, TgtRec2>::type;
  using Src = typename std::conditional<D == 3, SrcRec3, SrcRec2>::type;
  const Tgt *__restrict__ T = reinterpret_cast<const Tgt *>(tgtv);
  // the wave index is uniform: say so, so the target stream uses scalar loads
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int j = blockIdx.x * 64 + lane;
  const int jj = min(j, M - 1);
  const Src sr = reinterpret_cast<const Src *>(srcv)[jj];
  double m[3], cq[6];
#pragma unroll
  for (int a = 0; a < D; ++a) m[a] = (double)sr.mh[a] + (double)sr.ml[a];  // exact: m' to 2^-48
#pragma unroll
  for (int q = 0; q < (D == 3 ? 6 : 3); ++q) cq[q] = sr.q[q];
  const int c0 = blockIdx.y * chunk;
  const int c1 = min(c0 + chunk, N);
  const int per = (c1 - c0 + kCWaves - 1) / kCWaves;
  const int i0 = min(c0 + wave * per, c1);
  const int i1 = min(i0 + per, c1);
  double W[NW];
#pragma unroll
  for (int k = 0; k < NW; ++k) W[k] = 0.0;
  double o = -INFINITY;  // running scale (log2) of this lane's sums
  if (i0 < i1) {
    Tgt tn = T[i0];
    for (int s0 = i0; s0 < i1; s0 += kCSub) {
      float w[NW];
#pragma unroll
      for (int k = 0; k < NW; ++k) w[k] = 0.0f;
      const int s1 = min(s0 + kCSub, i1);
      for (int i = s0; i < s1; ++i) {
        const Tgt t = tn;
        tn = T[min(i + 1, i1 - 1)];
        double e1, e2;
        if constexpr (D == 3) {
          const double dx = (double)t.n[0] - m[0], dy = (double)t.n[1] - m[1], dz = (double)t.n[2] - m[2];
          const double xx = dx * dx, yy = dy * dy, zz = dz * dz, xy = dx * dy, xz = dx * dz, yz = dy * dz;
          const double mal1 = fma(cq[0], xx, fma(cq[1], yy, fma(cq[2], zz, fma(cq[3], xy, fma(cq[4], xz, cq[5] * yz)))));
          const double mal2 =
              fma(t.q[0], xx, fma(t.q[1], yy, fma(t.q[2], zz, fma(t.q[3], xy, fma(t.q[4], xz, t.q[5] * yz)))));
          e1 = fma(mal1, kNegHalfLog2eD, t.lg);
          e2 = fma(mal2, kNegHalfLog2eD, t.lg);
        } else {
          const double dx = (double)t.n[0] - m[0], dy = (double)t.n[1] - m[1];
          const double xx = dx * dx, yy = dy * dy, xy = dx * dy;
          const double mal1 = fma(cq[0], xx, fma(cq[1], yy, cq[2] * xy));
          const double mal2 = fma(t.q[0], xx, fma(t.q[1], yy, t.q[2] * xy));
          e1 = fma(mal1, kNegHalfLog2eD, t.lg);
          e2 = fma(mal2, kNegHalfLog2eD, t.lg);
        }
        const double em = fmax(e1, e2);
        if (em > o + kRescale) {
          // the lane's largest weight so far: rescale its sums (rare)
          const float f = __builtin_amdgcn_exp2f((float)(o - em));
#pragma unroll
          for (int k = 0; k < NW; ++k) {
            w[k] *= f;
            W[k] *= (double)f;
          }
          o = em;
        }
        const float P = __builtin_amdgcn_exp2f((float)(e1 - o)) + __builtin_amdgcn_exp2f((float)(e2 - o));
        if constexpr (D == 3) {
          w[0] += P;
#pragma unroll
          for (int a = 0; a < 3; ++a) w[1 + a] = fmaf(P, t.n[a], w[1 + a]);
#pragma unroll
          for (int a = 0; a < 6; ++a) w[4 + a] = fmaf(P, t.nn[a], w[4 + a]);
          w[10] = fmaf(P, t.s, w[10]);
#pragma unroll
          for (int a = 0; a < 3; ++a) w[11 + a] = fmaf(P, t.v[a], w[11 + a]);
#pragma unroll
          for (int a = 0; a < 6; ++a) w[14 + a] = fmaf(P, t.b[a], w[14 + a]);
        } else {
          w[0] += P;
#pragma unroll
          for (int a = 0; a < 2; ++a) w[1 + a] = fmaf(P, t.n[a], w[1 + a]);
#pragma unroll
          for (int a = 0; a < 3; ++a) w[3 + a] = fmaf(P, t.nn[a], w[3 + a]);
          w[6] = fmaf(P, t.s, w[6]);
#pragma unroll
          for (int a = 0; a < 2; ++a) w[7 + a] = fmaf(P, t.v[a], w[7 + a]);
#pragma unroll
          for (int a = 0; a < 3; ++a) w[9 + a] = fmaf(P, t.b[a], w[9 + a]);
        }
      }
#pragma unroll
      for (int k = 0; k < NW; ++k) W[k] += (double)w[k];
    }
  }
  // back to the true scale in FP64 (o = -inf: nothing was added)
  const double sc = o == -INFINITY ? 0.0 : exp2(o);
#pragma unroll
  for (int k = 0; k < NW; ++k) W[k] *= sc;
  __shared__ double red[2][NW][64];
  if (wave >= 2)
#pragma unroll
    for (int k = 0; k < NW; ++k) red[wave - 2][k][lane] = W[k];
  __syncthreads();
  if (wave < 2)
#pragma unroll
    for (int k = 0; k < NW; ++k) W[k] += red[wave][k][lane];
  __syncthreads();
  if (wave == 1)
#pragma unroll
    for (int k = 0; k < NW; ++k) red[0][k][lane] = W[k];
  __syncthreads();
  if (wave == 0 && j < M) {
    double *out = part + ((size_t)blockIdx.y * M + j) * NW;
#pragma unroll
    for (int k = 0; k < NW; ++k) out[k] = W[k] + red[0][k][lane];
  }
}

// Moments: mom[k] = sum_j prod of up to four feature entries of source j
// (index -1 = 1).  One workgroup per kMomSrc sources: their features (the W
// partials summed over the target chunks in a fixed order, times gY_j, plus
// m' and A' from k_compi_prep) are staged in LDS, then each thread
// accumulates its moments over them.  Block partials are summed later.
struct MomTab {
  short ix[400][4];
};

template <int D>
__global__ __launch_bounds__(kCBlock) void k_compi_mom(
    const double *__restrict__ wpart, int M, int nchunks,
    const void *__restrict__ srcv, const double *__restrict__ feat, int nmom,
    MomTab tab, double *__restrict__ mpart) {
  constexpr int NW = D == 3 ? 20 : 12;
  using Src = typename std::conditional<D == 3, SrcRec3, SrcRec2>::type;
  const Src *__restrict__ src = reinterpret_cast<const Src *>(srcv);
  __shared__ double fs[kMomSrc][kFeat];
  const int j0 = blockIdx.x * kMomSrc;
  for (int e = threadIdx.x; e < kMomSrc * kFeat; e += kCBlock) {
    const int jl = e / kFeat, k = e % kFeat, j = j0 + jl;
    double v = 0.0;
    if (j < M) {
      if (k < NW) {
        // eight loads in flight, summed in a fixed order
        const double *p = wpart + (size_t)j * NW + k;
        const size_t stride = (size_t)M * NW;
        double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        int c = 0;
        for (; c + 8 <= nchunks; c += 8) {
          double x[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) x[u] = p[(size_t)(c + u) * stride];
#pragma unroll
          for (int u = 0; u < 8; ++u) acc[u] += x[u];
        }
        for (; c < nchunks; ++c) acc[0] += p[(size_t)c * stride];
        const double sum = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
        v = sum * (double)src[j].g;
      } else if (k < NW + D + D * D) {  // m' and A'
        v = feat[(size_t)j * kFeat + k];
      }
    }
    fs[jl][k] = v;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nmom; k += kCBlock) {
    const int a = tab.ix[k][0], b = tab.ix[k][1], c = tab.ix[k][2], d = tab.ix[k][3];
    double s = 0.0;
    for (int jl = 0; jl < kMomSrc; ++jl) {
      double v = fs[jl][a];
      if (b >= 0) v *= fs[jl][b];
      if (c >= 0) v *= fs[jl][c];
      if (d >= 0) v *= fs[jl][d];
      s += v;
    }
    mpart[(size_t)blockIdx.x * nmom + k] = s;
  }
}

// One wave per moment: lanes take every 64th block partial, then a fixed-
// order wave sum.
__global__ __launch_bounds__(kCBlock) void k_compi_msum(
    const double *__restrict__ part, int nblocks, int nmom, double *__restrict__ out) {
  const int k = blockIdx.x * kCWaves + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (k >= nmom) return;  // uniform per wave
  double s = 0.0;
  for (int b = lane; b < nblocks; b += 64) s += part[(size_t)b * nmom + k];
  s = wave_sum(s);
  if (lane == 0) out[k] = s;
}

// ---------------------------------------------------------------------------
// Host side: the moment list and its polynomial expansion.
// ---------------------------------------------------------------------------
enum MomKind { kM0, kM1, kM2, kM3, kM4, kM5, kK1, kK2, kK3, kK4, kK5, kK6 };

struct Mom {
  MomKind kind;
  int i0, i1, i2, i3;  // index meaning depends on kind
  short ix[4];         // feature indices multiplied on the GPU
};

// Feature layout per source (doubles): W at 0.., m' at FM, A' (D x D,
// column-major) at FA.  W: [1][n D][nn sym][s][v D][S sym].
struct FeatLayout {
  int D, W1, Wn, Wnn, Ws, Wv, WS, FM, FA, nsym;
  explicit FeatLayout(int d) : D(d) {
    nsym = d * (d + 1) / 2;
    W1 = 0;
    Wn = 1;
    Wnn = Wn + d;
    Ws = Wnn + nsym;
    Wv = Ws + 1;
    WS = Wv + d;
    FM = WS + nsym;
    FA = FM + d;
  }
  int sym(int a, int b) const { return D == 3 ? sidx3(a, b) : sidx2(a, b); }
};

static std::vector<Mom> moment_list(int D) {
  FeatLayout L(D);
  std::vector<Mom> v;
  auto add = [&](MomKind k, int i0, int i1, int i2, int i3, int a, int b, int c, int d) {
    Mom m;
    m.kind = k;
    m.i0 = i0; m.i1 = i1; m.i2 = i2; m.i3 = i3;
    m.ix[0] = (short)a; m.ix[1] = (short)b; m.ix[2] = (short)c; m.ix[3] = (short)d;
    v.push_back(m);
  };
  auto A = [&](int b, int c) { return L.FA + b + D * c; };
  auto m_ = [&](int f) { return L.FM + f; };
  // d^T B d part
  add(kM0, 0, 0, 0, 0, L.Ws, -1, -1, -1);
  for (int a = 0; a < D; ++a)
    for (int b = 0; b < D; ++b) add(kM1, a, b, 0, 0, L.Wv + a, m_(b), -1, -1);
  for (int a = 0; a < D; ++a) add(kM2, a, 0, 0, 0, L.Wv + a, -1, -1, -1);
  for (int a = 0; a < D; ++a)
    for (int c = a; c < D; ++c)
      for (int b = 0; b < D; ++b)
        for (int d = b; d < D; ++d) add(kM3, a, c, b, d, L.WS + L.sym(a, c), m_(b), m_(d), -1);
  for (int a = 0; a < D; ++a)
    for (int c = a; c < D; ++c)
      for (int d = 0; d < D; ++d) add(kM4, a, c, d, 0, L.WS + L.sym(a, c), m_(d), -1, -1);
  for (int a = 0; a < D; ++a)
    for (int c = a; c < D; ++c) add(kM5, a, c, 0, 0, L.WS + L.sym(a, c), -1, -1, -1);
  // u^T A' u part
  for (int a = 0; a < D; ++a)
    for (int e = a; e < D; ++e)
      for (int b = 0; b < D; ++b)
        for (int c = 0; c < D; ++c) add(kK1, a, e, b, c, L.Wnn + L.sym(a, e), A(b, c), -1, -1);
  for (int x = 0; x < D; ++x)
    for (int b = 0; b < D; ++b)
      for (int c = 0; c < D; ++c) add(kK2, x, b, c, 0, L.Wn + x, A(b, c), -1, -1);
  for (int x = 0; x < D; ++x)
    for (int f = 0; f < D; ++f)
      for (int b = 0; b < D; ++b)
        for (int c = 0; c < D; ++c) add(kK3, x, f, b, c, L.Wn + x, m_(f), A(b, c), -1);
  for (int b = 0; b < D; ++b)
    for (int c = 0; c < D; ++c) add(kK4, b, c, 0, 0, L.W1, A(b, c), -1, -1);
  for (int f = 0; f < D; ++f)
    for (int b = 0; b < D; ++b)
      for (int c = 0; c < D; ++c) add(kK5, f, b, c, 0, L.W1, m_(f), A(b, c), -1);
  for (int f = 0; f < D; ++f)
    for (int g = f; g < D; ++g)
      for (int b = 0; b < D; ++b)
        for (int c = 0; c < D; ++c) add(kK6, f, g, b, c, L.W1, m_(f), m_(g), A(b, c));
  return v;
}

// Polynomial in the update variables, keyed by the sorted variable list.
struct Poly {
  int nv;
  std::vector<double> c;
  explicit Poly(int n) : nv(n), c((size_t)(n + 1) * (n + 1) * (n + 1) * (n + 1), 0.0) {}
  int key(int *v, int deg) const {
    std::sort(v, v + deg);
    int k = 0;
    for (int q = 0; q < 4; ++q) k = k * (nv + 1) + (q < deg ? v[q] : nv);
    return k;
  }
};

// A signed single variable: dt_a or an entry of dR.
struct SVar {
  int var;
  double sgn;
};

struct Vars {
  int D;
  SVar dt(int a) const { return {a, 1.0}; }
  SVar dR(int a, int b) const {
    if (D == 3) return {3 + 3 * a + b, 1.0};
    // DUGMA.m:67: dR = [[r11, -r21], [r21, r11]], r11 = var 2, r21 = var 3
    static const int vv[2][2] = {{2, 3}, {3, 2}};
    static const double ss[2][2] = {{1.0, -1.0}, {1.0, 1.0}};
    return {vv[a][b], ss[a][b]};
  }
};

static void addt(Poly &P, double coef, std::initializer_list<SVar> vs) {
  int v[4], deg = 0;
  double s = coef;
  for (const SVar &x : vs) {
    v[deg++] = x.var;
    s *= x.sgn;
  }
  P.c[P.key(v, deg)] += s;
}

static void expand(int D, const std::vector<Mom> &moms, const double *val, Poly &P) {
  Vars V{D};
  for (size_t k = 0; k < moms.size(); ++k) {
    const Mom &m = moms[k];
    const double x = val[k];
    if (x == 0.0) continue;
    switch (m.kind) {
      case kM0:
        addt(P, x, {});
        break;
      case kM1:  // - sum_ab dR_ab m_b v_a
        addt(P, -x, {V.dR(m.i0, m.i1)});
        break;
      case kM2:  // - dt_a v_a
        addt(P, -x, {V.dt(m.i0)});
        break;
      case kM3: {  // dR_ab dR_cd S_ac m_b m_d over ordered (a,c), (b,d)
        const int a = m.i0, c = m.i1, b = m.i2, d = m.i3;
        for (int p = 0; p < (a == c ? 1 : 2); ++p)
          for (int q = 0; q < (b == d ? 1 : 2); ++q) {
            const int aa = p ? c : a, cc = p ? a : c, bb = q ? d : b, dd = q ? b : d;
            addt(P, x, {V.dR(aa, bb), V.dR(cc, dd)});
          }
        break;
      }
      case kM4: {  // 2 dt_a dR_cd S_ac m_d over ordered (a,c)
        const int a = m.i0, c = m.i1, d = m.i2;
        for (int p = 0; p < (a == c ? 1 : 2); ++p) {
          const int aa = p ? c : a, cc = p ? a : c;
          addt(P, 2.0 * x, {V.dt(aa), V.dR(cc, d)});
        }
        break;
      }
      case kM5: {  // dt_a dt_c S_ac over ordered (a,c)
        const int a = m.i0, c = m.i1;
        for (int p = 0; p < (a == c ? 1 : 2); ++p) addt(P, x, {V.dt(p ? c : a), V.dt(p ? a : c)});
        break;
      }
      case kK1: {  // dR_ab dR_ec nn_ae A'_bc over ordered (a,e)
        const int a = m.i0, e = m.i1, b = m.i2, c = m.i3;
        for (int p = 0; p < (a == e ? 1 : 2); ++p)
          addt(P, x, {V.dR(p ? e : a, b), V.dR(p ? a : e, c)});
        break;
      }
      case kK2: {  // -(dR_xb dR_ec dt_e + dR_ab dR_xc dt_a)
        const int xx = m.i0, b = m.i1, c = m.i2;
        for (int e = 0; e < D; ++e) addt(P, -x, {V.dR(xx, b), V.dR(e, c), V.dt(e)});
        for (int a = 0; a < D; ++a) addt(P, -x, {V.dR(a, b), V.dR(xx, c), V.dt(a)});
        break;
      }
      case kK3: {  // -(dR_xb dR_ec dR_ef + dR_ab dR_xc dR_af)
        const int xx = m.i0, f = m.i1, b = m.i2, c = m.i3;
        for (int e = 0; e < D; ++e) addt(P, -x, {V.dR(xx, b), V.dR(e, c), V.dR(e, f)});
        for (int a = 0; a < D; ++a) addt(P, -x, {V.dR(a, b), V.dR(xx, c), V.dR(a, f)});
        break;
      }
      case kK4: {  // dR_ab dR_ec dt_a dt_e
        const int b = m.i0, c = m.i1;
        for (int a = 0; a < D; ++a)
          for (int e = 0; e < D; ++e) addt(P, x, {V.dR(a, b), V.dR(e, c), V.dt(a), V.dt(e)});
        break;
      }
      case kK5: {  // dR_ab dR_ec (dt_a dR_ef + dR_af dt_e)
        const int f = m.i0, b = m.i1, c = m.i2;
        for (int a = 0; a < D; ++a)
          for (int e = 0; e < D; ++e) {
            addt(P, x, {V.dR(a, b), V.dR(e, c), V.dt(a), V.dR(e, f)});
            addt(P, x, {V.dR(a, b), V.dR(e, c), V.dR(a, f), V.dt(e)});
          }
        break;
      }
      case kK6: {  // dR_ab dR_ec dR_af dR_eg over ordered (f,g)
        const int f = m.i0, g = m.i1, b = m.i2, c = m.i3;
        for (int p = 0; p < (f == g ? 1 : 2); ++p) {
          const int ff = p ? g : f, gg = p ? f : g;
          for (int a = 0; a < D; ++a)
            for (int e = 0; e < D; ++e) addt(P, x, {V.dR(a, b), V.dR(e, c), V.dR(a, ff), V.dR(e, gg)});
        }
        break;
      }
    }
  }
}

// The moment -> coefficient map is linear and depends only on D: coef[k] =
// sum over (k, moment, factor) entries of factor * moment.  Built once per D
// by expanding each moment alone with value 1; entries are ordered by
// coefficient, then moment, so each coefficient sums its moments in order.
struct ExpMap {
  std::vector<int> start;  // CSR over coefficients (monomial order)
  std::vector<int> mom;
  std::vector<double> fac;
};

static void build_map(int D, ExpMap &E) {
  const std::vector<Mom> moms = moment_list(D);
  const int nv = D == 3 ? 12 : 4;
  const int nm = D == 3 ? DUGMA_NMONO3 : DUGMA_NMONO2;
  Poly P(nv);
  std::vector<int> key(nm);
  for (int k = 0; k < nm; ++k) {
    const unsigned char *e = D == 3 ? DUGMA_MONO3[k] : DUGMA_MONO2[k];
    int v[4], deg = 0;
    for (int x = 0; x < nv; ++x)
      for (int r = 0; r < e[x]; ++r) v[deg++] = x;
    key[k] = P.key(v, deg);
  }
  std::vector<std::vector<std::pair<int, double>>> rows(nm);
  std::vector<double> one(moms.size(), 0.0);
  for (size_t m = 0; m < moms.size(); ++m) {
    one[m] = 1.0;
    std::fill(P.c.begin(), P.c.end(), 0.0);
    expand(D, moms, one.data(), P);
    one[m] = 0.0;
    for (int k = 0; k < nm; ++k)
      if (P.c[key[k]] != 0.0) rows[k].push_back({(int)m, P.c[key[k]]});
  }
  E.start.assign(nm + 1, 0);
  for (int k = 0; k < nm; ++k) {
    E.start[k + 1] = E.start[k] + (int)rows[k].size();
    for (const auto &r : rows[k]) {
      E.mom.push_back(r.first);
      E.fac.push_back(r.second);
    }
  }
}

static const ExpMap &expansion_map(int D) {
  static ExpMap maps[2];
  static std::once_flag once[2];
  std::call_once(once[D - 2], [D] { build_map(D, maps[D - 2]); });
  return maps[D - 2];
}

static int check_dims(dugma_ctx *c, double N, double M, double D) {
  if ((int)N != c->N || (int)M != c->M || (int)D != c->D)
    return fail(DUGMA_ERR_STATE, "N, M, D do not match the context");
  return DUGMA_OK;
}

}  // namespace dugma

using namespace dugma;

extern "C" int dugma_compi(dugma_ctx *ctx, const double *R, const double *t,
                           double N, double M, double D, double sig,
                           double *coef) {
  if (!ctx || !R || !t || !coef) return fail(DUGMA_ERR_ARG, "dugma_compi: null argument");
  int rc = check_dims(ctx, N, M, D);
  if (rc) return rc;
  if (!(sig > 0) || !std::isfinite(sig)) return fail(DUGMA_ERR_ARG, "dugma_compi: sig must be positive");
  dugma_ctx *c = ctx;
  const int Dd = c->D, n = c->N, m = c->M;
  CompiPose ps;
  memset(&ps, 0, sizeof(ps));
  for (int a = 0; a < Dd * Dd; ++a) ps.R[a] = R[a];
  for (int a = 0; a < Dd; ++a) ps.t[a] = t[a];
  const double is = 1.0 / sig;
  float f = (float)is;
  if (Dd == 2) {
    if ((double)f > is) f = std::nextafter(f, -INFINITY);  // __double2float_rd
    ps.Isig = f;
    ps.gsc = f;
  } else {
    ps.Isig = f;
    ps.gsc = std::sqrt(f * f * f);  // FP32 as compi.cu:662
  }

  const std::vector<Mom> moms = moment_list(Dd);
  const int nmom = (int)moms.size();
  const int NW = Dd == 3 ? 20 : 12;
  const size_t trec = Dd == 3 ? sizeof(TgtRec3) : sizeof(TgtRec2);
  const size_t srec = Dd == 3 ? sizeof(SrcRec3) : sizeof(SrcRec2);
  // launch shapes: 64 sources per workgroup of k_compi_w, target chunks so
  // that about 2,048 workgroups run, each wave with at least 64 targets
  const int sblocks = (m + 63) / 64;
  int nch = (2048 + sblocks - 1) / sblocks;
  nch = std::max(1, std::min(nch, std::max(1, n / (kCWaves * 64))));
  const int chunk = (n + nch - 1) / nch;
  const int nchunks = (n + chunk - 1) / chunk;
  const int mblocks = (m + kMomSrc - 1) / kMomSrc;
  // scratch: tgt records | src | feat | W partials | moment partials | moments
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t o_src = al(trec * n);
  const size_t o_feat = o_src + al(srec * (size_t)m);
  const size_t o_part = o_feat + al(sizeof(double) * kFeat * (size_t)m);
  const size_t o_mp = o_part + al(sizeof(double) * (size_t)nchunks * m * NW);
  const size_t o_mom = o_mp + al(sizeof(double) * (size_t)mblocks * nmom);
  rc = c->scratch2.ensure(o_mom + al(sizeof(double) * nmom));
  if (rc) return rc;
  char *base = (char *)c->scratch2.ptr;
  void *tgt = base;
  void *src = base + o_src;
  double *feat = (double *)(base + o_feat);
  double *wpart = (double *)(base + o_part);
  double *mpart = (double *)(base + o_mp);
  MomTab tab;
  memset(&tab, 0xff, sizeof(tab));
  for (int k = 0; k < nmom; ++k)
    for (int q = 0; q < 4; ++q) tab.ix[k][q] = moms[k].ix[q];
  const ExpMap &E = expansion_map(Dd);

  dim3 blk(kCBlock);
  const unsigned pgrid = (unsigned)((std::max(n, m) + kCBlock - 1) / kCBlock);
  if (Dd == 3) {
    hipLaunchKernelGGL(k_compi_prep<3>, dim3(pgrid), blk, 0, c->stream, c->dX, c->dXCinv, c->dgX, n, c->dY,
                       c->dYCinv, c->dgY, m, ps, tgt, src, feat);
    hipLaunchKernelGGL(k_compi_w<3>, dim3(sblocks, nchunks), blk, 0, c->stream, src, m, tgt, n, chunk, wpart);
    hipLaunchKernelGGL(k_compi_mom<3>, dim3(mblocks), blk, 0, c->stream, wpart, m, nchunks, src, feat, nmom, tab, mpart);
  } else {
    hipLaunchKernelGGL(k_compi_prep<2>, dim3(pgrid), blk, 0, c->stream, c->dX, c->dXCinv, c->dgX, n, c->dY,
                       c->dYCinv, c->dgY, m, ps, tgt, src, feat);
    hipLaunchKernelGGL(k_compi_w<2>, dim3(sblocks, nchunks), blk, 0, c->stream, src, m, tgt, n, chunk, wpart);
    hipLaunchKernelGGL(k_compi_mom<2>, dim3(mblocks), blk, 0, c->stream, wpart, m, nchunks, src, feat, nmom, tab, mpart);
  }
  // moment sums straight into the context's pinned result slots
  double *slots = c->hres + dugma_ctx::kResMoments;
  preset_slots(slots, nmom);
  hipLaunchKernelGGL(k_compi_msum, dim3((nmom + kCWaves - 1) / kCWaves), blk, 0, c->stream, mpart, mblocks, nmom,
                     c->dres + dugma_ctx::kResMoments);
  DUGMA_HIP(hipGetLastError());
  const int rcw = wait_slots(c, slots, nmom);
  if (rcw) return rcw;
  const volatile double *hm = slots;

  const int nm = Dd == 3 ? DUGMA_NMONO3 : DUGMA_NMONO2;
  for (int k = 0; k < nm; ++k) {
    double v = 0.0;
    for (int e = E.start[k]; e < E.start[k + 1]; ++e) v += E.fac[e] * hm[E.mom[e]];
    coef[k] = v;
  }
  return DUGMA_OK;
}
