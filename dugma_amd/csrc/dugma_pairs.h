// dugma_pairs.h -- per-pair arithmetic of the convolved-uncertainty energy
// and its adjoint, for the four point-record kinds.
//
// For a stationary point p (in VGPRs) and a streamed point q (wave-uniform,
// in SGPRs): d = p - q, S = Sp + Sq,
//   e = det(S)^-1/2 exp(-d^T S^-1 d / 2)          (times (2 pi)^-D/2 later)
// and, with w = S^-1 d, the adjoint accumulators
//   EW += e w,  B += e (w w^T - S^-1)              (full covariance)
//   EW += e w,  G  += (e/s)(|d|^2/s - D)          (isotropic, S = s I)
// from which dE/dp = -c EW and dE/dSp = c B / 2, dE/ds_p = c G / 2.
// FP32 throughout; one v_rsq (or v_rcp) and one v_exp per pair.
#pragma once

#include "dugma_internal.h"

namespace dugma {

constexpr float kNegHalfLog2e = -0.72134752044448170368f;  // -log2(e)/2

// Isotropic records hold their variance times ln 4 = 2 ln 2, so that
// exp2(-|d|^2 / s') = exp(-|d|^2 / 2s) takes no constant multiply per pair.
// Per pair, e then comes out short by (ln 4)^(D/2) and e/s by
// (ln 4)^(D/2 + 1); the kernel epilogue restores both in FP64 (ESCALE,
// GSCALE below).  The full-covariance records keep plain covariances: their
// adjoint term S^-1 d d^T S^-1 - S^-1 would not scale uniformly.
constexpr double kLn4 = 1.3862943611198906;
constexpr float kLn4f = 1.38629436f;
constexpr double kLn4Pow1p5 = 1.632236874939246;
constexpr double kLn4Pow2 = 1.9218120556728056;
constexpr double kLn4Pow2p5 = 2.2627607757402286;

__device__ __forceinline__ float fexp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Sum of v over the 64 lanes of a wave, returned in every lane.  DPP row
// rotations by 8, 4, 2, 1 sum each row of 16 lanes in registers (no LDS
// traffic), then the four row sums are read from lanes 0, 16, 32, 48 and
// added in a fixed order, so the result is deterministic.
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_move<0x128>(v);  // row_ror:8
  v += dpp_move<0x124>(v);  // row_ror:4
  v += dpp_move<0x122>(v);  // row_ror:2
  v += dpp_move<0x121>(v);  // row_ror:1
  const int lo = __double2loint(v), hi = __double2hiint(v);
  const double r0 = __hiloint2double(__builtin_amdgcn_readlane(hi, 0), __builtin_amdgcn_readlane(lo, 0));
  const double r1 = __hiloint2double(__builtin_amdgcn_readlane(hi, 16), __builtin_amdgcn_readlane(lo, 16));
  const double r2 = __hiloint2double(__builtin_amdgcn_readlane(hi, 32), __builtin_amdgcn_readlane(lo, 32));
  const double r3 = __hiloint2double(__builtin_amdgcn_readlane(hi, 48), __builtin_amdgcn_readlane(lo, 48));
  return (r0 + r1) + (r2 + r3);
}

// Sums of v over the low and the high 32 lanes of a wave, in every lane:
// DPP row rotations by 8, 4, 2, 1 sum each row of 16 lanes, then rows 0 + 1
// and rows 2 + 3 are added from lanes 0, 16, 32, 48 (fixed order).
__device__ __forceinline__ void half_wave_sums(double v, double &lo, double &hi) {
  v += dpp_move<0x128>(v);  // row_ror:8
  v += dpp_move<0x124>(v);  // row_ror:4
  v += dpp_move<0x122>(v);  // row_ror:2
  v += dpp_move<0x121>(v);  // row_ror:1
  const int l = __double2loint(v), h = __double2hiint(v);
  const double r0 = __hiloint2double(__builtin_amdgcn_readlane(h, 0), __builtin_amdgcn_readlane(l, 0));
  const double r1 = __hiloint2double(__builtin_amdgcn_readlane(h, 16), __builtin_amdgcn_readlane(l, 16));
  const double r2 = __hiloint2double(__builtin_amdgcn_readlane(h, 32), __builtin_amdgcn_readlane(l, 32));
  const double r3 = __hiloint2double(__builtin_amdgcn_readlane(h, 48), __builtin_amdgcn_readlane(l, 48));
  lo = r0 + r1;
  hi = r2 + r3;
}

// ESCALE / GSCALE: factors that restore the energy / the gradient
// accumulators of a record kind (1 for full covariance).
template <class Rec> struct RecTraits;
template <> struct RecTraits<Rec3F> {
  static constexpr int D = 3, NACC = kAcc3F;
  static constexpr bool FULL = true;
  static constexpr double ESCALE = 1.0, GSCALE = 1.0;
};
template <> struct RecTraits<Rec3I> {
  static constexpr int D = 3, NACC = kAcc3I;
  static constexpr bool FULL = false;
  static constexpr double ESCALE = kLn4Pow1p5, GSCALE = kLn4Pow2p5;
};
template <> struct RecTraits<Rec2F> {
  static constexpr int D = 2, NACC = kAcc2F;
  static constexpr bool FULL = true;
  static constexpr double ESCALE = 1.0, GSCALE = 1.0;
};
template <> struct RecTraits<Rec2I> {
  static constexpr int D = 2, NACC = kAcc2I;
  static constexpr bool FULL = false;
  static constexpr double ESCALE = kLn4, GSCALE = kLn4Pow2;
};

// GRAD: 0 energy only; 1 pose adjoint; 2 per-point adjoint (isotropic
// records then also accumulate G).  COMP: how the FP32 hi/lo position split
// enters d = p - q:  0 hi parts only; 1 only the streamed point has a lo part
// (d = (p.h - q.h) - q.l); 2 only the stationary point has one
// (d = (p.h - q.h) + p.l).  p.h - q.h is exact for nearby points, so d keeps
// the low bits of the FP64 source transform.
#define DUGMA_DIFF(c)                            \
  float d##c = p.h##c - q.h##c;                  \
  if (COMP == 1) d##c = d##c - q.l##c;           \
  if (COMP == 2) d##c = d##c + p.l##c;

// ---- 3-D, full covariance -------------------------------------------------
template <int GRAD, int COMP>
__device__ __forceinline__ void pair_eval(const Rec3F &p, const Rec3F &q, float *acc) {
  DUGMA_DIFF(x)
  DUGMA_DIFF(y)
  DUGMA_DIFF(z)
  const float a11 = p.s11 + q.s11, a12 = p.s12 + q.s12, a13 = p.s13 + q.s13;
  const float a22 = p.s22 + q.s22, a23 = p.s23 + q.s23, a33 = p.s33 + q.s33;
  // adjugate (symmetric) and determinant
  const float c11 = fmaf(a22, a33, -(a23 * a23));
  const float c12 = fmaf(a13, a23, -(a12 * a33));
  const float c13 = fmaf(a12, a23, -(a13 * a22));
  const float c22 = fmaf(a11, a33, -(a13 * a13));
  const float c23 = fmaf(a12, a13, -(a11 * a23));
  const float c33 = fmaf(a11, a22, -(a12 * a12));
  const float det = fmaf(a11, c11, fmaf(a12, c12, a13 * c13));
  const float u1 = fmaf(c11, dx, fmaf(c12, dy, c13 * dz));
  const float u2 = fmaf(c12, dx, fmaf(c22, dy, c23 * dz));
  const float u3 = fmaf(c13, dx, fmaf(c23, dy, c33 * dz));
  const float qa = fmaf(dx, u1, fmaf(dy, u2, dz * u3));  // d^T adj(S) d
  const float r = frsq(det);
  const float r2 = r * r;  // 1/det
  const float e = fexp2((qa * r2) * kNegHalfLog2e) * r;
  acc[0] += e;
  if (GRAD) {
    // w = u/det = v:  EW += e w = er2 u,  B += e (w w^T - adj/det)
    // = er2 (v u^T - adj), one accumulator per entry of the symmetric B
    const float er2 = e * r2;
    acc[1] = fmaf(er2, u1, acc[1]);
    acc[2] = fmaf(er2, u2, acc[2]);
    acc[3] = fmaf(er2, u3, acc[3]);
    const float v1 = r2 * u1, v2 = r2 * u2, v3 = r2 * u3;
    acc[4] = fmaf(er2, fmaf(v1, u1, -c11), acc[4]);
    acc[5] = fmaf(er2, fmaf(v1, u2, -c12), acc[5]);
    acc[6] = fmaf(er2, fmaf(v1, u3, -c13), acc[6]);
    acc[7] = fmaf(er2, fmaf(v2, u2, -c22), acc[7]);
    acc[8] = fmaf(er2, fmaf(v2, u3, -c23), acc[8]);
    acc[9] = fmaf(er2, fmaf(v3, u3, -c33), acc[9]);
  }
}

// ---- 3-D, isotropic covariance --------------------------------------------
template <int GRAD, int COMP>
__device__ __forceinline__ void pair_eval(const Rec3I &p, const Rec3I &q, float *acc) {
  DUGMA_DIFF(x)
  DUGMA_DIFF(y)
  DUGMA_DIFF(z)
  const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
  const float s = p.s + q.s;  // ln 4 times the variance
  const float rs = frsq(s);
  const float inv = rs * rs;  // 1/s
  const float e = fexp2(-(d2 * inv)) * (rs * inv);  // s^-3/2
  acc[0] += e;
  if (GRAD) {
    const float einv = e * inv;
    acc[1] = fmaf(einv, dx, acc[1]);
    acc[2] = fmaf(einv, dy, acc[2]);
    acc[3] = fmaf(einv, dz, acc[3]);
    if (GRAD == 2) acc[4] = fmaf(einv, fmaf(d2 * inv, kLn4f, -3.0f), acc[4]);
    if (GRAD == 3) {
      // B = sum e (w w^T - S^-1) in two parts (the ln 4 restored with the
      // gradient's): [4..9] sum (e / s'^2) ln 4 d d^T, upper triangle row by
      // row, and [10] sum e / s'
      const float cw = einv * inv * kLn4f;
      const float cx = cw * dx, cy = cw * dy, cz = cw * dz;
      acc[4] = fmaf(cx, dx, acc[4]);
      acc[5] = fmaf(cx, dy, acc[5]);
      acc[6] = fmaf(cx, dz, acc[6]);
      acc[7] = fmaf(cy, dy, acc[7]);
      acc[8] = fmaf(cy, dz, acc[8]);
      acc[9] = fmaf(cz, dz, acc[9]);
      acc[10] += einv;
    }
  }
}

// ---- 2-D, full covariance -------------------------------------------------
template <int GRAD, int COMP>
__device__ __forceinline__ void pair_eval(const Rec2F &p, const Rec2F &q, float *acc) {
  DUGMA_DIFF(x)
  DUGMA_DIFF(y)
  const float a11 = p.s11 + q.s11, a12 = p.s12 + q.s12, a22 = p.s22 + q.s22;
  const float det = fmaf(a11, a22, -(a12 * a12));
  const float u1 = fmaf(a22, dx, -(a12 * dy));
  const float u2 = fmaf(a11, dy, -(a12 * dx));
  const float qa = fmaf(dx, u1, dy * u2);
  const float r = frsq(det);
  const float r2 = r * r;
  const float e = fexp2((qa * r2) * kNegHalfLog2e) * r;
  acc[0] += e;
  if (GRAD) {
    // adj = [[a22, -a12], [-a12, a11]];  B += er2 (v u^T - adj), v = u/det
    const float er2 = e * r2;
    acc[1] = fmaf(er2, u1, acc[1]);
    acc[2] = fmaf(er2, u2, acc[2]);
    const float v1 = r2 * u1, v2 = r2 * u2;
    acc[3] = fmaf(er2, fmaf(v1, u1, -a22), acc[3]);
    acc[4] = fmaf(er2, fmaf(v1, u2, a12), acc[4]);
    acc[5] = fmaf(er2, fmaf(v2, u2, -a11), acc[5]);
  }
}

// ---- 2-D, isotropic covariance --------------------------------------------
template <int GRAD, int COMP>
__device__ __forceinline__ void pair_eval(const Rec2I &p, const Rec2I &q, float *acc) {
  DUGMA_DIFF(x)
  DUGMA_DIFF(y)
  const float d2 = fmaf(dx, dx, dy * dy);
  const float inv = frcp(p.s + q.s);  // 1 / (ln 4 times the variance)
  const float e = fexp2(-(d2 * inv)) * inv;  // s^-1
  acc[0] += e;
  if (GRAD) {
    const float einv = e * inv;
    acc[1] = fmaf(einv, dx, acc[1]);
    acc[2] = fmaf(einv, dy, acc[2]);
    if (GRAD == 2) acc[3] = fmaf(einv, fmaf(d2 * inv, kLn4f, -2.0f), acc[3]);
  }
}

#undef DUGMA_DIFF

// ---- pose-gradient contribution of one stationary target ------------------
// The energy is invariant under one rigid motion of both clouds, so the
// gradient for moving the sources (rotation about t, then translation) is
// minus the gradient for moving the targets the same way:
//   dE/dt = c sum_i EW_i
//   dE/dw = c sum_i [ (y_i - t) x EW_i + vee(Sy_i B_i - B_i Sy_i) ]
// (2-D: dE/dtheta = c sum_i [ EW^T J (y - t) - ((Sy B)_12 - (B Sy)_12) ]).
// r: [E, rotation dofs, translation dofs].
__device__ __forceinline__ void pose_contrib(const Rec3F &p, const double *a,
                                             const double *t, double *r) {
  const double px = (double)p.hx + (double)p.lx - t[0];
  const double py = (double)p.hy + (double)p.ly - t[1];
  const double pz = (double)p.hz + (double)p.lz - t[2];
  const double ex = a[1], ey = a[2], ez = a[3];
  const double b11 = a[4], b12 = a[5], b13 = a[6];
  const double b22 = a[7], b23 = a[8], b33 = a[9];
  const double s11 = p.s11, s12 = p.s12, s13 = p.s13, s22 = p.s22, s23 = p.s23, s33 = p.s33;
  // K = S B - B S (antisymmetric); vee(K) = (K32, K13, K21)
  const double sb32 = s13 * b12 + s23 * b22 + s33 * b23;  // (S B)_32
  const double bs32 = b13 * s12 + b23 * s22 + b33 * s23;  // (B S)_32
  const double sb13 = s11 * b13 + s12 * b23 + s13 * b33;
  const double bs13 = b11 * s13 + b12 * s23 + b13 * s33;
  const double sb21 = s12 * b11 + s22 * b12 + s23 * b13;
  const double bs21 = b12 * s11 + b22 * s12 + b23 * s13;
  r[0] += a[0];
  r[1] += py * ez - pz * ey + (sb32 - bs32);
  r[2] += pz * ex - px * ez + (sb13 - bs13);
  r[3] += px * ey - py * ex + (sb21 - bs21);
  r[4] += ex;
  r[5] += ey;
  r[6] += ez;
}

__device__ __forceinline__ void pose_contrib(const Rec3I &p, const double *a,
                                             const double *t, double *r) {
  const double px = (double)p.hx + (double)p.lx - t[0];
  const double py = (double)p.hy + (double)p.ly - t[1];
  const double pz = (double)p.hz + (double)p.lz - t[2];
  const double ex = a[1], ey = a[2], ez = a[3];
  r[0] += a[0];
  r[1] += py * ez - pz * ey;
  r[2] += pz * ex - px * ez;
  r[3] += px * ey - py * ex;
  r[4] += ex;
  r[5] += ey;
  r[6] += ez;
}

__device__ __forceinline__ void pose_contrib(const Rec2F &p, const double *a,
                                             const double *t, double *r) {
  const double px = (double)p.hx + (double)p.lx - t[0];
  const double py = (double)p.hy + (double)p.ly - t[1];
  const double ex = a[1], ey = a[2];
  const double b11 = a[3], b12 = a[4], b22 = a[5];
  const double s11 = p.s11, s12 = p.s12, s22 = p.s22;
  r[0] += a[0];
  r[1] += ey * px - ex * py - ((s11 * b12 + s12 * b22) - (b11 * s12 + b12 * s22));
  r[2] += ex;
  r[3] += ey;
}

__device__ __forceinline__ void pose_contrib(const Rec2I &p, const double *a,
                                             const double *t, double *r) {
  const double px = (double)p.hx + (double)p.lx - t[0];
  const double py = (double)p.hy + (double)p.ly - t[1];
  r[0] += a[0];
  r[1] += a[2] * px - a[1] * py;
  r[2] += a[1];
  r[3] += a[2];
}

}  // namespace dugma
