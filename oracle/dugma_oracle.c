/*
 * dugma_oracle.c -- CPU restatement of the DUGMA registration hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle the HIP path is tested against, never
 * the product): only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load this file's library.  Built with -ffp-contract=off so every FP32 operation written
 * below rounds exactly once, in the order written.
 *
 * Pinning (what checks each function beyond restating it):
 *  dor_prep_inverse   closed forms (identity, diagonal, inverse*C = I) and the
 *                     Kinect NaN / 1e6-variance fixtures
 *  dor_update_sig     brute force in numpy with the same FP32 order; closed
 *                     forms (Y = X, R = I -> 0; single shifted pair)
 *  dor_knn            numpy brute force with a stable (distance, index) sort
 *  dor_knn_sigma      closed forms on lattices (mean of the k nearest
 *                     distances, self excluded) and the dor_knn lists
 *  dor_energy         scipy.stats.multivariate_normal.pdf per pair, rigid and
 *                     permutation invariance, isotropic closed form,
 *                     Gaussian-product quadrature; gradients by central
 *                     differences and by float64 torch autograd
 *  dor_compi          golden vectors evaluated from compi.cu's own per-pair
 *                     coefficient expressions, and F(dT) == brute force
 *                     surrogate at random dT
 *  dor_compi_moments  equal, coefficient by coefficient, to dor_compi and to
 *                     the golden vectors; end to end through the DUGMA.m loop,
 *                     which with it reproduces the reference's published
 *                     Kinect poses
 *  dor_transform_points  closed forms through dor_update_sig (Y = X, R = I
 *                     gives 0) and the brute-force nearest-distance check
 *  dor_tree_sum_f32   a hand-worked 5-element case on which the pairwise
 *                     order and a left-to-right sum differ
 *  dor_energy_fd_pose central differences against the analytic pose
 *                     gradient, which float64 torch autograd also checks
 *  dor_pair_weight    summed over all pairs at the identity pose, equal to
 *                     the surrogate's value there
 *  dor_surrogate_eval equal to the 637 / 45 coefficient polynomial at
 *                     random dT (the polynomial identity)
 *  dor_poly_eval      closed-form monomial evaluation (tests/test_poly.py)
 *  dor_constraints    zero on rotations, non-zero off them, the 2-D form
 *                     on a unit vector
 */
#include "dugma_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../dugma_amd/csrc/dugma_monomials.inc"

int dor_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------ */
/* init: inverse covariance and g = sqrtf(1/det)   (init.cu:22-190)    */
/* ------------------------------------------------------------------ */
int dor_prep_inverse(int D, int n, const double *C, float *inv, float *g) {
  int bad = 0;
  for (int k = 0; k < n; ++k) {
    const double *c = C + (size_t)k * D * D;
    float *o = inv + (size_t)k * D * D;
    if (D == 2) {
      /* init.cu:33-57 */
      float a1 = (float)c[0], a2 = (float)c[1], a3 = (float)c[2], a4 = (float)c[3];
      float d = a1 * a4 - a2 * a3;
      if (d > 0) {
        o[0] = a4 / d;
        o[1] = (-a2) / d;
        o[2] = (-a3) / d;
        o[3] = a1 / d;
        g[k] = sqrtf(1.0f / d);
      } else {
        for (int q = 0; q < 4; ++q) o[q] = 100000.0f;
        g[k] = 0.0f;
        ++bad;
      }
    } else {
      /* init.cu:62-100, column-major a_rc = c[(r-1) + 3(c-1)] */
      float a11 = (float)c[0], a21 = (float)c[1], a31 = (float)c[2];
      float a12 = (float)c[3], a22 = (float)c[4], a32 = (float)c[5];
      float a13 = (float)c[6], a23 = (float)c[7], a33 = (float)c[8];
      float d = a11 * a22 * a33 - a11 * a23 * a32 - a12 * a21 * a33 +
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
        g[k] = sqrtf(1.0f / d);
      } else {
        for (int q = 0; q < 9; ++q) o[q] = 100000.0f;
        g[k] = 0.0f;
        ++bad;
      }
    }
  }
  return bad;
}

/* ------------------------------------------------------------------ */
/* update_sig: quantised 1-NN mean squared distance                   */
/* ------------------------------------------------------------------ */
void dor_transform_points(int D, int M, const float *Yf, const double *R,
                          const double *t, float *out) {
  for (int j = 0; j < M; ++j) {
    const float *y = Yf + (size_t)j * D;
    for (int a = 0; a < D; ++a) {
      /* row a of R (column-major): R[a + D*b] */
      double s = R[a] * (double)y[0];
      for (int b = 1; b < D; ++b) s = s + R[a + D * b] * (double)y[b];
      s = s + t[a];
      out[(size_t)j * D + a] = (float)s;
    }
  }
}

/* __float2uint_rn: round half to even, saturating; NaN and negatives -> 0. */
static unsigned f2u_rn(float s) {
  if (!(s > 0.0f)) return 0u;
  if (s >= 4294967296.0f) return 4294967295u;
  return (unsigned)rintf(s);
}

float dor_tree_sum_f32(int n, const float *v) {
  if (n <= 0) return 0.0f;
  float *a = (float *)malloc(sizeof(float) * (size_t)n);
  memcpy(a, v, sizeof(float) * (size_t)n);
  long L = n, stride = 1;
  while (L > 1) {
    long half = (L + 1) / 2;
    for (long idx = 0; idx < half; ++idx) {
      if ((L & 1) == 0 || idx != half - 1)
        a[idx * 2 * stride] = a[idx * 2 * stride] + a[idx * 2 * stride + stride];
    }
    L = half;
    stride *= 2;
  }
  float s = a[0];
  free(a);
  return s;
}

float dor_update_sig(int D, int N, const float *Xf, int M, const float *Yf,
                     const double *R, const double *t, unsigned *q) {
  float *m = (float *)malloc(sizeof(float) * (size_t)M * D + 4);
  dor_transform_points(D, M, Yf, R, t, m);
  float *f = (float *)malloc(sizeof(float) * (size_t)N + 4);
  const float DD = (float)D;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; ++i) {
    const float *n = Xf + (size_t)i * D;
    unsigned best = 4000000000u;
    for (int j = 0; j < M; ++j) {
      const float *mj = m + (size_t)j * D;
      float e0 = n[0] - mj[0], e1 = n[1] - mj[1];
      float s2 = e0 * e0 + e1 * e1;
      if (D == 3) {
        float e2 = n[2] - mj[2];
        s2 = s2 + e2 * e2;
      }
      float s = s2 / DD;
      s = s * 1000000.0f;
      unsigned u = f2u_rn(s);
      if (u < best) best = u;
    }
    if (q) q[i] = best;
    f[i] = ((float)best) / 1000000.0f;
  }
  float sum = dor_tree_sum_f32(N, f);
  free(m);
  free(f);
  return sum / (float)N;
}

/* ------------------------------------------------------------------ */
/* exact k-NN with lowest-index tie-break                             */
/* ------------------------------------------------------------------ */
void dor_knn(int D, int nq, const float *Q, int np, const float *P, int k,
             int self_exclude, int *idx, float *d2) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nq; ++i) {
    float bd[64];
    int bi[64];
    for (int s = 0; s < k; ++s) {
      bd[s] = INFINITY;
      bi[s] = -1;
    }
    const float *q = Q + (size_t)i * D;
    for (int j = 0; j < np; ++j) {
      if (self_exclude && j == i) continue;
      const float *p = P + (size_t)j * D;
      float dx = q[0] - p[0], dy = q[1] - p[1];
      float dd;
      if (D == 3) {
        float dz = q[2] - p[2];
        dd = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
      } else {
        dd = fmaf(dy, dy, dx * dx);
      }
      /* candidates arrive in increasing j: (dd, j) < (bd, bi) iff dd < bd */
      if (!(dd < bd[k - 1])) continue;
      int s = k - 1;
      while (s > 0 && dd < bd[s - 1]) {
        bd[s] = bd[s - 1];
        bi[s] = bi[s - 1];
        --s;
      }
      bd[s] = dd;
      bi[s] = j;
    }
    for (int s = 0; s < k; ++s) {
      idx[(size_t)i * k + s] = bi[s];
      d2[(size_t)i * k + s] = bd[s];
    }
  }
}

/* Per-point sigma from the k nearest neighbours within one cloud: the
 * lists of dor_knn (self excluded), then the mean distance nearest first. */
/* The k <= 64 nearest other points of point `self` of P, in (distance,
 * index) order (as dor_knn ranks them); missing ones: +inf, -1. */
static void knn_self(int D, int self, int np, const float *P, int k, float *bd, int *bi) {
  for (int s = 0; s < k; ++s) {
    bd[s] = INFINITY;
    bi[s] = -1;
  }
  const float *q = P + (size_t)self * D;
  for (int j = 0; j < np; ++j) {
    if (j == self) continue;
    const float *p = P + (size_t)j * D;
    float dx = q[0] - p[0], dy = q[1] - p[1];
    float dd;
    if (D == 3) {
      float dz = q[2] - p[2];
      dd = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    } else {
      dd = fmaf(dy, dy, dx * dx);
    }
    if (!(dd < bd[k - 1])) continue;
    int s = k - 1;
    while (s > 0 && dd < bd[s - 1]) {
      bd[s] = bd[s - 1];
      bi[s] = bi[s - 1];
      --s;
    }
    bd[s] = dd;
    bi[s] = j;
  }
}

void dor_knn_sigma(int D, int nq, const int *qidx, int np, const float *P,
                   int k, float *sigma) {
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < nq; ++i) {
    float bd[64];
    int bi[64];
    knn_self(D, qidx ? qidx[i] : i, np, P, k, bd, bi);
    float sum = sqrtf(bd[0]);
    for (int s = 1; s < k; ++s) sum = sum + sqrtf(bd[s]);
    sigma[i] = sum / (float)k;
  }
}

void dor_knn_cov(int D, int nq, const int *qidx, int np, const float *P, int k,
                 float scale, float floor, float *sigma, float *cov) {
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < nq; ++i) {
    float bd[64];
    int bi[64];
    knn_self(D, qidx ? qidx[i] : i, np, P, k, bd, bi);
    float sum = sqrtf(bd[0]);
    for (int s = 1; s < k; ++s) sum = sum + sqrtf(bd[s]);
    const float sg = sum / (float)k;
    sigma[i] = sg;
    float *o = cov + (size_t)i * D * D;
    if (bi[k - 1] < 0) {
      for (int e = 0; e < D * D; ++e) o[e] = NAN;
      continue;
    }
    float mu[3] = {0, 0, 0};
    for (int a = 0; a < D; ++a) {
      float m = P[(size_t)bi[0] * D + a];
      for (int s = 1; s < k; ++s) m = m + P[(size_t)bi[s] * D + a];
      mu[a] = m / (float)k;
    }
    const float fl = floor * (sg * sg);
    for (int a = 0; a < D; ++a)
      for (int b = 0; b < D; ++b) {
        float c = 0.0f;
        for (int s = 0; s < k; ++s) {
          const float u = P[(size_t)bi[s] * D + a] - mu[a], v = P[(size_t)bi[s] * D + b] - mu[b];
          c = c + u * v;
        }
        float cab = scale * (c / (float)k);
        if (a == b) cab = cab + fl;
        o[a + D * b] = cab;
      }
  }
}

/* ------------------------------------------------------------------ */
/* convolved-uncertainty energy and its adjoint (FP64)                 */
/* ------------------------------------------------------------------ */
static const double kInvSqrt2Pi = 0.39894228040143267794; /* (2 pi)^-1/2 */

/* Validity rule shared with the library's init: finite coordinates, finite
 * covariance with positive leading principal minors. */
static int valid_point(int D, const double *p, const double *S) {
  for (int a = 0; a < D; ++a)
    if (!isfinite(p[a])) return 0;
  for (int a = 0; a < D * D; ++a)
    if (!isfinite(S[a])) return 0;
  if (!(S[0] > 0)) return 0;
  double m2 = S[0] * S[D + 1] - S[1] * S[D];
  if (!(m2 > 0)) return 0;
  if (D == 3) {
    double det = S[0] * (S[4] * S[8] - S[7] * S[5]) -
                 S[3] * (S[1] * S[8] - S[7] * S[2]) +
                 S[6] * (S[1] * S[5] - S[4] * S[2]);
    if (!(det > 0)) return 0;
  }
  return 1;
}

/* One pair: S = Sa + Sb, d = ya - xb.  Returns e = N(d; 0, S) and fills
 * w = S^-1 d and Sinv. */
static double pair_conv(int D, const double *ya, const double *Sa,
                        const double *xb, const double *Sb, double *w,
                        double *Sinv) {
  double S[9], d[3];
  for (int a = 0; a < D * D; ++a) S[a] = Sa[a] + Sb[a];
  for (int a = 0; a < D; ++a) d[a] = ya[a] - xb[a];
  double det;
  if (D == 3) {
    /* symmetrise so the inverse is exactly symmetric */
    double s11 = S[0], s22 = S[4], s33 = S[8];
    double s12 = 0.5 * (S[3] + S[1]), s13 = 0.5 * (S[6] + S[2]),
           s23 = 0.5 * (S[7] + S[5]);
    double c11 = s22 * s33 - s23 * s23, c12 = s13 * s23 - s12 * s33,
           c13 = s12 * s23 - s13 * s22, c22 = s11 * s33 - s13 * s13,
           c23 = s12 * s13 - s11 * s23, c33 = s11 * s22 - s12 * s12;
    det = s11 * c11 + s12 * c12 + s13 * c13;
    double id = 1.0 / det;
    Sinv[0] = c11 * id; Sinv[1] = c12 * id; Sinv[2] = c13 * id;
    Sinv[3] = c12 * id; Sinv[4] = c22 * id; Sinv[5] = c23 * id;
    Sinv[6] = c13 * id; Sinv[7] = c23 * id; Sinv[8] = c33 * id;
  } else {
    double s11 = S[0], s22 = S[3], s12 = 0.5 * (S[1] + S[2]);
    det = s11 * s22 - s12 * s12;
    double id = 1.0 / det;
    Sinv[0] = s22 * id; Sinv[1] = -s12 * id; Sinv[2] = -s12 * id;
    Sinv[3] = s11 * id;
  }
  double q = 0;
  for (int a = 0; a < D; ++a) {
    double s = 0;
    for (int b = 0; b < D; ++b) s += Sinv[a + D * b] * d[b];
    w[a] = s;
    q += d[a] * s;
  }
  double norm = (D == 3) ? kInvSqrt2Pi * kInvSqrt2Pi * kInvSqrt2Pi
                         : kInvSqrt2Pi * kInvSqrt2Pi;
  return norm / sqrt(det) * exp(-0.5 * q);
}

static void transform_sources(int D, int M, const double *X, const double *SX,
                              const double *R, const double *t, double *xp,
                              double *Sxp) {
  for (int j = 0; j < M; ++j) {
    const double *x = X + (size_t)j * D;
    const double *S = SX + (size_t)j * D * D;
    for (int a = 0; a < D; ++a) {
      double s = t[a];
      for (int b = 0; b < D; ++b) s += R[a + D * b] * x[b];
      xp[(size_t)j * D + a] = s;
    }
    /* R S R^T */
    double RS[9];
    for (int a = 0; a < D; ++a)
      for (int c = 0; c < D; ++c) {
        double s = 0;
        for (int b = 0; b < D; ++b) s += R[a + D * b] * S[b + D * c];
        RS[a + D * c] = s;
      }
    for (int a = 0; a < D; ++a)
      for (int c = 0; c < D; ++c) {
        double s = 0;
        for (int b = 0; b < D; ++b) s += RS[a + D * b] * R[c + D * b];
        Sxp[(size_t)j * D * D + a + D * c] = s;
      }
  }
}

double dor_energy(int D, int N, const double *Y, const double *SY, int M,
                  const double *X, const double *SX, const double *R,
                  const double *t, double *gpose, double *gY, double *gSY,
                  double *gX, double *gSX, double *eT) {
  const int DD = D * D;
  double *xp = (double *)malloc(sizeof(double) * ((size_t)M * D + 1));
  double *Sxp = (double *)malloc(sizeof(double) * ((size_t)M * DD + 1));
  char *vx = (char *)malloc((size_t)M + 1);
  char *vy = (char *)malloc((size_t)N + 1);
  transform_sources(D, M, X, SX, R, t, xp, Sxp);
  for (int j = 0; j < M; ++j)
    vx[j] = (char)valid_point(D, X + (size_t)j * D, SX + (size_t)j * DD);
  for (int i = 0; i < N; ++i)
    vy[i] = (char)valid_point(D, Y + (size_t)i * D, SY + (size_t)i * DD);

  double E = 0;
  /* target pass: energy, dE/dy_i, dE/dSy_i */
#pragma omp parallel for schedule(static) reduction(+ : E)
  for (int i = 0; i < N; ++i) {
    double ei = 0, ga[3] = {0, 0, 0}, gS[9] = {0};
    if (vy[i]) {
      for (int j = 0; j < M; ++j) {
        if (!vx[j]) continue;
        double w[3], Si[9];
        double e = pair_conv(D, Y + (size_t)i * D, SY + (size_t)i * DD,
                             xp + (size_t)j * D, Sxp + (size_t)j * DD, w, Si);
        ei += e;
        for (int a = 0; a < D; ++a) ga[a] -= e * w[a];
        for (int a = 0; a < D; ++a)
          for (int b = 0; b < D; ++b)
            gS[a + D * b] += 0.5 * e * (w[a] * w[b] - Si[a + D * b]);
      }
    }
    E += ei;
    if (eT) eT[i] = ei;
    if (gY)
      for (int a = 0; a < D; ++a) gY[(size_t)i * D + a] = ga[a];
    if (gSY)
      for (int a = 0; a < DD; ++a) gSY[(size_t)i * DD + a] = gS[a];
  }

  if (gpose || gX || gSX) {
    double *gxp = (double *)calloc((size_t)M * D + 1, sizeof(double));
    double *gSxp = (double *)calloc((size_t)M * DD + 1, sizeof(double));
    /* source pass: dE/dx'_j, dE/dSx'_j */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < M; ++j) {
      if (!vx[j]) continue;
      double ga[3] = {0, 0, 0}, gS[9] = {0};
      for (int i = 0; i < N; ++i) {
        if (!vy[i]) continue;
        double w[3], Si[9];
        double e = pair_conv(D, Y + (size_t)i * D, SY + (size_t)i * DD,
                             xp + (size_t)j * D, Sxp + (size_t)j * DD, w, Si);
        for (int a = 0; a < D; ++a) ga[a] += e * w[a];
        for (int a = 0; a < D; ++a)
          for (int b = 0; b < D; ++b)
            gS[a + D * b] += 0.5 * e * (w[a] * w[b] - Si[a + D * b]);
      }
      for (int a = 0; a < D; ++a) gxp[(size_t)j * D + a] = ga[a];
      for (int a = 0; a < DD; ++a) gSxp[(size_t)j * DD + a] = gS[a];
    }
    if (gpose) {
      double gw[3] = {0, 0, 0}, gt[3] = {0, 0, 0}, gth = 0;
      for (int j = 0; j < M; ++j) {
        const double *g = gxp + (size_t)j * D;
        const double *H = gSxp + (size_t)j * DD;
        const double *Sp = Sxp + (size_t)j * DD;
        double p[3];
        for (int a = 0; a < D; ++a) {
          p[a] = xp[(size_t)j * D + a] - t[a];
          gt[a] += g[a];
        }
        if (D == 3) {
          gw[0] += p[1] * g[2] - p[2] * g[1];
          gw[1] += p[2] * g[0] - p[0] * g[2];
          gw[2] += p[0] * g[1] - p[1] * g[0];
          /* K = Sp H - H Sp ; gw -= 2 vee(K) */
          double K[9];
          for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 3; ++c) {
              double s = 0;
              for (int b = 0; b < 3; ++b)
                s += Sp[a + 3 * b] * H[b + 3 * c] - H[a + 3 * b] * Sp[b + 3 * c];
              K[a + 3 * c] = s;
            }
          gw[0] -= 2.0 * K[2 + 3 * 1];
          gw[1] -= 2.0 * K[0 + 3 * 2];
          gw[2] -= 2.0 * K[1 + 3 * 0];
        } else {
          /* J = [[0,-1],[1,0]]: dE = g^T J p + tr(H (J Sp - Sp J)) */
          gth += g[0] * (-p[1]) + g[1] * p[0];
          double JS[4], SJ[4];
          /* (J Sp)_ac = sum_b J_ab Sp_bc */
          JS[0] = -Sp[1]; JS[2] = -Sp[3];
          JS[1] = Sp[0];  JS[3] = Sp[2];
          /* (Sp J)_ac = sum_b Sp_ab J_bc */
          SJ[0] = Sp[2];  SJ[2] = -Sp[0];
          SJ[1] = Sp[3];  SJ[3] = -Sp[1];
          for (int a = 0; a < 2; ++a)
            for (int c = 0; c < 2; ++c)
              gth += H[c + 2 * a] * (JS[a + 2 * c] - SJ[a + 2 * c]);
        }
      }
      if (D == 3) {
        for (int a = 0; a < 3; ++a) {
          gpose[a] = gw[a];
          gpose[3 + a] = gt[a];
        }
      } else {
        gpose[0] = gth;
        gpose[1] = gt[0];
        gpose[2] = gt[1];
      }
    }
    if (gX || gSX) {
      for (int j = 0; j < M; ++j) {
        const double *g = gxp + (size_t)j * D;
        const double *H = gSxp + (size_t)j * DD;
        if (gX)
          for (int b = 0; b < D; ++b) {
            double s = 0;
            for (int a = 0; a < D; ++a) s += R[a + D * b] * g[a];
            gX[(size_t)j * D + b] = s;
          }
        if (gSX) {
          /* R^T H R */
          for (int b = 0; b < D; ++b)
            for (int c = 0; c < D; ++c) {
              double s = 0;
              for (int a = 0; a < D; ++a)
                for (int e = 0; e < D; ++e)
                  s += R[a + D * b] * H[a + D * e] * R[e + D * c];
              gSX[(size_t)j * DD + b + D * c] = s;
            }
        }
      }
    }
    free(gxp);
    free(gSxp);
  }
  free(xp);
  free(Sxp);
  free(vx);
  free(vy);
  return E;
}

/* Rodrigues: exp([w]x) */
static void rot_exp(const double *w, double *Rm /* column-major */) {
  double th = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  double A, B;
  if (th < 1e-12) {
    A = 1.0;
    B = 0.5;
  } else {
    A = sin(th) / th;
    B = (1.0 - cos(th)) / (th * th);
  }
  double K[9] = {0, w[2], -w[1], -w[2], 0, w[0], w[1], -w[0], 0};
  for (int a = 0; a < 3; ++a)
    for (int c = 0; c < 3; ++c) {
      double k2 = 0;
      for (int b = 0; b < 3; ++b) k2 += K[a + 3 * b] * K[b + 3 * c];
      Rm[a + 3 * c] = (a == c ? 1.0 : 0.0) + A * K[a + 3 * c] + B * k2;
    }
}

static double energy_at(int D, int N, const double *Y, const double *SY, int M,
                        const double *X, const double *SX, const double *R0,
                        const double *t0, const double *p) {
  double R[9], t[3];
  if (D == 3) {
    double Rw[9];
    rot_exp(p, Rw);
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < 3; ++c) {
        double s = 0;
        for (int b = 0; b < 3; ++b) s += Rw[a + 3 * b] * R0[b + 3 * c];
        R[a + 3 * c] = s;
      }
    for (int a = 0; a < 3; ++a) t[a] = t0[a] + p[3 + a];
  } else {
    double c = cos(p[0]), s = sin(p[0]);
    double Rt[4] = {c, s, -s, c};
    for (int a = 0; a < 2; ++a)
      for (int cc = 0; cc < 2; ++cc)
        R[a + 2 * cc] = Rt[a] * R0[2 * cc] + Rt[a + 2] * R0[1 + 2 * cc];
    t[0] = t0[0] + p[1];
    t[1] = t0[1] + p[2];
  }
  return dor_energy(D, N, Y, SY, M, X, SX, R, t, NULL, NULL, NULL, NULL, NULL,
                    NULL);
}

double dor_energy_fd_pose(int D, int N, const double *Y, const double *SY,
                          int M, const double *X, const double *SX,
                          const double *R, const double *t, double h,
                          double *gpose) {
  int np = (D == 3) ? 6 : 3;
  double p[6] = {0, 0, 0, 0, 0, 0};
  double E0 = energy_at(D, N, Y, SY, M, X, SX, R, t, p);
  for (int k = 0; k < np; ++k) {
    p[k] = h;
    double ep = energy_at(D, N, Y, SY, M, X, SX, R, t, p);
    p[k] = -h;
    double em = energy_at(D, N, Y, SY, M, X, SX, R, t, p);
    p[k] = 0;
    gpose[k] = (ep - em) / (2 * h);
  }
  return E0;
}

/* ------------------------------------------------------------------ */
/* compi: surrogate polynomial coefficients                            */
/* ------------------------------------------------------------------ */
/* Monomials of degree <= 4 are keyed by their sorted variable list padded
 * with the sentinel NV: code = ((v0*B + v1)*B + v2)*B + v3, B = NV + 1. */
typedef struct {
  int n;
  unsigned char deg[32];
  unsigned char v[32][4];
  double c[32];
} spoly;

static void sp_add_term(spoly *p, int deg, const unsigned char *v, double c) {
  for (int k = 0; k < p->n; ++k) {
    if (p->deg[k] != deg) continue;
    int same = 1;
    for (int q = 0; q < deg; ++q)
      if (p->v[k][q] != v[q]) same = 0;
    if (same) {
      p->c[k] += c;
      return;
    }
  }
  int k = p->n++;
  p->deg[k] = (unsigned char)deg;
  for (int q = 0; q < 4; ++q) p->v[k][q] = q < deg ? v[q] : 0;
  p->c[k] = c;
}

static void sort_small(unsigned char *v, int n) {
  for (int a = 1; a < n; ++a) {
    unsigned char x = v[a];
    int b = a - 1;
    while (b >= 0 && v[b] > x) {
      v[b + 1] = v[b];
      --b;
    }
    v[b + 1] = x;
  }
}

static int mono_code(int nv, int deg, const unsigned char *v) {
  int B = nv + 1, code = 0;
  for (int q = 0; q < 4; ++q) code = code * B + (q < deg ? v[q] : nv);
  return code;
}

/* dense[code] += s * a * b */
static void acc_product(double *dense, int nv, double s, const spoly *a,
                        const spoly *b) {
  for (int x = 0; x < a->n; ++x)
    for (int y = 0; y < b->n; ++y) {
      unsigned char v[4];
      int deg = a->deg[x] + b->deg[y];
      for (int q = 0; q < a->deg[x]; ++q) v[q] = a->v[x][q];
      for (int q = 0; q < b->deg[y]; ++q) v[a->deg[x] + q] = b->v[y][q];
      sort_small(v, deg);
      dense[mono_code(nv, deg, v)] += s * a->c[x] * b->c[y];
    }
}

/* Device state of compi.cu after its shared-memory staging: scaled inverses
 * and scaled g factors (compi.cu:642-677 for 3-D, 106-148 for 2-D). */
typedef struct {
  int D;
  float Isig;
  float gsc; /* multiplier of g */
} compi_scale;

static compi_scale compi_scaling(int D, double sig) {
  compi_scale s;
  s.D = D;
  double is = 1.0 / sig;
  float f = (float)is;
  if (D == 2) {
    /* __double2float_rd (compi.cu:110) */
    if ((double)f > is) f = nextafterf(f, -INFINITY);
    s.Isig = f;
    s.gsc = f;
  } else {
    /* __double2float_rn (compi.cu:646); sqrtf(Isig^3) (compi.cu:662) */
    s.Isig = f;
    s.gsc = sqrtf(f * f * f);
  }
  return s;
}

/* Per pair: P (compi.cu:747-772 / 193-211), m' = fl32(R m + t) and
 * A' = fl32(C) (compi.cu:775-791 / 213-228). */
static double pair_state(int D, const float *n, const float *bs, float gxs,
                         const float *m, const float *as, float gys,
                         const double *R, const double *t, float *mp,
                         float *Ap) {
  double d[3], C[9];
  if (D == 3) {
    double r[3][3];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) r[a][b] = R[a + 3 * b];
    /* a_lk = as[(l) + 3(k)] (column-major) */
    for (int a = 0; a < 3; ++a)
      d[a] = (double)n[a] - t[a] - (double)m[0] * r[a][0] -
             (double)m[1] * r[a][1] - (double)m[2] * r[a][2];
    for (int p = 0; p < 3; ++p)
      for (int q = 0; q < 3; ++q) {
        double s = 0;
        for (int k = 0; k < 3; ++k) {
          double inner = 0;
          for (int l = 0; l < 3; ++l) inner += (double)as[l + 3 * k] * r[q][l];
          s += r[p][k] * inner;
        }
        C[p + 3 * q] = s; /* c_pq */
      }
  } else {
    double r11 = R[0], r21 = R[1], r12 = R[2], r22 = R[3];
    double a11 = as[0], a21 = as[1], a12 = as[2], a22 = as[3];
    d[0] = (double)n[0] - t[0] - (double)m[0] * r11 - (double)m[1] * r12;
    d[1] = (double)n[1] - t[1] - (double)m[0] * r21 - (double)m[1] * r22;
    C[0] = a11 * r11 * r11 + a22 * r12 * r12 + a12 * r11 * r12 + a21 * r11 * r12;
    C[2] = a11 * r11 * r21 + a12 * r11 * r22 + a21 * r12 * r21 + a22 * r12 * r22;
    C[1] = a11 * r11 * r21 + a12 * r12 * r21 + a21 * r11 * r22 + a22 * r12 * r22;
    C[3] = a11 * r21 * r21 + a22 * r22 * r22 + a12 * r21 * r22 + a21 * r21 * r22;
  }
  double mal1 = 0, mal2 = 0;
  for (int p = 0; p < D; ++p)
    for (int q = 0; q < D; ++q) {
      mal1 += d[p] * C[p + D * q] * d[q];
      mal2 += d[p] * (double)bs[p + D * q] * d[q];
    }
  float gg = gxs * gys;
  double P = (double)gg * (exp(-0.5 * mal1) + exp(-0.5 * mal2));
  for (int a = 0; a < D; ++a) {
    double s = R[a] * (double)m[0];
    for (int b = 1; b < D; ++b) s = s + R[a + D * b] * (double)m[b];
    s = s + t[a];
    mp[a] = (float)s;
  }
  for (int a = 0; a < D * D; ++a) Ap[a] = (float)C[a];
  return P;
}

static void scaled_state(int D, const float *inv, float g, compi_scale sc,
                         float *out, float *gout) {
  for (int a = 0; a < D * D; ++a) out[a] = inv[a] * sc.Isig;
  *gout = g * sc.gsc;
}

double dor_pair_weight(int D, const float *n, const float *binv, float gx,
                       const float *m, const float *ainv, float gy,
                       const double *R, const double *t, double sig) {
  compi_scale sc = compi_scaling(D, sig);
  float bs[9], as[9], gxs, gys, mp[3], Ap[9];
  scaled_state(D, binv, gx, sc, bs, &gxs);
  scaled_state(D, ainv, gy, sc, as, &gys);
  return pair_state(D, n, bs, gxs, m, as, gys, R, t, mp, Ap);
}

/* dR entry (a,b) as a signed single variable.  3-D: variable 3 + 3a + b.
 * 2-D (DUGMA.m:67): [[r11, -r21], [r21, r11]], r11 = var 2, r21 = var 3. */
static void dR_entry(int D, int a, int b, int *var, double *sgn) {
  if (D == 3) {
    *var = 3 + 3 * a + b;
    *sgn = 1.0;
  } else {
    static const int vv[2][2] = {{2, 3}, {3, 2}};
    static const double ss[2][2] = {{1.0, -1.0}, {1.0, 1.0}};
    *var = vv[a][b];
    *sgn = ss[a][b];
  }
}

/* Output vector in energy3D.m:15 / energy2D.m:7 order from the dense
 * monomial table; returns the number of monomials outside that list whose
 * value exceeds 1e-12 of the largest listed one (terms that cancel
 * mathematically leave rounding residues when summed in another order). */
static int gather_coef(int D, const double *dense, double *coef) {
  const int nv = (D == 3) ? 12 : 4, B = nv + 1, ncode = B * B * B * B;
  int nm = (D == 3) ? DUGMA_NMONO3 : DUGMA_NMONO2;
  char *used = (char *)calloc((size_t)ncode, 1);
  for (int k = 0; k < nm; ++k) {
    const unsigned char *e = (D == 3) ? DUGMA_MONO3[k] : DUGMA_MONO2[k];
    unsigned char v[4];
    int deg = 0;
    for (int x = 0; x < nv; ++x)
      for (int r = 0; r < e[x]; ++r) v[deg++] = (unsigned char)x;
    int code = mono_code(nv, deg, v);
    coef[k] = dense[code];
    used[code] = 1;
  }
  double big = 0.0;
  for (int k = 0; k < nm; ++k) big = fmax(big, fabs(coef[k]));
  int missing = 0;
  for (int c = 0; c < ncode; ++c)
    if (!used[c] && fabs(dense[c]) > 1e-12 * big) ++missing;
  free(used);
  return missing;
}

int dor_compi(int D, int N, const float *Xf, const float *XCinv,
              const float *gX, int M, const float *Yf, const float *YCinv,
              const float *gY, const double *R, const double *t, double sig,
              double *coef) {
  const int nv = (D == 3) ? 12 : 4;
  const int B = nv + 1;
  const int ncode = B * B * B * B;
  const int DD = D * D;
  compi_scale sc = compi_scaling(D, sig);
  double *dense = (double *)calloc((size_t)ncode, sizeof(double));
  for (int i = 0; i < N; ++i) {
    float bs[9], gxs;
    scaled_state(D, XCinv + (size_t)i * DD, gX[i], sc, bs, &gxs);
    const float *n = Xf + (size_t)i * D;
    for (int j = 0; j < M; ++j) {
      float as[9], gys, mp[3], Ap[9];
      scaled_state(D, YCinv + (size_t)j * DD, gY[j], sc, as, &gys);
      double P = pair_state(D, n, bs, gxs, Yf + (size_t)j * D, as, gys, R, t,
                            mp, Ap);
      if (P == 0.0) continue;
      /* d_a = n_a - t_a - sum_b dR_ab m'_b */
      spoly d[3], u[3];
      for (int a = 0; a < D; ++a) {
        d[a].n = 0;
        unsigned char v[4];
        sp_add_term(&d[a], 0, v, (double)n[a]);
        v[0] = (unsigned char)a;
        sp_add_term(&d[a], 1, v, -1.0);
        for (int b = 0; b < D; ++b) {
          int var;
          double s;
          dR_entry(D, a, b, &var, &s);
          v[0] = (unsigned char)var;
          sp_add_term(&d[a], 1, v, -s * (double)mp[b]);
        }
      }
      /* u_c = sum_a dR_ac d_a */
      for (int c = 0; c < D; ++c) {
        u[c].n = 0;
        for (int a = 0; a < D; ++a) {
          int var;
          double s;
          dR_entry(D, a, c, &var, &s);
          for (int k = 0; k < d[a].n; ++k) {
            unsigned char v[4];
            int deg = d[a].deg[k];
            for (int q = 0; q < deg; ++q) v[q] = d[a].v[k][q];
            v[deg] = (unsigned char)var;
            sort_small(v, deg + 1);
            sp_add_term(&u[c], deg + 1, v, s * d[a].c[k]);
          }
        }
      }
      for (int p = 0; p < D; ++p)
        for (int q = 0; q < D; ++q) {
          acc_product(dense, nv, P * (double)Ap[p + D * q], &u[p], &u[q]);
          acc_product(dense, nv, P * (double)bs[p + D * q], &d[p], &d[q]);
        }
    }
  }
  int missing = gather_coef(D, dense, coef);
  free(dense);
  return missing;
}

/* ------------------------------------------------------------------ */
/* compi in moment form (FP64): per source j the P-weighted target     */
/* moments, then the polynomial of each source                          */
/* ------------------------------------------------------------------ */
static void sp_clear(spoly *p) { p->n = 0; }

/* dst += s * src */
static void sp_axpy(spoly *dst, double s, const spoly *src) {
  for (int k = 0; k < src->n; ++k) sp_add_term(dst, src->deg[k], src->v[k], s * src->c[k]);
}

/* out = a * b */
static void sp_mul(spoly *out, const spoly *a, const spoly *b) {
  sp_clear(out);
  for (int x = 0; x < a->n; ++x)
    for (int y = 0; y < b->n; ++y) {
      unsigned char v[4];
      int deg = a->deg[x] + b->deg[y];
      for (int q = 0; q < a->deg[x]; ++q) v[q] = a->v[x][q];
      for (int q = 0; q < b->deg[y]; ++q) v[a->deg[x] + q] = b->v[y][q];
      sort_small(v, deg);
      sp_add_term(out, deg, v, a->c[x] * b->c[y]);
    }
}

/* dense[code] += s * a */
static void dense_add(double *dense, int nv, double s, const spoly *a) {
  for (int x = 0; x < a->n; ++x) dense[mono_code(nv, a->deg[x], a->v[x])] += s * a->c[x];
}

/* Moments of source j (FP64): S0 = sum_i P, S1 = sum P n, S2 = sum P n n^T,
 * SB = sum P B, SnB_b = sum P n_a B_ab, SBn_a = sum P B_ab n_b,
 * SnBn = sum P n^T B n; plus m'_j and A'_j.  Stride kMomStride doubles. */
enum { kMomStride = 1 + 3 + 9 + 9 + 3 + 3 + 1 + 3 + 9 };

int dor_compi_moments(int D, int N, const float *Xf, const float *XCinv,
                      const float *gX, int M, const float *Yf,
                      const float *YCinv, const float *gY, const double *R,
                      const double *t, double sig, double *coef) {
  const int DD = D * D;
  const int nv = (D == 3) ? 12 : 4;
  const int B = nv + 1, ncode = B * B * B * B;
  compi_scale sc = compi_scaling(D, sig);
  double *mom = (double *)calloc((size_t)M * kMomStride + 1, sizeof(double));
  float *bsa = (float *)malloc(sizeof(float) * ((size_t)N * DD + 1));
  float *gxa = (float *)malloc(sizeof(float) * ((size_t)N + 1));
  for (int i = 0; i < N; ++i)
    scaled_state(D, XCinv + (size_t)i * DD, gX[i], sc, bsa + (size_t)i * DD, gxa + i);
#pragma omp parallel for schedule(dynamic, 8)
  for (int j = 0; j < M; ++j) {
    float as[9], gys, mp[3], Ap[9];
    scaled_state(D, YCinv + (size_t)j * DD, gY[j], sc, as, &gys);
    double S0 = 0, S1[3] = {0}, S2[9] = {0}, SB[9] = {0}, SnB[3] = {0}, SBn[3] = {0}, SnBn = 0;
    for (int i = 0; i < N; ++i) {
      const float *n = Xf + (size_t)i * D;
      const float *bs = bsa + (size_t)i * DD;
      double P = pair_state(D, n, bs, gxa[i], Yf + (size_t)j * D, as, gys, R, t, mp, Ap);
      if (P == 0.0) continue;
      double nd[3];
      for (int a = 0; a < D; ++a) nd[a] = (double)n[a];
      S0 += P;
      for (int a = 0; a < D; ++a) {
        S1[a] += P * nd[a];
        for (int b = 0; b < D; ++b) {
          const double Bab = (double)bs[a + D * b];
          S2[a + D * b] += P * nd[a] * nd[b];
          SB[a + D * b] += P * Bab;
          SnB[b] += P * nd[a] * Bab;
          SBn[a] += P * Bab * nd[b];
          SnBn += P * nd[a] * Bab * nd[b];
        }
      }
    }
    /* m'_j and A'_j do not depend on i: the values of the last pair (or of
     * a pair evaluated here when every P was zero) */
    pair_state(D, Xf, bsa, gxa[0], Yf + (size_t)j * D, as, gys, R, t, mp, Ap);
    double *o = mom + (size_t)j * kMomStride;
    o[0] = S0;
    for (int a = 0; a < 3; ++a) {
      o[1 + a] = S1[a];
      o[22 + a] = SnB[a];
      o[25 + a] = SBn[a];
      o[29 + a] = a < D ? (double)mp[a] : 0.0;
    }
    for (int a = 0; a < 9; ++a) {
      o[4 + a] = S2[a];
      o[13 + a] = SB[a];
      o[32 + a] = a < DD ? (double)Ap[a] : 0.0;
    }
    o[28] = SnBn;
  }
  free(bsa);
  free(gxa);
  /* F_j(dR, dt) = SnBn - SnB.v - SBn.v + v^T SB v
   *             + sum_pq K_pq (S2_pq - S1_p v_q - S1_q v_p + S0 v_p v_q),
   * v = dR m'_j + dt, K = dR A'_j dR^T; summed over j in order */
  double *dense = (double *)calloc((size_t)ncode, sizeof(double));
  for (int j = 0; j < M; ++j) {
    const double *o = mom + (size_t)j * kMomStride;
    const double S0 = o[0], *S1 = o + 1, *S2 = o + 4, *SB = o + 13, *SnB = o + 22, *SBn = o + 25,
                 SnBn = o[28], *mp = o + 29, *Ap = o + 32;
    if (S0 == 0.0) continue;
    spoly v[3], vv[3][3], K[3][3], tmp, w;
    unsigned char var1[4];
    for (int a = 0; a < D; ++a) {
      sp_clear(&v[a]);
      var1[0] = (unsigned char)a;
      sp_add_term(&v[a], 1, var1, 1.0); /* dt_a */
      for (int b = 0; b < D; ++b) {
        int var;
        double sg;
        dR_entry(D, a, b, &var, &sg);
        var1[0] = (unsigned char)var;
        sp_add_term(&v[a], 1, var1, sg * mp[b]);
      }
    }
    for (int p = 0; p < D; ++p)
      for (int q = 0; q < D; ++q) {
        sp_mul(&vv[p][q], &v[p], &v[q]);
        sp_clear(&K[p][q]);
        for (int k = 0; k < D; ++k)
          for (int l = 0; l < D; ++l) {
            int va, vb;
            double sa, sb2;
            dR_entry(D, p, k, &va, &sa);
            dR_entry(D, q, l, &vb, &sb2);
            unsigned char v2[4] = {(unsigned char)(va < vb ? va : vb), (unsigned char)(va < vb ? vb : va), 0, 0};
            sp_add_term(&K[p][q], 2, v2, sa * sb2 * Ap[k + D * l]);
          }
      }
    unsigned char none[4] = {0, 0, 0, 0};
    sp_clear(&tmp);
    sp_add_term(&tmp, 0, none, SnBn);
    for (int a = 0; a < D; ++a) sp_axpy(&tmp, -(SnB[a] + SBn[a]), &v[a]);
    dense_add(dense, nv, 1.0, &tmp);
    for (int a = 0; a < D; ++a)
      for (int b = 0; b < D; ++b) dense_add(dense, nv, SB[a + D * b], &vv[a][b]);
    for (int p = 0; p < D; ++p)
      for (int q = 0; q < D; ++q) {
        sp_clear(&w);
        sp_add_term(&w, 0, none, S2[p + D * q]);
        sp_axpy(&w, -S1[p], &v[q]);
        sp_axpy(&w, -S1[q], &v[p]);
        sp_axpy(&w, S0, &vv[p][q]);
        acc_product(dense, nv, 1.0, &K[p][q], &w);
      }
  }
  free(mom);
  int missing = gather_coef(D, dense, coef);
  free(dense);
  return missing;
}

double dor_surrogate_eval(int D, int N, const float *Xf, const float *XCinv,
                          const float *gX, int M, const float *Yf,
                          const float *YCinv, const float *gY, const double *R,
                          const double *t, double sig, const double *dT) {
  const int DD = D * D;
  compi_scale sc = compi_scaling(D, sig);
  double dR[9], dt[3];
  if (D == 3) {
    for (int a = 0; a < 3; ++a) dt[a] = dT[a];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) dR[a + 3 * b] = dT[3 + 3 * a + b];
  } else {
    dt[0] = dT[0];
    dt[1] = dT[1];
    dR[0] = dT[2];
    dR[1] = dT[3];
    dR[2] = -dT[3];
    dR[3] = dT[2];
  }
  double F = 0;
  for (int i = 0; i < N; ++i) {
    float bs[9], gxs;
    scaled_state(D, XCinv + (size_t)i * DD, gX[i], sc, bs, &gxs);
    const float *n = Xf + (size_t)i * D;
    for (int j = 0; j < M; ++j) {
      float as[9], gys, mp[3], Ap[9];
      scaled_state(D, YCinv + (size_t)j * DD, gY[j], sc, as, &gys);
      double P = pair_state(D, n, bs, gxs, Yf + (size_t)j * D, as, gys, R, t,
                            mp, Ap);
      double d[3], u[3];
      for (int a = 0; a < D; ++a) {
        double s = (double)n[a] - dt[a];
        for (int b = 0; b < D; ++b) s -= dR[a + D * b] * (double)mp[b];
        d[a] = s;
      }
      for (int c = 0; c < D; ++c) {
        double s = 0;
        for (int a = 0; a < D; ++a) s += dR[a + D * c] * d[a];
        u[c] = s;
      }
      double val = 0;
      for (int p = 0; p < D; ++p)
        for (int q = 0; q < D; ++q)
          val += u[p] * (double)Ap[p + D * q] * u[q] +
                 d[p] * (double)bs[p + D * q] * d[q];
      F += P * val;
    }
  }
  return F;
}

double dor_poly_eval(int D, const double *coef, const double *dT) {
  int nm = (D == 3) ? DUGMA_NMONO3 : DUGMA_NMONO2;
  int nv = (D == 3) ? 12 : 4;
  double F = 0;
  for (int k = 0; k < nm; ++k) {
    const unsigned char *e = (D == 3) ? DUGMA_MONO3[k] : DUGMA_MONO2[k];
    double m = 1.0;
    for (int x = 0; x < nv; ++x)
      for (int r = 0; r < e[x]; ++r) m *= dT[x];
    F += coef[k] * m;
  }
  return F;
}

int dor_constraints(int D, const double *dT, double *ceq) {
  if (D == 2) {
    ceq[0] = dT[2] * dT[2] + dT[3] * dT[3] - 1.0;
    return 1;
  }
  /* fcontr3D.m:16-26: R row-major from dT(4:12); ceq(1) = det R - 1;
   * ceq(2..10) = (R'R - I) in MATLAB linear (column-major) order. */
  double r[3][3];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) r[a][b] = dT[3 + 3 * a + b];
  ceq[0] = r[0][0] * (r[1][1] * r[2][2] - r[1][2] * r[2][1]) -
           r[0][1] * (r[1][0] * r[2][2] - r[1][2] * r[2][0]) +
           r[0][2] * (r[1][0] * r[2][1] - r[1][1] * r[2][0]) - 1.0;
  for (int c = 0; c < 3; ++c)
    for (int a = 0; a < 3; ++a) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += r[k][a] * r[k][c];
      ceq[1 + a + 3 * c] = s - (a == c ? 1.0 : 0.0);
    }
  return 10;
}
