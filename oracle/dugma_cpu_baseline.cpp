// dugma_cpu_baseline.cpp -- the CPU baseline of the registration hot path.
//
// TEST INFRASTRUCTURE ONLY.  bench.py's cpu_baseline leg times these loops
// on the host cores; tests/test_cpu_baseline.py checks them against the
// oracle (dugma_oracle.c).  The product (libdugma_hip.so) never links or
// calls this library.
//
// An OpenMP restatement of the path's per-pair loops, written for speed on
// x86-64 (AVX2 + FMA, vectorised expf/exp from glibc's libmvec under
// -ffast-math), one parallel loop over the outer point set, the inner loop
// vectorised:
//   - the convolved energy, forward and fused with the 6-DoF pose gradient,
//     in FP32 per pair (FP64 per-target sums) and in FP64; central finite
//     differences of the forward pass (13 passes per gradient);
//   - update_sig's quantised nearest-distance loop (update_sig.cu:238-260,
//     79-91, 272-315, 384-385);
//   - compi as the reference computes it: per pair the weight P_ij
//     (compi.cu:747-772, FP64 exp) and the expansion of
//     P_ij [u^T A'_j u + d^T B_i d] into the 637 polynomial coefficients
//     (compi.cu:775-3403), here from that closed form, about 1,500
//     multiply-adds per pair into per-thread FP64 accumulators;
//   - compi restructured: per source j the P-weighted sums of 20 target
//     features, then one expansion per source.
// Every timed entry point runs once untimed, then `reps` times, and returns
// the median wall time (std::chrono::steady_clock).  Inputs are finite
// (-ffast-math): the bench's workloads have no NaN or infinite point.
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../dugma_amd/csrc/dugma_monomials.inc"

namespace {

template <class F>
double median_time(int reps, F &&fn) {
  fn();  // untimed warm-up
  std::vector<double> ts;
  for (int r = 0; r < std::max(reps, 1); ++r) {
    auto t0 = std::chrono::steady_clock::now();
    fn();
    ts.push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
  }
  std::sort(ts.begin(), ts.end());
  return ts[ts.size() / 2];
}

// ---------------------------------------------------------------------------
// Convolved energy  E = sum_i sum_j N(y_i - (R x_j + t); S_i + R S_j R^T)
// ---------------------------------------------------------------------------

// Sources transformed by the pose (FP64 arithmetic, stored as T), SoA.
template <class T>
struct Src {
  std::vector<T> x, y, z, s11, s12, s13, s22, s23, s33;  // s11 alone when isotropic
};

template <class T>
void transform(int M, const double *X, const double *SX, bool iso, const double *R, const double *t,
               Src<T> &o) {
  o.x.resize(M); o.y.resize(M); o.z.resize(M); o.s11.resize(M);
  if (!iso) { o.s12.resize(M); o.s13.resize(M); o.s22.resize(M); o.s23.resize(M); o.s33.resize(M); }
#pragma omp parallel for schedule(static)
  for (int j = 0; j < M; ++j) {
    const double *p = X + 3 * (size_t)j;
    double q[3];
    for (int a = 0; a < 3; ++a) q[a] = R[a] * p[0] + R[a + 3] * p[1] + R[a + 6] * p[2] + t[a];
    o.x[j] = (T)q[0]; o.y[j] = (T)q[1]; o.z[j] = (T)q[2];
    if (iso) {
      o.s11[j] = (T)SX[j];
      continue;
    }
    const double *S = SX + 9 * (size_t)j;  // column-major block
    double RS[9], Sp[9];
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < 3; ++c) {
        double s = 0;
        for (int b = 0; b < 3; ++b) s += R[a + 3 * b] * S[b + 3 * c];
        RS[a + 3 * c] = s;
      }
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < 3; ++c) {
        double s = 0;
        for (int b = 0; b < 3; ++b) s += RS[a + 3 * b] * R[c + 3 * b];
        Sp[a + 3 * c] = s;
      }
    o.s11[j] = (T)Sp[0]; o.s22[j] = (T)Sp[4]; o.s33[j] = (T)Sp[8];
    o.s12[j] = (T)(0.5 * (Sp[3] + Sp[1]));
    o.s13[j] = (T)(0.5 * (Sp[6] + Sp[2]));
    o.s23[j] = (T)(0.5 * (Sp[7] + Sp[5]));
  }
}

constexpr double kInvTwoPi32 = 0.063493635934240969785;  // (2 pi)^(-3/2)
constexpr int kBlk = 512;  // sources per FP32 partial sum

// Per-target sums: E_i, EW_i = sum_j e S^-1 d, and (GRAD, full covariance)
// B_i = sum_j e (S^-1 d d^T S^-1 - S^-1) as 6 symmetric entries.  d = y - x'.
template <class T, bool ISO, bool GRAD>
void target_sums(const double *y, const double *Sy, int M, const Src<T> &s, double out[10]) {
  const T y0 = (T)y[0], y1 = (T)y[1], y2 = (T)y[2];
  T a11, a12 = 0, a13 = 0, a22 = 0, a23 = 0, a33 = 0;
  if (ISO) {
    a11 = (T)Sy[0];
  } else {
    a11 = (T)Sy[0]; a22 = (T)Sy[4]; a33 = (T)Sy[8];
    a12 = (T)(0.5 * (Sy[3] + Sy[1])); a13 = (T)(0.5 * (Sy[6] + Sy[2])); a23 = (T)(0.5 * (Sy[7] + Sy[5]));
  }
  for (int k = 0; k < 10; ++k) out[k] = 0;
  const T *X = s.x.data(), *Yv = s.y.data(), *Z = s.z.data(), *S11 = s.s11.data();
  const T *S12 = s.s12.data(), *S13 = s.s13.data(), *S22 = s.s22.data(), *S23 = s.s23.data(),
          *S33 = s.s33.data();
  const T c0 = (T)kInvTwoPi32;
  for (int j0 = 0; j0 < M; j0 += kBlk) {
    const int j1 = std::min(M, j0 + kBlk);
    T e = 0, w0 = 0, w1 = 0, w2 = 0, b11 = 0, b12 = 0, b13 = 0, b22 = 0, b23 = 0, b33 = 0;
    if (ISO) {
#pragma omp simd reduction(+ : e, w0, w1, w2)
      for (int j = j0; j < j1; ++j) {
        const T dx = y0 - X[j], dy = y1 - Yv[j], dz = y2 - Z[j];
        const T v = a11 + S11[j];
        const T iv = (T)1 / v;
        const T q = (dx * dx + dy * dy + dz * dz) * iv;
        const T ee = c0 * iv * std::sqrt(iv) * std::exp((T)-0.5 * q);
        e += ee;
        if (GRAD) {
          const T we = ee * iv;
          w0 += we * dx; w1 += we * dy; w2 += we * dz;
        }
      }
    } else {
#pragma omp simd reduction(+ : e, w0, w1, w2, b11, b12, b13, b22, b23, b33)
      for (int j = j0; j < j1; ++j) {
        const T dx = y0 - X[j], dy = y1 - Yv[j], dz = y2 - Z[j];
        const T s11 = a11 + S11[j], s12 = a12 + S12[j], s13 = a13 + S13[j];
        const T s22 = a22 + S22[j], s23 = a23 + S23[j], s33 = a33 + S33[j];
        const T c11 = s22 * s33 - s23 * s23, c12 = s13 * s23 - s12 * s33, c13 = s12 * s23 - s13 * s22;
        const T c22 = s11 * s33 - s13 * s13, c23 = s12 * s13 - s11 * s23, c33 = s11 * s22 - s12 * s12;
        const T det = s11 * c11 + s12 * c12 + s13 * c13;
        const T idet = (T)1 / det;
        const T u0 = c11 * dx + c12 * dy + c13 * dz;
        const T u1 = c12 * dx + c22 * dy + c23 * dz;
        const T u2 = c13 * dx + c23 * dy + c33 * dz;
        const T q = (dx * u0 + dy * u1 + dz * u2) * idet;
        const T ee = c0 * std::sqrt(idet) * std::exp((T)-0.5 * q);
        e += ee;
        if (GRAD) {
          const T x0 = u0 * idet, x1 = u1 * idet, x2 = u2 * idet;  // S^-1 d
          w0 += ee * x0; w1 += ee * x1; w2 += ee * x2;
          const T ei = ee * idet;
          b11 += ee * x0 * x0 - ei * c11; b12 += ee * x0 * x1 - ei * c12; b13 += ee * x0 * x2 - ei * c13;
          b22 += ee * x1 * x1 - ei * c22; b23 += ee * x1 * x2 - ei * c23; b33 += ee * x2 * x2 - ei * c33;
        }
      }
    }
    out[0] += e; out[1] += w0; out[2] += w1; out[3] += w2;
    out[4] += b11; out[5] += b12; out[6] += b13; out[7] += b22; out[8] += b23; out[9] += b33;
  }
}

// E and (GRAD) the pose gradient [dE/dw, dE/dt] for R <- exp([w]x) R from
// target-side sums alone (the energy is invariant under one rigid motion of
// both clouds):  dE/dt = sum_i EW_i,
//   dE/dw = sum_i (y_i - t) x EW_i + vee(S_i B_i - B_i S_i).
template <class T, bool ISO, bool GRAD>
double energy_pass(int N, const double *Y, const double *SY, int M, const Src<T> &s, const double *t,
                   double *gpose) {
  double E = 0, g[6] = {0, 0, 0, 0, 0, 0};
#pragma omp parallel
  {
    double El = 0, gl[6] = {0, 0, 0, 0, 0, 0};
#pragma omp for schedule(dynamic, 16) nowait
    for (int i = 0; i < N; ++i) {
      double o[10];
      const double *Sy = ISO ? SY + i : SY + 9 * (size_t)i;
      target_sums<T, ISO, GRAD>(Y + 3 * (size_t)i, Sy, M, s, o);
      El += o[0];
      if (GRAD) {
        const double q0 = Y[3 * i] - t[0], q1 = Y[3 * i + 1] - t[1], q2 = Y[3 * i + 2] - t[2];
        gl[0] += q1 * o[3] - q2 * o[2];
        gl[1] += q2 * o[1] - q0 * o[3];
        gl[2] += q0 * o[2] - q1 * o[1];
        gl[3] += o[1]; gl[4] += o[2]; gl[5] += o[3];
        if (!ISO) {
          // K = S B - B S (antisymmetric), vee(K) = (K32, K13, K21)
          const double *C = Sy;
          double S[9] = {C[0], 0.5 * (C[1] + C[3]), 0.5 * (C[2] + C[6]), 0, C[4], 0.5 * (C[5] + C[7]), 0, 0, C[8]};
          S[3] = S[1]; S[6] = S[2]; S[7] = S[5];
          const double B[9] = {o[4], o[5], o[6], o[5], o[7], o[8], o[6], o[8], o[9]};
          auto K = [&](int a, int c) {
            double v = 0;
            for (int b = 0; b < 3; ++b) v += S[a + 3 * b] * B[b + 3 * c] - B[a + 3 * b] * S[b + 3 * c];
            return v;
          };
          gl[0] += K(2, 1); gl[1] += K(0, 2); gl[2] += K(1, 0);
        }
      }
    }
#pragma omp critical
    {
      E += El;
      for (int k = 0; k < 6; ++k) g[k] += gl[k];
    }
  }
  if (GRAD && gpose)
    for (int k = 0; k < 6; ++k) gpose[k] = g[k];
  return E;
}

void expw(const double *w, const double *R, double *Rn) {
  const double th = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  double K[9] = {0, w[2], -w[1], -w[2], 0, w[0], w[1], -w[0], 0};  // [w]x column-major
  double K2[9], E[9];
  for (int a = 0; a < 3; ++a)
    for (int c = 0; c < 3; ++c) {
      double s = 0;
      for (int b = 0; b < 3; ++b) s += K[a + 3 * b] * K[b + 3 * c];
      K2[a + 3 * c] = s;
    }
  const double A = th > 0 ? std::sin(th) / th : 1.0, B = th > 0 ? (1 - std::cos(th)) / (th * th) : 0.5;
  for (int k = 0; k < 9; ++k) E[k] = (k % 4 == 0 ? 1.0 : 0.0) + A * K[k] + B * K2[k];
  for (int a = 0; a < 3; ++a)
    for (int c = 0; c < 3; ++c) {
      double s = 0;
      for (int b = 0; b < 3; ++b) s += E[a + 3 * b] * R[b + 3 * c];
      Rn[a + 3 * c] = s;
    }
}

template <class T, bool ISO>
double energy_mode(int mode, int N, const double *Y, const double *SY, int M, const double *X,
                   const double *SX, const double *R, const double *t, double h, double *gpose) {
  Src<T> s;
  transform<T>(M, X, SX, ISO, R, t, s);
  if (mode == 0) return energy_pass<T, ISO, false>(N, Y, SY, M, s, t, nullptr);
  if (mode == 1) return energy_pass<T, ISO, true>(N, Y, SY, M, s, t, gpose);
  // central differences over (w, t): 12 shifted forward passes + the centre
  const double E = energy_pass<T, ISO, false>(N, Y, SY, M, s, t, nullptr);
  for (int k = 0; k < 6; ++k) {
    double f[2];
    for (int sgn = 0; sgn < 2; ++sgn) {
      double w[3] = {0, 0, 0}, tn[3] = {t[0], t[1], t[2]}, Rn[9];
      const double st = sgn ? -h : h;
      if (k < 3) w[k] = st; else tn[k - 3] += st;
      expw(w, R, Rn);
      transform<T>(M, X, SX, ISO, Rn, tn, s);
      f[sgn] = energy_pass<T, ISO, false>(N, Y, SY, M, s, tn, nullptr);
    }
    if (gpose) gpose[k] = (f[0] - f[1]) / (2 * h);
  }
  return E;
}

// ---------------------------------------------------------------------------
// compi
// ---------------------------------------------------------------------------

// Monomial code of up to 4 variables (0..11), sorted, base 13 with 12 = none.
inline int mcode(int a, int b, int c, int d) {
  int v[4] = {a, b, c, d};
  std::sort(v, v + 4);
  return ((v[0] * 13 + v[1]) * 13 + v[2]) * 13 + v[3];
}
constexpr int kNone = 12;  // "no variable" sorts last
constexpr int kCodes = 13 * 13 * 13 * 13;

struct MonoMap {
  std::vector<int16_t> slot;  // code -> position in energy3D.m order, -1 outside
  MonoMap() : slot(kCodes, -1) {
    for (int k = 0; k < DUGMA_NMONO3; ++k) {
      int v[4] = {kNone, kNone, kNone, kNone}, n = 0;
      for (int x = 0; x < 12; ++x)
        for (int r = 0; r < DUGMA_MONO3[k][x]; ++r) v[n++] = x;
      slot[mcode(v[0], v[1], v[2], v[3])] = (int16_t)k;
    }
  }
};

// Variables: t1 t2 t3 = 0..2, r_ab (row-major dR) = 3 + 3a + b.
inline int rv(int a, int b) { return 3 + 3 * a + b; }

// The expansion tables, fixed for D = 3.  With d_a = n_a - t_a - sum_b m_b r_ab
// and u_c = sum_a r_ac d_a, u_c has 15 terms:
//   k = a      : n_a r_ac
//   k = 3 + a  : -t_a r_ac
//   k = 6+3a+b : -m_b r_ac r_ab
// and d_a has 5: n_a, -t_a, -m_b r_ab.  uu[p][q][k1][k2] / dd[a][b][k1][k2]
// give the output slot of the product of two terms.
struct Tables {
  MonoMap mm;
  int16_t uu[3][3][15][15];
  int16_t dd[3][3][5][5];
  // a term as up to two variables
  static void uterm(int c, int k, int &v0, int &v1) {
    if (k < 3) { v0 = rv(k, c); v1 = kNone; }
    else if (k < 6) { v0 = rv(k - 3, c); v1 = k - 3; }
    else { const int a = (k - 6) / 3, b = (k - 6) % 3; v0 = rv(a, c); v1 = rv(a, b); }
  }
  // a product outside the 637 monomials (none arises) lands in a spare slot
  static int16_t spare(int16_t s) { return s < 0 ? (int16_t)DUGMA_NMONO3 : s; }
  static void dterm(int a, int k, int &v0) {
    v0 = k == 0 ? kNone : (k == 1 ? a : rv(a, k - 2));
  }
  Tables() {
    for (int p = 0; p < 3; ++p)
      for (int q = 0; q < 3; ++q)
        for (int k1 = 0; k1 < 15; ++k1)
          for (int k2 = 0; k2 < 15; ++k2) {
            int a0, a1, b0, b1;
            uterm(p, k1, a0, a1);
            uterm(q, k2, b0, b1);
            uu[p][q][k1][k2] = spare(mm.slot[mcode(a0, a1, b0, b1)]);
          }
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        for (int k1 = 0; k1 < 5; ++k1)
          for (int k2 = 0; k2 < 5; ++k2) {
            int v0, v1;
            dterm(a, k1, v0);
            dterm(b, k2, v1);
            dd[a][b][k1][k2] = spare(mm.slot[mcode(v0, v1, kNone, kNone)]);
          }
  }
};

const Tables &tables() {
  static const Tables T;
  return T;
}

struct CompiScale {
  float Isig, gsc;
};
CompiScale compi_scale(double sig) {
  CompiScale s;
  s.Isig = (float)(1.0 / sig);                // __double2float_rn (compi.cu:646)
  s.gsc = std::sqrt(s.Isig * s.Isig * s.Isig);  // compi.cu:662
  return s;
}

// Per source j: m' = fl32(R y_j + t), A'_j = fl32(R A_j^T R^T) (FP64), and
// the scaled A_j, gY_j (compi.cu:655-662, 775-791).
struct SrcState {
  double m[3];   // FP64 copy of fl32(R y + t)
  double y[3];   // FP32 source point
  double C[9];   // FP64 R A^T R^T (P's quadratic form)
  double Ap[9];  // fl32(C)
  float g;
};

void source_states(int M, const float *Yf, const float *YCinv, const float *gY, const double *R,
                   const double *t, CompiScale sc, std::vector<SrcState> &o) {
  o.resize(M);
#pragma omp parallel for schedule(static)
  for (int j = 0; j < M; ++j) {
    SrcState &s = o[j];
    float A[9];
    for (int k = 0; k < 9; ++k) A[k] = YCinv[9 * (size_t)j + k] * sc.Isig;
    s.g = gY[j] * sc.gsc;
    for (int a = 0; a < 3; ++a) s.y[a] = Yf[3 * (size_t)j + a];
    for (int p = 0; p < 3; ++p)
      for (int q = 0; q < 3; ++q) {
        double v = 0;
        for (int k = 0; k < 3; ++k) {
          double inner = 0;
          for (int l = 0; l < 3; ++l) inner += (double)A[l + 3 * k] * R[q + 3 * l];
          v += R[p + 3 * k] * inner;
        }
        s.C[p + 3 * q] = v;
        s.Ap[p + 3 * q] = (double)(float)v;
      }
    for (int a = 0; a < 3; ++a) {
      double v = R[a] * s.y[0] + R[a + 3] * s.y[1] + R[a + 6] * s.y[2] + t[a];
      s.m[a] = (double)(float)v;
    }
  }
}

struct TgtState {
  double n[3];
  double B[9];  // fl32(inv(CX) Isig)
  float g;
};

void target_states(int N, const float *Xf, const float *XCinv, const float *gX, CompiScale sc,
                   std::vector<TgtState> &o) {
  o.resize(N);
  for (int i = 0; i < N; ++i) {
    for (int a = 0; a < 3; ++a) o[i].n[a] = Xf[3 * (size_t)i + a];
    for (int k = 0; k < 9; ++k) o[i].B[k] = (double)(XCinv[9 * (size_t)i + k] * sc.Isig);
    o[i].g = gX[i] * sc.gsc;
  }
}

// P_ij (compi.cu:747-772): FP64 Mahalanobis forms, FP64 exp.
inline double pair_weight(const TgtState &x, const SrcState &s, const double *R, const double *t) {
  double d[3];
  for (int a = 0; a < 3; ++a)
    d[a] = x.n[a] - t[a] - s.y[0] * R[a] - s.y[1] * R[a + 3] - s.y[2] * R[a + 6];
  double m1 = 0, m2 = 0;
  for (int p = 0; p < 3; ++p)
    for (int q = 0; q < 3; ++q) {
      m1 += d[p] * s.C[p + 3 * q] * d[q];
      m2 += d[p] * x.B[p + 3 * q] * d[q];
    }
  const float gg = x.g * s.g;
  return (double)gg * (std::exp(-0.5 * m1) + std::exp(-0.5 * m2));
}

// One pair (or, with moments, one source): acc += the 637 coefficients of
//   w [u^T A' u + d^T B d].
// Term values of u_c: k < 3: n_a, 3..5: -1, 6..14: -m_b.  Products of two
// n-terms use NN[a][a'] (sum of P n_a n_a'), n x constant uses N1[a], and
// constant x constant uses W0 (per pair: NN = n n^T, N1 = n, W0 = 1).
struct Moments {
  double W0, N1[3], NN[9];
  double s, v[3], B[9];  // d^T B d: sum P n^T B n, sum P (B + B^T) n, sum P sym(B)
};

void expand(const Moments &mo, const double *m, const double *Ap, double *acc) {
  const Tables &T = tables();
  // weights of the 15 x 15 term products (the same for every (p, q))
  double c[15];
  for (int k = 0; k < 15; ++k) c[k] = k < 3 ? 1.0 : (k < 6 ? -1.0 : -m[(k - 6) % 3]);
  double Wm[225];
  for (int k1 = 0; k1 < 15; ++k1)
    for (int k2 = 0; k2 < 15; ++k2) {
      double w;
      if (k1 < 3 && k2 < 3) w = mo.NN[k1 + 3 * k2];
      else if (k1 < 3) w = mo.N1[k1] * c[k2];
      else if (k2 < 3) w = mo.N1[k2] * c[k1];
      else w = mo.W0 * c[k1] * c[k2];
      Wm[15 * k1 + k2] = w;
    }
  // u^T A' u: 9 x 225 multiply-adds into the output slots
  for (int p = 0; p < 3; ++p)
    for (int q = 0; q < 3; ++q) {
      const double ap = Ap[p + 3 * q];
      const int16_t *sl = &T.uu[p][q][0][0];
      for (int k = 0; k < 225; ++k) acc[sl[k]] += ap * Wm[k];
    }
  // d^T B d = n^T B n + sum_b v_b l_b + l^T sym(B) l,  l_a = -t_a - sum_b m_b r_ab
  // constant of term k >= 1 of d_a: -1 (t_a), -m_{k-2} (r_a,k-2)
  auto dc = [&](int k) { return k == 1 ? -1.0 : -m[k - 2]; };
  acc[T.dd[0][0][0][0]] += mo.s;
  for (int b = 0; b < 3; ++b)
    for (int k = 1; k < 5; ++k) acc[T.dd[b][0][k][0]] += mo.v[b] * dc(k);
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      const double bab = mo.B[a + 3 * b];
      for (int k1 = 1; k1 < 5; ++k1)
        for (int k2 = 1; k2 < 5; ++k2)
          acc[T.dd[a][b][k1][k2]] += bab * dc(k1) * dc(k2);
    }
}

// Per-pair moments of one target (weight P).
inline void pair_moments(const TgtState &x, double P, Moments &mo) {
  mo.W0 = P;
  for (int a = 0; a < 3; ++a) mo.N1[a] = P * x.n[a];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) mo.NN[a + 3 * b] = P * x.n[a] * x.n[b];
  double s = 0;
  for (int a = 0; a < 3; ++a) {
    double va = 0;
    for (int b = 0; b < 3; ++b) {
      s += x.n[a] * x.B[a + 3 * b] * x.n[b];
      va += (x.B[a + 3 * b] + x.B[b + 3 * a]) * x.n[b];
    }
    mo.v[a] = P * va;
  }
  mo.s = P * s;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) mo.B[a + 3 * b] = P * 0.5 * (x.B[a + 3 * b] + x.B[b + 3 * a]);
}

void compi(int form, int N, const float *Xf, const float *XCinv, const float *gX, int M, const float *Yf,
           const float *YCinv, const float *gY, const double *R, const double *t, double sig, double *coef) {
  const CompiScale sc = compi_scale(sig);
  std::vector<SrcState> src;
  std::vector<TgtState> tgt;
  source_states(M, Yf, YCinv, gY, R, t, sc, src);
  target_states(N, Xf, XCinv, gX, sc, tgt);
  (void)tables();
  const int nt = omp_get_max_threads();
  constexpr int kAcc = DUGMA_NMONO3 + 1;  // + the spare slot
  std::vector<double> accs((size_t)nt * kAcc, 0.0);
#pragma omp parallel
  {
    double *acc = accs.data() + (size_t)omp_get_thread_num() * kAcc;
    if (form == 0) {
      // the reference's per-pair expansion
#pragma omp for schedule(dynamic, 4)
      for (int i = 0; i < N; ++i)
        for (int j = 0; j < M; ++j) {
          // every pair is expanded, P_ij = 0 included, as compi.cu:775-3350 does
          const double P = pair_weight(tgt[i], src[j], R, t);
          Moments mo;
          pair_moments(tgt[i], P, mo);
          expand(mo, src[j].m, src[j].Ap, acc);
        }
    } else {
      // restructured: per source, P-weighted target features, one expansion
#pragma omp for schedule(dynamic, 16)
      for (int j = 0; j < M; ++j) {
        Moments S;
        std::memset(&S, 0, sizeof(S));
        for (int i = 0; i < N; ++i) {
          const double P = pair_weight(tgt[i], src[j], R, t);
          const TgtState &x = tgt[i];
          S.W0 += P;
          for (int a = 0; a < 3; ++a) S.N1[a] += P * x.n[a];
          for (int a = 0; a < 3; ++a)
            for (int b = a; b < 3; ++b) S.NN[a + 3 * b] += P * x.n[a] * x.n[b];
          double s = 0;
          for (int a = 0; a < 3; ++a) {
            double va = 0;
            for (int b = 0; b < 3; ++b) {
              s += x.n[a] * x.B[a + 3 * b] * x.n[b];
              va += (x.B[a + 3 * b] + x.B[b + 3 * a]) * x.n[b];
            }
            S.v[a] += P * va;
          }
          S.s += P * s;
          for (int a = 0; a < 3; ++a)
            for (int b = a; b < 3; ++b) S.B[a + 3 * b] += P * 0.5 * (x.B[a + 3 * b] + x.B[b + 3 * a]);
        }
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < a; ++b) {
            S.NN[a + 3 * b] = S.NN[b + 3 * a];
            S.B[a + 3 * b] = S.B[b + 3 * a];
          }
        expand(S, src[j].m, src[j].Ap, acc);
      }
    }
  }
  for (int k = 0; k < DUGMA_NMONO3; ++k) {
    double s = 0;
    for (int r = 0; r < nt; ++r) s += accs[(size_t)r * kAcc + k];
    coef[k] = s;
  }
}

// ---------------------------------------------------------------------------
// update_sig (update_sig.cu:238-260, 79-91, 272-315, 384-385)
// ---------------------------------------------------------------------------
float update_sig(int N, const float *Xf, int M, const float *Yf, const double *R, const double *t) {
  std::vector<float> mx(M), my(M), mz(M);
  for (int j = 0; j < M; ++j) {
    const float *y = Yf + 3 * (size_t)j;
    mx[j] = (float)(R[0] * y[0] + R[3] * y[1] + R[6] * y[2] + t[0]);
    my[j] = (float)(R[1] * y[0] + R[4] * y[1] + R[7] * y[2] + t[1]);
    mz[j] = (float)(R[2] * y[0] + R[5] * y[1] + R[8] * y[2] + t[2]);
  }
  std::vector<float> v(N);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; ++i) {
    const float n0 = Xf[3 * (size_t)i], n1 = Xf[3 * (size_t)i + 1], n2 = Xf[3 * (size_t)i + 2];
    float best = 3.4e38f;
#pragma omp simd reduction(min : best)
    for (int j = 0; j < M; ++j) {
      const float a = n0 - mx[j], b = n1 - my[j], c = n2 - mz[j];
      best = std::min(best, a * a + b * b + c * c);
    }
    // quantise the minimum (monotone, so the min of the quantised values)
    const float s = (best / 3.0f) * 1000000.0f;
    const uint32_t q = s >= 4e9f ? 4000000000u : (uint32_t)std::nearbyint(s);
    v[i] = (float)q / 1000000.0f;
  }
  // pairwise tree sum (update_sig.cu:291-307)
  int n = N;
  while (n > 1) {
    const int h = n / 2;
    for (int k = 0; k < h; ++k) v[k] = v[2 * k] + v[2 * k + 1];
    if (n & 1) v[h] = v[n - 1];
    n = h + (n & 1);
  }
  return N > 0 ? v[0] / (float)N : 0.0f;
}

}  // namespace

extern "C" {

int dcb_num_threads(void) { return omp_get_max_threads(); }

// prec 32 | 64; mode 0 forward, 1 fused (E + pose gradient), 2 central
// differences (step h); iso: SY / SX hold one variance per point, else
// 3x3 column-major blocks.  3-D.  Returns E; *median_s the median time.
double dcb_energy(int prec, int mode, int iso, int N, const double *Y, const double *SY, int M,
                  const double *X, const double *SX, const double *R, const double *t, double h,
                  double *gpose, int reps, double *median_s) {
  double E = 0;
  auto run = [&]() {
    if (prec == 32)
      E = iso ? energy_mode<float, true>(mode, N, Y, SY, M, X, SX, R, t, h, gpose)
              : energy_mode<float, false>(mode, N, Y, SY, M, X, SX, R, t, h, gpose);
    else
      E = iso ? energy_mode<double, true>(mode, N, Y, SY, M, X, SX, R, t, h, gpose)
              : energy_mode<double, false>(mode, N, Y, SY, M, X, SX, R, t, h, gpose);
  };
  const double ts = median_time(reps, run);
  if (median_s) *median_s = ts;
  return E;
}

// compi, 3-D (637 coefficients in energy3D.m order): form 0 per-pair
// expansion, form 1 per-source moments.
void dcb_compi(int form, int N, const float *Xf, const float *XCinv, const float *gX, int M, const float *Yf,
               const float *YCinv, const float *gY, const double *R, const double *t, double sig, double *coef,
               int reps, double *median_s) {
  const double ts = median_time(reps, [&]() { compi(form, N, Xf, XCinv, gX, M, Yf, YCinv, gY, R, t, sig, coef); });
  if (median_s) *median_s = ts;
}

float dcb_update_sig(int N, const float *Xf, int M, const float *Yf, const double *R, const double *t, int reps,
                     double *median_s) {
  float sig = 0;
  const double ts = median_time(reps, [&]() { sig = update_sig(N, Xf, M, Yf, R, t); });
  if (median_s) *median_s = ts;
  return sig;
}

}  // extern "C"
