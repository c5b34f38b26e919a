// dugma_register.cpp -- host side of the reference's registration loop
// (DUGMA.m:19-137, normalize.m:3-26) over the C ABI, and the surrogate
// polynomial's objective / gradient / minimiser (energy3D.m, energy2D.m,
// fcontr3D.m, fcontr2D.m) that MATLAB's fmincon provides in the reference.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "dugma_internal.h"
#include "dugma_monomials.inc"

namespace dugma {

// ---- the polynomial of energy3D.m / energy2D.m ---------------------------
static void mono_table(int D, int *nm, int *nv) {
  *nm = D == 3 ? DUGMA_NMONO3 : DUGMA_NMONO2;
  *nv = D == 3 ? 12 : 4;
}

static const unsigned char *mono(int D, int k) {
  return D == 3 ? DUGMA_MONO3[k] : DUGMA_MONO2[k];
}

// Monomials as lists of at most four variable factors (with repetition),
// built once per D from the exponent table.
struct MonoList {
  int nm = 0, nv = 0;
  std::vector<int> deg;
  std::vector<std::array<int, 4>> var;
};

static const MonoList &mono_list(int D) {
  static MonoList lists[2];
  static std::once_flag once[2];
  std::call_once(once[D - 2], [D] {
    MonoList &L = lists[D - 2];
    mono_table(D, &L.nm, &L.nv);
    L.deg.resize(L.nm);
    L.var.resize(L.nm);
    for (int k = 0; k < L.nm; ++k) {
      const unsigned char *e = mono(D, k);
      int d = 0;
      std::array<int, 4> v{0, 0, 0, 0};
      for (int x = 0; x < L.nv; ++x)
        for (int r = 0; r < e[x] && d < 4; ++r) v[d++] = x;
      L.deg[k] = d;
      L.var[k] = v;
    }
  });
  return lists[D - 2];
}

// F = sum_k coef_k prod_v dT_v^e_kv and, if asked, dF/ddT (nv) and the full
// Hessian d2F/ddT2 (nv x nv, row-major).  Per monomial the products leave
// out one factor (gradient) or two (Hessian); a repeated variable is counted
// once per factor position, which gives the powers' derivatives.
static double poly_value(int D, const double *coef, const double *dT, double *grad,
                         double *hess = nullptr) {
  const MonoList &L = mono_list(D);
  const int nv = L.nv;
  if (grad)
    for (int v = 0; v < nv; ++v) grad[v] = 0.0;
  if (hess)
    for (int v = 0; v < nv * nv; ++v) hess[v] = 0.0;
  double F = 0.0;
  for (int k = 0; k < L.nm; ++k) {
    const double c = coef[k];
    if (c == 0.0) continue;
    const int d = L.deg[k];
    const std::array<int, 4> &v = L.var[k];
    double f[4] = {1.0, 1.0, 1.0, 1.0};
    for (int q = 0; q < d; ++q) f[q] = dT[v[q]];
    F += c * ((f[0] * f[1]) * (f[2] * f[3]));
    if (grad)
      for (int q = 0; q < d; ++q) {
        double m = c;
        for (int r = 0; r < d; ++r)
          if (r != q) m *= f[r];
        grad[v[q]] += m;
      }
    if (hess)
      for (int q = 0; q < d; ++q)
        for (int r = q + 1; r < d; ++r) {
          double m = c;
          for (int t = 0; t < d; ++t)
            if (t != q && t != r) m *= f[t];
          hess[v[q] * nv + v[r]] += m;
          hess[v[r] * nv + v[q]] += m;
        }
  }
  return F;
}

// dT of a pose increment: 3-D [dt (3), dR row-major (9)]; 2-D [dt, r11, r21].
static void pack_dT(int D, const double *dR, const double *dt, double *dT) {
  if (D == 3) {
    for (int a = 0; a < 3; ++a) dT[a] = dt[a];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) dT[3 + 3 * a + b] = dR[a + 3 * b];
  } else {
    dT[0] = dt[0];
    dT[1] = dt[1];
    dT[2] = dR[0];
    dT[3] = dR[1];
  }
}

static void rot_exp3(const double *w, double *R) {
  const double th = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  const double A = th < 1e-12 ? 1.0 : std::sin(th) / th;
  const double B = th < 1e-12 ? 0.5 : (1.0 - std::cos(th)) / (th * th);
  const double K[9] = {0, w[2], -w[1], -w[2], 0, w[0], w[1], -w[0], 0};
  for (int a = 0; a < 3; ++a)
    for (int c = 0; c < 3; ++c) {
      double k2 = 0;
      for (int b = 0; b < 3; ++b) k2 += K[a + 3 * b] * K[b + 3 * c];
      R[a + 3 * c] = (a == c ? 1.0 : 0.0) + A * K[a + 3 * c] + B * k2;
    }
}

static void matmul(int D, const double *A, const double *B, double *C) {
  for (int a = 0; a < D; ++a)
    for (int c = 0; c < D; ++c) {
      double s = 0;
      for (int b = 0; b < D; ++b) s += A[a + D * b] * B[b + D * c];
      C[a + D * c] = s;
    }
}

// Pose increment from parameters p (3-D: w, dt; 2-D: theta, dt) applied on
// the left of dR0:  dR = exp([w]x) dR0, dt = dt0 + p_t.
static void apply(int D, const double *dR0, const double *dt0, const double *p,
                  double *dR, double *dt) {
  if (D == 3) {
    double E[9];
    rot_exp3(p, E);
    matmul(3, E, dR0, dR);
    for (int a = 0; a < 3; ++a) dt[a] = dt0[a] + p[3 + a];
  } else {
    const double c = std::cos(p[0]), s = std::sin(p[0]);
    const double E[4] = {c, s, -s, c};
    matmul(2, E, dR0, dR);
    dt[0] = dt0[0] + p[1];
    dt[1] = dt0[1] + p[2];
  }
}

// Objective on the constraint set (dR a rotation) as a function of the
// parameters p of a left increment (3-D: w, dt; 2-D: theta, dt) of (dR, dt).
static double obj_p(int D, const double *coef, const double *dR0, const double *dt0,
                    const double *p) {
  double dR[9], dt[3], dT[12];
  apply(D, dR0, dt0, p, dR, dt);
  pack_dT(D, dR, dt, dT);
  return poly_value(D, coef, dT, nullptr);
}

// Value, gradient and Hessian in p at p = 0, from the polynomial's own
// gradient g and Hessian H in dT by the chain rule: with J = d dT / d p,
// grad_p = J^T g and hess_p = J^T H J + sum_v g_v d2 dT_v / dp dp.
// 3-D: d dR / d w_a = [e_a]x dR, d2 dR / d w_a d w_b = (K_a K_b + K_b K_a) dR / 2;
// 2-D: dR = Rot(theta) dR0, so d (r11, r21) / d theta = (-r21, r11) and the
// second derivative is (-r11, -r21).
static double obj_p_newton(int D, const double *coef, const double *dR, const double *dt, double *gp,
                           double *Hp) {
  const int nv = D == 3 ? 12 : 4, np = D == 3 ? 6 : 3;
  double dT[12], g[12], H[144], J[12][6];
  pack_dT(D, dR, dt, dT);
  const double F = poly_value(D, coef, dT, g, H);
  for (int v = 0; v < nv; ++v)
    for (int a = 0; a < np; ++a) J[v][a] = 0.0;
  double S2[3][3][9];  // second derivatives of dR (column-major) in w_a, w_b
  if (D == 3) {
    double K[3][9], KR[3][9];
    for (int a = 0; a < 3; ++a) {
      double ea[3] = {0, 0, 0};
      ea[a] = 1.0;
      const double Ka[9] = {0, ea[2], -ea[1], -ea[2], 0, ea[0], ea[1], -ea[0], 0};
      memcpy(K[a], Ka, sizeof(Ka));
      matmul(3, K[a], dR, KR[a]);
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) J[3 + 3 * r + c][a] = KR[a][r + 3 * c];
      J[a][3 + a] = 1.0;
    }
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        double KaKbR[9], KbKaR[9];
        matmul(3, K[a], KR[b], KaKbR);
        matmul(3, K[b], KR[a], KbKaR);
        for (int q = 0; q < 9; ++q) S2[a][b][q] = 0.5 * (KaKbR[q] + KbKaR[q]);
      }
  } else {
    J[2][0] = -dR[1];
    J[3][0] = dR[0];
    J[0][1] = 1.0;
    J[1][2] = 1.0;
  }
  for (int a = 0; a < np; ++a) {
    double s = 0.0;
    for (int v = 0; v < nv; ++v) s += J[v][a] * g[v];
    gp[a] = s;
  }
  for (int a = 0; a < np; ++a)
    for (int b = 0; b < np; ++b) {
      double s = 0.0;
      for (int v = 0; v < nv; ++v) {
        if (J[v][a] == 0.0) continue;
        double hv = 0.0;
        for (int u = 0; u < nv; ++u) hv += H[v * nv + u] * J[u][b];
        s += J[v][a] * hv;
      }
      Hp[a * np + b] = s;
    }
  if (D == 3) {
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        double s = 0.0;
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) s += g[3 + 3 * r + c] * S2[a][b][r + 3 * c];
        Hp[a * np + b] += s;
      }
  } else {
    Hp[0] += g[2] * (-dR[0]) + g[3] * (-dR[1]);
  }
  return F;
}

// Damped Newton on p with the analytic Hessian, re-centred on the accepted
// iterate each step.
static int minimise(int D, const double *coef, double *dR, double *dt, double *Fout) {
  const int np = D == 3 ? 6 : 3;
  for (int a = 0; a < D * D; ++a) dR[a] = (a % (D + 1) == 0) ? 1.0 : 0.0;
  for (int a = 0; a < D; ++a) dt[a] = 0.0;
  double zero[6] = {0, 0, 0, 0, 0, 0};
  double F = obj_p(D, coef, dR, dt, zero);
  if (!std::isfinite(F)) return -1;
  double lambda = 1e-3;
  for (int it = 0; it < 100; ++it) {
    double g[6], H[36];
    obj_p_newton(D, coef, dR, dt, g, H);
    bool accepted = false;
    for (int tries = 0; tries < 30 && !accepted; ++tries) {
      // solve (H + lambda (|diag| + 1e-12) I) s = -g by Gaussian elimination
      double A[6][7];
      for (int a = 0; a < np; ++a) {
        for (int b = 0; b < np; ++b) A[a][b] = H[a * np + b];
        A[a][a] += lambda * (std::fabs(H[a * np + a]) + 1e-12);
        A[a][np] = -g[a];
      }
      bool ok = true;
      for (int c = 0; c < np && ok; ++c) {
        int piv = c;
        for (int r = c + 1; r < np; ++r)
          if (std::fabs(A[r][c]) > std::fabs(A[piv][c])) piv = r;
        if (std::fabs(A[piv][c]) < 1e-300) {
          ok = false;
          break;
        }
        for (int k = 0; k <= np; ++k) std::swap(A[c][k], A[piv][k]);
        for (int r = 0; r < np; ++r) {
          if (r == c) continue;
          const double f = A[r][c] / A[c][c];
          for (int k = c; k <= np; ++k) A[r][k] -= f * A[c][k];
        }
      }
      if (ok) {
        double s[6];
        for (int a = 0; a < np; ++a) s[a] = A[a][np] / A[a][a];
        const double Fn = obj_p(D, coef, dR, dt, s);
        if (std::isfinite(Fn) && Fn <= F) {
          double nR[9], nt[3];
          apply(D, dR, dt, s, nR, nt);
          memcpy(dR, nR, sizeof(double) * D * D);
          memcpy(dt, nt, sizeof(double) * D);
          double sn = 0;
          for (int a = 0; a < np; ++a) sn += s[a] * s[a];
          const double dF = F - Fn;
          F = Fn;
          lambda = std::max(lambda * 0.3, 1e-12);
          accepted = true;
          if (sn < 1e-28 || dF <= 1e-15 * std::fabs(F)) {
            *Fout = F;
            return 1;
          }
          break;
        }
      }
      lambda *= 10.0;
    }
    if (!accepted) break;  // no decrease possible: at a (local) minimum
  }
  *Fout = F;
  return 1;
}

// rotm2eul, 'ZYX' (MATLAB Robotics Toolbox convention): R = Rz(a) Ry(b) Rx(c)
static void rotm2eul_zyx(const double *R, double *eul) {
  const double r11 = R[0], r21 = R[1], r31 = R[2], r32 = R[5], r33 = R[8];
  const double r12 = R[3], r22 = R[4];
  const double sy = std::sqrt(r11 * r11 + r21 * r21);
  if (sy > 1e-12) {
    eul[0] = std::atan2(r21, r11);
    eul[1] = std::atan2(-r31, sy);
    eul[2] = std::atan2(r32, r33);
  } else {
    eul[0] = std::atan2(-r12, r22);
    eul[1] = std::atan2(-r31, sy);
    eul[2] = 0.0;
  }
}

}  // namespace dugma

using namespace dugma;

extern "C" double dugma_poly_eval(int D, const double *coef, const double *dT, double *grad) {
  if ((D != 2 && D != 3) || !coef || !dT) return NAN;
  return poly_value(D, coef, dT, grad);
}

extern "C" int dugma_poly_minimize(int D, const double *coef, double *dT, double *F) {
  if ((D != 2 && D != 3) || !coef || !dT) return fail(DUGMA_ERR_ARG, "dugma_poly_minimize: bad argument");
  double dR[9], dt[3], Fv = 0;
  const int ok = minimise(D, coef, dR, dt, &Fv);
  pack_dT(D, dR, dt, dT);
  if (F) *F = Fv;
  return ok > 0 ? DUGMA_OK : fail(DUGMA_ERR_STATE, "dugma_poly_minimize: objective is not finite");
}

extern "C" int dugma_register(const double *X, const double *X_C, const double *Y,
                              const double *Y_C, int N, int M, int D,
                              const double *R0, const double *t0,
                              const dugma_register_opts *o, double *R, double *t,
                              int *iterations, double *sig_out) {
  if (!X || !X_C || !Y || !Y_C || !R0 || !t0 || !o || !R || !t)
    return fail(DUGMA_ERR_ARG, "dugma_register: null argument");
  if (D != 2 && D != 3) return fail(DUGMA_ERR_ARG, "dugma_register: D must be 2 or 3");
  // normalize.m:3-26: centre each cloud, divide both by the larger RMS radius
  std::vector<double> x(X, X + (size_t)N * D), y(Y, Y + (size_t)M * D);
  double xd[3] = {0, 0, 0}, yd[3] = {0, 0, 0};
  for (int i = 0; i < N; ++i)
    for (int a = 0; a < D; ++a) xd[a] += x[(size_t)i * D + a];
  for (int j = 0; j < M; ++j)
    for (int a = 0; a < D; ++a) yd[a] += y[(size_t)j * D + a];
  for (int a = 0; a < D; ++a) {
    xd[a] /= N;
    yd[a] /= M;
  }
  double xs = 0, ys = 0;
  for (int i = 0; i < N; ++i)
    for (int a = 0; a < D; ++a) {
      x[(size_t)i * D + a] -= xd[a];
      xs += x[(size_t)i * D + a] * x[(size_t)i * D + a];
    }
  for (int j = 0; j < M; ++j)
    for (int a = 0; a < D; ++a) {
      y[(size_t)j * D + a] -= yd[a];
      ys += y[(size_t)j * D + a] * y[(size_t)j * D + a];
    }
  xs = std::sqrt(xs / N);
  ys = std::sqrt(ys / M);
  const double scale = xs > ys ? xs : ys;
  for (double &v : x) v /= scale;
  for (double &v : y) v /= scale;
  // DUGMA.m:44-47 (covariances are passed unscaled, as the reference does)
  dugma_ctx *ctx = nullptr;
  double ok = 0;
  int rc = dugma_init(x.data(), X_C, y.data(), Y_C, N, M, D, &ctx, &ok);
  if (rc) return rc;
  double Rc[9], tc[3];
  memcpy(Rc, R0, sizeof(double) * D * D);
  memcpy(tc, t0, sizeof(double) * D);
  double sig = 0;
  rc = dugma_update_sig(ctx, Rc, tc, N, M, D, &sig);
  double angle[3] = {10000, 10000, 10000}, angle0[3];
  double tprev[3];
  memcpy(tprev, tc, sizeof(tc));
  bool con_angle = true, failed = false;
  int iteration = 1;
  std::vector<double> coef(D == 3 ? DUGMA_NMONO3 : DUGMA_NMONO2);
  auto tdiff2 = [&]() {
    double s = 0;
    for (int a = 0; a < D; ++a) s += (tprev[a] - tc[a]) * (tprev[a] - tc[a]);
    return s;
  };
  // DUGMA.m:57 loop condition
  while (!rc && iteration <= o->max_iter && (con_angle || tdiff2() > o->acc_trans) && sig > o->acc_sig) {
    rc = dugma_compi(ctx, Rc, tc, N, M, D, sig, coef.data());
    if (rc) break;
    double dR[9], dt[3], F = 0;
    if (minimise(D, coef.data(), dR, dt, &F) < 0) {
      failed = true;  // DUGMA.m:84-87: exitflag < 0
      break;
    }
    // DUGMA.m:67-78: R <- dR R, t <- dt + t
    double Rn[9];
    matmul(D, dR, Rc, Rn);
    memcpy(Rc, Rn, sizeof(double) * D * D);
    memcpy(tprev, tc, sizeof(tc));
    for (int a = 0; a < D; ++a) tc[a] += dt[a];
    memcpy(angle0, angle, sizeof(angle));
    if (D == 3) {
      rotm2eul_zyx(Rc, angle);
      con_angle = std::fabs(angle0[0] - angle[0]) > o->acc_rot ||
                  std::fabs(angle0[1] - angle[1]) > o->acc_rot ||
                  std::fabs(angle0[2] - angle[2]) > o->acc_rot;
    } else {
      // DUGMA.m:71: asind(R(2)) in degrees
      angle[0] = std::asin(std::max(-1.0, std::min(1.0, Rc[1]))) * 57.29577951308232;
      con_angle = std::fabs(angle0[0] - angle[0]) > o->acc_rot;
    }
    rc = dugma_update_sig(ctx, Rc, tc, N, M, D, &sig);
    if (o->verbose)
      std::printf("Iteration=%d Sigma=%g F=%g\n", iteration, sig, F);
    ++iteration;
  }
  double okf = 0;
  dugma_fr_m(ctx, &okf);
  if (rc) return rc;
  // DUGMA.m:116-123
  if (failed) {
    for (int a = 0; a < D * D; ++a) R[a] = (a % (D + 1) == 0) ? 1.0 : 0.0;
    for (int a = 0; a < D; ++a) t[a] = 0.0;
  } else {
    memcpy(R, Rc, sizeof(double) * D * D);
    for (int a = 0; a < D; ++a) {
      double ry = 0;
      for (int b = 0; b < D; ++b) ry += Rc[a + D * b] * yd[b];
      t[a] = xd[a] + tc[a] * scale - ry;
    }
  }
  if (iterations) *iterations = iteration - 1;
  if (sig_out) *sig_out = sig;
  return DUGMA_OK;
}

// ---- registration on the convolved energy ----------------------------------

namespace dugma {

// Pose after a step s in scaled parameters (rotation part times L):
// R <- exp([s_w / L]x) R (2-D: Rot(s_0 / L) R), t <- t + s_t.
static void energy_step(int D, const double *R, const double *t, const double *s, double L,
                        double *Rn, double *tn) {
  double p[6];
  const int nr = D == 3 ? 3 : 1, np = D == 3 ? 6 : 3;
  for (int a = 0; a < np; ++a) p[a] = a < nr ? s[a] / L : s[a];
  const double zero[3] = {0, 0, 0};
  double dR[9], dt[3];
  apply(D, R, zero, p, dR, dt);  // dR = exp(..) R, dt = the translation part
  memcpy(Rn, dR, sizeof(double) * D * D);
  for (int a = 0; a < D; ++a) tn[a] = t[a] + dt[a];
}

}  // namespace dugma

// The inverse of the Hessian of f = -log E in the driver's parameters
// q = [w L, dt], from E, its gradient g and Hessian He in (w, dt) (row-major
// 6 x 6): H_f = -He / E + g g^T / E^2, a factor 1/L per rotation index.
// False unless H_f is positive definite (Cholesky pivots above 1e-12 of its
// largest diagonal entry).
static bool newton_inverse_q(double E, const double *g, const double *He, double L, double *Hinv) {
  double H[6][6], dmax = 0.0;
  for (int a = 0; a < 6; ++a)
    for (int b = 0; b < 6; ++b)
      H[a][b] = (-He[6 * a + b] / E + g[a] * g[b] / (E * E)) * (a < 3 ? 1.0 / L : 1.0) * (b < 3 ? 1.0 / L : 1.0);
  for (int a = 0; a < 6; ++a) dmax = std::max(dmax, H[a][a]);
  if (!(dmax > 0.0) || !std::isfinite(dmax)) return false;
  for (int j = 0; j < 6; ++j) {
    double d = H[j][j];
    for (int k = 0; k < j; ++k) d -= H[j][k] * H[j][k];
    if (!(d > 1e-12 * dmax)) return false;
    d = std::sqrt(d);
    H[j][j] = d;
    for (int i = j + 1; i < 6; ++i) {
      double v = H[i][j];
      for (int k = 0; k < j; ++k) v -= H[i][k] * H[j][k];
      H[i][j] = v / d;
    }
  }
  for (int c = 0; c < 6; ++c) {
    double y[6], x[6];
    for (int i = 0; i < 6; ++i) {
      double v = i == c ? 1.0 : 0.0;
      for (int k = 0; k < i; ++k) v -= H[i][k] * y[k];
      y[i] = v / H[i][i];
    }
    for (int i = 5; i >= 0; --i) {
      double v = y[i];
      for (int k = i + 1; k < 6; ++k) v -= H[k][i] * x[k];
      x[i] = v / H[i][i];
    }
    for (int i = 0; i < 6; ++i) Hinv[6 * i + c] = x[i];
  }
  return true;
}

extern "C" int dugma_register_energy(dugma_ctx *ctx, const double *R0, const double *t0,
                                     const dugma_register_energy_opts *opts, double *R, double *t,
                                     int *iterations, double *E_out) {
  if (!ctx || !R0 || !t0 || !R || !t) return fail(DUGMA_ERR_ARG, "dugma_register_energy: null argument");
  dugma_register_energy_opts op = {200, 1e-7, 1e-14, DUGMA_COV_AUTO};
  if (opts) op = *opts;
  if (op.max_iter < 0 || !(op.tol_rot >= 0) || !(op.tol_trans >= 0))
    return fail(DUGMA_ERR_ARG, "dugma_register_energy: max_iter and tolerances must be >= 0");
  const int D = ctx->D, DD = D * D, np = D == 3 ? 6 : 3, nr = D == 3 ? 3 : 1;
  if (ctx->nS == 0 || ctx->nT == 0)
    return fail(DUGMA_ERR_STATE, "dugma_register_energy: no valid target or source point");
  // Length scales from the kept source points (host copy, scaled units):
  // L, the RMS radius about their centroid; w0, the per-axis kernel width
  // (source variance counted twice for the target's).
  double L = 0.0, w0 = 0.0;
  {
    int rc0 = refresh_host_sources(ctx);
    if (rc0) return rc0;
    const double *src = ctx->hSsrc.data();
    const size_t n = ctx->nS, st = (size_t)(D + DD);
    double c[3] = {0, 0, 0}, var = 0.0;
    for (size_t k = 0; k < n; ++k)
      for (int a = 0; a < D; ++a) c[a] += src[k * st + a] / (double)n;
    for (size_t k = 0; k < n; ++k) {
      for (int a = 0; a < D; ++a) {
        const double d = src[k * st + a] - c[a];
        L += d * d / (double)n;
        var += src[k * st + D + a * (D + 1)] / (double)(n * D);
      }
    }
    L = std::sqrt(L) / ctx->scale;
    w0 = std::sqrt(2.0 * std::max(var, 0.0)) / ctx->scale;
  }
  if (!(L > 0) || !std::isfinite(L)) L = 1.0;
  if (!(w0 > 0) || !std::isfinite(w0)) w0 = 1e-3 * L;
  const double step_max = 0.25 * L;

  // 3-D: each pass also sums the pose Hessian (dugma_energy_grad_hess), and
  // an accepted pose takes the Newton direction where -log E's Hessian is
  // positive definite; else, and in 2-D, BFGS.  DUGMA_REG_NEWTON=0 in the
  // environment keeps BFGS alone (for comparisons).
  const char *nv = getenv("DUGMA_REG_NEWTON");
  const bool newton = D == 3 && !(nv && nv[0] == '0');
  auto eval = [&](const double *Rx, const double *tx, double *Ex, double *gx, double *Hx) -> int {
    return newton ? dugma_energy_grad_hess(ctx, Rx, tx, op.cov_mode, Ex, gx, Hx)
                  : dugma_energy_grad(ctx, Rx, tx, op.cov_mode, Ex, gx);
  };
  double cR[9], ct[3], E, g[6], He[36];
  memcpy(cR, R0, sizeof(double) * DD);
  memcpy(ct, t0, sizeof(double) * D);
  int rc = eval(cR, ct, &E, g, He);
  if (rc) return rc;
  if (!(E > 0) || !std::isfinite(E))
    return fail(DUGMA_ERR_STATE, "dugma_register_energy: the energy is zero at the start pose");
  // gradient of f = -log E in the scaled parameters
  auto grad_q = [&](double e, const double *gp, double *gq) {
    for (int a = 0; a < np; ++a) gq[a] = a < nr ? -gp[a] / (e * L) : -gp[a] / e;
  };
  double gq[6], H[36];
  grad_q(E, g, gq);
  auto reset = [&]() {
    for (int a = 0; a < np * np; ++a) H[a] = (a % (np + 1) == 0) ? 1.0 : 0.0;
  };
  reset();
  bool scaled = false;
  if (newton && newton_inverse_q(E, g, He, L, H)) scaled = true;
  int it = 0;
  while (it < op.max_iter) {
    double d[6], slope = 0.0, dn = 0.0;
    for (int a = 0; a < np; ++a) {
      double v = 0.0;
      for (int b = 0; b < np; ++b) v -= H[a * np + b] * gq[b];
      d[a] = v;
      slope += gq[a] * v;
    }
    if (!(slope < 0.0)) {  // not a descent direction: restart from the gradient
      reset();
      scaled = false;
      slope = 0.0;
      for (int a = 0; a < np; ++a) {
        d[a] = -gq[a];
        slope -= gq[a] * gq[a];
      }
      if (!(slope < 0.0)) break;  // zero gradient
    }
    for (int a = 0; a < np; ++a) dn += d[a] * d[a];
    dn = std::sqrt(dn);
    double alpha = scaled ? 1.0 : w0 / dn;
    if (alpha * dn > step_max) alpha = step_max / dn;
    if (scaled) {
      // the quasi-Newton step is already below the tolerances: converged at
      // the current pose, without a pass to evaluate that step
      double pa = 0.0, pt = 0.0;
      for (int a = 0; a < np; ++a) {
        const double v = alpha * d[a];
        if (a < nr)
          pa += (v / L) * (v / L);
        else
          pt += v * v;
      }
      if (std::sqrt(pa) < op.tol_rot && pt < op.tol_trans) break;
    }
    const double f = -std::log(E);
    double s[6], Rn[9], tn[3], En = 0.0, gn[6], Hen[36];
    bool ok = false;
    for (int k = 0; k < 40 && !ok; ++k) {
      for (int a = 0; a < np; ++a) s[a] = alpha * d[a];
      energy_step(D, cR, ct, s, L, Rn, tn);
      rc = eval(Rn, tn, &En, gn, Hen);
      if (rc) return rc;
      if (En > 0 && std::isfinite(En) && -std::log(En) <= f + 1e-4 * alpha * slope)
        ok = true;
      else
        alpha *= 0.5;
    }
    if (!ok) break;  // no decrease along the direction: stationary to working precision
    ++it;
    double gqn[6], y[6], sy = 0.0, yy = 0.0, ss = 0.0;
    grad_q(En, gn, gqn);
    for (int a = 0; a < np; ++a) {
      y[a] = gqn[a] - gq[a];
      sy += s[a] * y[a];
      yy += y[a] * y[a];
      ss += s[a] * s[a];
    }
    if (sy > 1e-12 * std::sqrt(ss * yy)) {
      if (!scaled) {  // first curvature pair: scale the initial inverse Hessian
        for (int a = 0; a < np * np; ++a) H[a] *= sy / yy;
        scaled = true;
      }
      // H <- (I - r s y^T) H (I - r y s^T) + r s s^T, r = 1 / s^T y
      const double r = 1.0 / sy;
      double Hy[6], yHy = 0.0;
      for (int a = 0; a < np; ++a) {
        double v = 0.0;
        for (int b = 0; b < np; ++b) v += H[a * np + b] * y[b];
        Hy[a] = v;
        yHy += y[a] * v;
      }
      for (int a = 0; a < np; ++a)
        for (int b = 0; b < np; ++b)
          H[a * np + b] += -r * (Hy[a] * s[b] + s[a] * Hy[b]) + (r * r * yHy + r) * s[a] * s[b];
    }
    double Hn[36];
    if (newton && newton_inverse_q(En, gn, Hen, L, Hn)) {
      memcpy(H, Hn, sizeof(H));
      scaled = true;
    }
    memcpy(cR, Rn, sizeof(double) * DD);
    memcpy(ct, tn, sizeof(double) * D);
    E = En;
    memcpy(gq, gqn, sizeof(gq));
    double ang = 0.0, tt = 0.0;
    for (int a = 0; a < np; ++a) {
      if (a < nr)
        ang += (s[a] / L) * (s[a] / L);
      else
        tt += s[a] * s[a];
    }
    if (std::sqrt(ang) < op.tol_rot && tt < op.tol_trans) break;
  }
  memcpy(R, cR, sizeof(double) * DD);
  memcpy(t, ct, sizeof(double) * D);
  if (iterations) *iterations = it;
  if (E_out) *E_out = E;
  return DUGMA_OK;
}
