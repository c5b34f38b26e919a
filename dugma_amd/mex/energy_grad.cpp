// energy_grad.cpp -- MEX gateway for an fmincon objective with analytic
// gradient on the convolved energy:
//   [negE, g] = energy_grad(ad0, R, t, dT)
// evaluates -E at the pose (dR R, t + dt) given fmincon's 12 (3-D) / 4 (2-D)
// variables dT (energy3D.m / energy2D.m layout) and returns -dE/ddT when
// nlhs > 1 (for optimset('GradObj','on') / SpecifyObjectiveGradient).
#include <vector>

#include "dugma_mex_handle.hpp"

using namespace dugma_mex;

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
  if (nrhs != 4) mexErrMsgTxt("energy_grad: expected ad0, R, t, dT");
  Handle *h = from_mat(prhs[0]);
  int N, M, D;
  check(dugma_ctx_dims(h->ctx, &N, &M, &D));
  const double *R = mxGetPr(prhs[1]), *t = mxGetPr(prhs[2]), *dT = mxGetPr(prhs[3]);
  double dR[9], dt[3], Rn[9], tn[3];
  if (D == 3) {
    for (int a = 0; a < 3; ++a) {
      dt[a] = dT[a];
      for (int b = 0; b < 3; ++b) dR[a + 3 * b] = dT[3 + 3 * a + b];
    }
  } else {
    dt[0] = dT[0];
    dt[1] = dT[1];
    dR[0] = dT[2];
    dR[1] = dT[3];
    dR[2] = -dT[3];
    dR[3] = dT[2];
  }
  for (int a = 0; a < D; ++a) {
    tn[a] = t[a] + dt[a];
    for (int c = 0; c < D; ++c) {
      double s = 0;
      for (int b = 0; b < D; ++b) s += dR[a + D * b] * R[b + D * c];
      Rn[a + D * c] = s;
    }
  }
  double E = 0.0;
  std::vector<double> g(D == 3 ? 12 : 4, 0.0);
  if (nlhs > 1) {
    // gradient of E at (Rn, tn) with respect to a further dT' at identity,
    // composed with d(dR R)/d(dR) = (.) R (dT enters as dR R).
    std::vector<double> gid(D == 3 ? 12 : 4);
    check(dugma_energy_grad_points(h->ctx, Rn, tn, DUGMA_COV_AUTO, &E, nullptr, nullptr, nullptr, nullptr,
                                   nullptr, gid.data()));
    // With Rn = dR R, a change Delta of dR moves Rn by Delta R = (Delta dR^-1) Rn,
    // so dE/d(dR) = G dR^-T where G = dE/d(left perturbation) at Rn.
    if (D == 3) {
      double G[9], Ginv[9];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) G[a + 3 * b] = gid[3 + 3 * a + b];
      // dR^-T for a near-rotation dR: use the adjugate / det
      const double *m = dR;
      double det = m[0] * (m[4] * m[8] - m[7] * m[5]) - m[3] * (m[1] * m[8] - m[7] * m[2]) +
                   m[6] * (m[1] * m[5] - m[4] * m[2]);
      double inv[9] = {(m[4] * m[8] - m[5] * m[7]) / det, (m[2] * m[7] - m[1] * m[8]) / det,
                       (m[1] * m[5] - m[2] * m[4]) / det, (m[5] * m[6] - m[3] * m[8]) / det,
                       (m[0] * m[8] - m[2] * m[6]) / det, (m[2] * m[3] - m[0] * m[5]) / det,
                       (m[3] * m[7] - m[4] * m[6]) / det, (m[1] * m[6] - m[0] * m[7]) / det,
                       (m[0] * m[4] - m[1] * m[3]) / det};
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          double s = 0;
          for (int c = 0; c < 3; ++c) s += G[a + 3 * c] * inv[b + 3 * c];  // (G inv^T)_ab
          Ginv[a + 3 * b] = s;
        }
      for (int a = 0; a < 3; ++a) g[a] = -gid[a];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) g[3 + 3 * a + b] = -Ginv[a + 3 * b];
    } else {
      // dR = a I + b J (J the quarter turn): dE/d(dR) = G dR^-T = G dR / (a^2 + b^2),
      // and the r11 / r21 combinations of G dR follow from those of G.
      const double a = dT[2], b = dT[3], rho = a * a + b * b;
      g[0] = -gid[0];
      g[1] = -gid[1];
      g[2] = -(a * gid[2] - b * gid[3]) / rho;
      g[3] = -(a * gid[3] + b * gid[2]) / rho;
    }
  } else {
    // full-covariance kernels, like the gradient pass: fmincon also evaluates
    // dR that are not rotations, and only these kernels carry the source
    // covariance through as (dR R) S (dR R)^T (isotropic kernels keep s I)
    check(dugma_energy(h->ctx, Rn, tn, DUGMA_COV_FULL, &E));
  }
  plhs[0] = mxCreateDoubleMatrix(1, 1, mxREAL);
  mxGetPr(plhs[0])[0] = -E;
  if (nlhs > 1) {
    plhs[1] = mxCreateDoubleMatrix((mwSize)g.size(), 1, mxREAL);
    for (size_t k = 0; k < g.size(); ++k) mxGetPr(plhs[1])[k] = g[k];
  }
}
