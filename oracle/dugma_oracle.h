/*
 * dugma_oracle.h -- CPU restatement of the DUGMA registration hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the oracle the HIP path is
 * tested against: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product (libdugma_hip.so) never links or calls it.
 *
 * Conventions shared by every function:
 *  - D is 2 or 3.  Point arrays are point-major (x,y[,z] per point), which
 *    is the memory order of a MATLAB D-by-N column-major matrix.
 *  - Covariance arrays hold one D-by-D column-major block per point, which is
 *    the memory order of the reference's D-by-(D*N) X_C / Y_C matrices
 *    (dataset_all/Read Me:7-8, init.cu:63-71).
 *  - R is D-by-D column-major (MATLAB), t has D entries.
 *  - "target" = reference X (fixed, N points), "source" = reference Y
 *    (moving, M points); the model is X ~ R*Y + t (DUGMA.m:4-7).
 */
#ifndef DUGMA_ORACLE_H
#define DUGMA_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* init.cu:22-190 (detX_gpu/detY_gpu) after init.cu:192-234 (d2f_*):
 * FP64 -> FP32, adjugate inverse and g = sqrtf(1/det) in FP32 with the
 * reference's operation order; det <= 0 or NaN -> inverse 100000 in every
 * entry, g = 0.  Division is IEEE here (the reference builds with
 * -use_fast_math).  Returns the number of fallback points. */
int dor_prep_inverse(int D, int n, const double *C, float *inv, float *g);

/* update_sig.cu:247-252: m_j = fl32(R*y_j + t) computed in FP64 in the
 * order r_a1*y1 + r_a2*y2 (+ r_a3*y3) + t_a, no fused multiply-add. */
void dor_transform_points(int D, int M, const float *Yf, const double *R,
                          const double *t, float *out);

/* update_sig.cu:238-269, 79-91, 272-315, 385 (quantised 1-NN sigma).
 * q[i] = min(4e9, min_j rn_u32(fl32(fl32(|n_i - m_j|^2) / D) * 1e6)),
 * sig = fl32(tree_sum_i(fl32(q_i) / 1e6f)) / N with the reference's pairwise
 * tree order.  q may be NULL. */
float dor_update_sig(int D, int N, const float *Xf, int M, const float *Yf,
                     const double *R, const double *t, unsigned *q);

/* Float pairwise tree sum in the order of update_sig.cu:291-307 (n_c = 1). */
float dor_tree_sum_f32(int n, const float *v);

/* Exact k-nearest neighbours (k <= 64) of each query among candidates.
 * Squared distance in FP32, fixed order: fmaf(dz,dz,fmaf(dy,dy,dx*dx))
 * (2-D: fmaf(dy,dy,dx*dx)), dx = q.x - p.x.  Ranking is lexicographic on
 * (distance, index): ties go to the lowest index.  self_exclude skips j == i.
 * Missing neighbours are idx -1, d2 = +inf. */
void dor_knn(int D, int nq, const float *Q, int np, const float *P, int k,
             int self_exclude, int *idx, float *d2);

/* Per-point sigma within one cloud P (np points): for each query point
 * qidx[i] (i < nq; qidx NULL: point i), its k <= 64 nearest other points of
 * P ranked as dor_knn ranks them (self excluded), and
 *   sigma[i] = (sqrtf(d2_1) + ... + sqrtf(d2_k)) / k   in FP32, nearest first
 * (+inf when P has fewer than k other points). */
void dor_knn_sigma(int D, int nq, const int *qidx, int np, const float *P,
                   int k, float *sigma);

/* Per-point surface covariance from the same k <= 64 nearest neighbours
 * (self excluded, dor_knn's ranking), in FP32, nearest first:
 *   sigma[i] as dor_knn_sigma;
 *   mu_a = (p_1a + ... + p_ka) / k,
 *   c_ab = ((p_1a - mu_a)(p_1b - mu_b) + ... + (p_ka - mu_a)(p_kb - mu_b)) / k,
 *   C_ab = scale c_ab + (a == b ? floor (sigma sigma) : 0),
 * cov[i] column-major D x D; NaN when P has fewer than k other points. */
void dor_knn_cov(int D, int nq, const int *qidx, int np, const float *P, int k,
                 float scale, float floor, float *sigma, float *cov);

/* Convolved-uncertainty energy (new; no reference function computes it):
 *   E = sum_i sum_j N(y_i - (R x_j + t); Sy_i + R Sx_j R^T)
 * in FP64.  Gradient outputs may be NULL:
 *   gpose : 3-D [dE/dw (3), dE/dt (3)] for the left perturbation
 *           R <- exp([w]x) R; 2-D [dE/dtheta, dE/dtx, dE/dty]
 *   gY, gSY : dE/dy_i (D) and dE/dSy_i (D*D symmetric, entries treated as
 *             independent variables)
 *   gX, gSX : dE/dx_j and dE/dSx_j (in the source's own frame)
 *   eT : per-target energy sum_j N(...)
 * Returns E. */
double dor_energy(int D, int N, const double *Y, const double *SY, int M,
                  const double *X, const double *SX, const double *R,
                  const double *t, double *gpose, double *gY, double *gSY,
                  double *gX, double *gSX, double *eT);

/* Central finite differences of dor_energy over the pose (same layout as
 * gpose), step h; 2*(#pose dofs) + 1 evaluations.  Returns E at the centre. */
double dor_energy_fd_pose(int D, int N, const double *Y, const double *SY,
                          int M, const double *X, const double *SX,
                          const double *R, const double *t, double h,
                          double *gpose);

/* compi.cu:632-3407 (3-D) / 96-460 (2-D): coefficient vector of the
 * surrogate polynomial F(dR,dt) = sum_ij P_ij [u^T A'_j u + d^T B_i d],
 * d = n_i - dR m'_j - dt, u = dR^T d, expanded pair by pair in FP64 from the
 * reference's FP32 device state (points, prepped inverses, g) and written in
 * energy3D.m:15 / energy2D.m:7 order (637 / 45 values).
 * Returns the number of expanded monomials missing from that list (0). */
int dor_compi(int D, int N, const float *Xf, const float *XCinv,
              const float *gX, int M, const float *Yf, const float *YCinv,
              const float *gY, const double *R, const double *t, double sig,
              double *coef);

/* The same coefficient vector from moments (FP64 throughout): for each
 * source j the P_ij-weighted target moments sum_i P, P n, P n n^T, P B,
 * P n^T B, P B n, P n^T B n (P_ij as dor_compi computes it), then the
 * polynomial F_j(dR, dt) = sum_i P_ij [u^T A'_j u + d^T B_i d] expanded
 * from them, the sources taken in order.  O(N M) pair work instead of
 * O(N M x 637). */
int dor_compi_moments(int D, int N, const float *Xf, const float *XCinv,
                      const float *gX, int M, const float *Yf,
                      const float *YCinv, const float *gY, const double *R,
                      const double *t, double sig, double *coef);

/* Brute-force value of the same surrogate at dT (3-D: t1 t2 t3 r11 .. r33
 * row-major dR; 2-D: t1 t2 r11 r21). */
double dor_surrogate_eval(int D, int N, const float *Xf, const float *XCinv,
                          const float *gX, int M, const float *Yf,
                          const float *YCinv, const float *gY, const double *R,
                          const double *t, double sig, const double *dT);

/* energy3D.m:15-16 / energy2D.m:7-8: F = coef . B(dT). */
double dor_poly_eval(int D, const double *coef, const double *dT);

/* fcontr3D.m:14-26 / fcontr2D.m:6-7: equality constraints (10 / 1). */
int dor_constraints(int D, const double *dT, double *ceq);

/* Pair-weight P_ij of compi.cu:747-772 (3-D) / 193-211 (2-D) for one pair,
 * exposed for tests. */
double dor_pair_weight(int D, const float *n, const float *binv, float gx,
                       const float *m, const float *ainv, float gy,
                       const double *R, const double *t, double sig);

/* Number of threads the OpenMP loops use (1 when built without OpenMP). */
int dor_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
