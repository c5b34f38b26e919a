/*
 * dugma.h -- C ABI of the MI355X (gfx950) DUGMA registration hot path.
 *
 * Plain pointers and sizes only.  Every entry point returns a status code
 * (DUGMA_OK = 0, negative on error) and never terminates the process;
 * dugma_last_error() returns the message of the calling thread's last error.
 *
 * Layout conventions (those of the reference's MEX gateways):
 *  - a point cloud is D-by-N column-major double, i.e. x,y[,z] of point 0,
 *    then point 1, ... (init.cpp:43-60, mxGetPr of X);
 *  - covariances are D-by-(D*N) column-major double: the D-by-D block of
 *    point k is contiguous, column-major (dataset_all/Read Me:7-8);
 *  - R is D-by-D column-major double, t has D doubles;
 *  - "target" is the reference's fixed cloud X (N points), "source" is its
 *    moving cloud Y (M points); the model is X ~ R*Y + t (DUGMA.m:4-7).
 *
 * Threading: a context may be used by one host thread at a time; calls on
 * one context are serialised on the context's own HIP stream.
 */
#ifndef DUGMA_H
#define DUGMA_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct dugma_ctx dugma_ctx;

enum {
  DUGMA_OK = 0,
  DUGMA_ERR_ARG = -1,     /* bad argument (null pointer, D not 2/3, sizes) */
  DUGMA_ERR_DEVICE = -2,  /* HIP runtime error or no device */
  DUGMA_ERR_ALLOC = -3,   /* device or pinned allocation failed */
  DUGMA_ERR_SIZE = -4,    /* point count beyond the supported maximum */
  DUGMA_ERR_STATE = -5    /* mismatch with the context (N, M or D) */
};

/* Largest point count per cloud accepted by dugma_init (the reference caps
 * at 3,000,000 and exits: compi.cu:535-539, update_sig.cu:341-345). */
#define DUGMA_MAX_POINTS 16777216

/* ---- reference MEX entry points ---------------------------------------- */

/* [ok, h0..h5] = init(X, X_C, Y, Y_C)   (init.cpp:36-83, init.cu:249-348).
 * Uploads both clouds, keeps the FP32 points, FP32 inverse covariances and
 * g = sqrtf(1/det C) of the reference (det <= 0 or NaN: inverse 100000 in
 * every entry, g = 0), plus the data the convolved energy needs.  Writes
 * the new context to *out; *ok (may be NULL) receives 1.0 like init's first
 * output. */
int dugma_init(const double *X, const double *X_C, const double *Y,
               const double *Y_C, int N, int M, int D, dugma_ctx **out,
               double *ok);

/* [sig, h0..h5] = update_sig(h0..h5, R, t, N, M, D)
 * (update_sig.cpp:8-66, update_sig.cu:319-394).  N, M, D are passed as
 * doubles like the MEX arguments and must match the context.
 * sig = (1/N) sum_i min_j |X_i - fl32(R Y_j + t)|^2 / D, each minimum
 * quantised to 1e-6 (round to nearest, capped at 4000), summed in FP32 in
 * the reference's pairwise order. */
int dugma_update_sig(dugma_ctx *ctx, const double *R, const double *t,
                     double N, double M, double D, double *sig);

/* [coef, h0..h5] = compi(h0..h5, R, t, N, M, D, sig)
 * (compi.cpp:8-74, compi.cu:511-628).  Writes the 637 (3-D) or 45 (2-D)
 * coefficients of the surrogate polynomial in energy3D.m:15 / energy2D.m:7
 * monomial order. */
int dugma_compi(dugma_ctx *ctx, const double *R, const double *t, double N,
                double M, double D, double sig, double *coef);

/* ok = fr_m(h0..h5)   (fr_m.cpp:8-35, fr_m.cu:14-31).  Frees the context's
 * device memory.  Unlike the reference it does not reset the device. */
int dugma_fr_m(dugma_ctx *ctx, double *ok);

/* ---- the surrogate polynomial (energy3D.m / energy2D.m) ----------------- */

/* F(dT) = coef . B(dT) over the 637 (3-D) or 45 (2-D) monomials of
 * energy3D.m:15 / energy2D.m:7, dT = [t1 t2 t3 r11 .. r33] (dR row-major) or
 * [t1 t2 r11 r21].  grad (12 or 4 doubles, may be NULL) receives dF/ddT:
 * the objective gradient fmincon would otherwise take by finite differences.
 * Returns NaN for bad arguments. */
double dugma_poly_eval(int D, const double *coef, const double *dT, double *grad);

/* Minimise F over dR in SO(D) (fcontr3D.m / fcontr2D.m constraints, kept by
 * construction) and dt, from dT = identity: damped Newton on a left
 * rotation perturbation and dt.  Replaces fmincon (DUGMA.m:65,73); writes the
 * minimiser to dT and the value to *F. */
int dugma_poly_minimize(int D, const double *coef, double *dT, double *F);

/* ---- the registration loop (DUGMA.m) ------------------------------------ */

typedef struct {
  int max_iter;        /* Max_Iteration */
  double acc_rot;      /* accuracy_rotation (radians of ZYX Euler angles; degrees in 2-D) */
  double acc_trans;    /* accuracy_translation (squared step) */
  double acc_sig;      /* accuracy_sig */
  int verbose;         /* print one line per iteration like DUGMA.m:109 */
} dugma_register_opts;

/* [R, t] = DUGMA(X, Y, X_C, Y_C, R0, t0, ...)   (DUGMA.m:19-137): normalise
 * the points (normalize.m), init, update_sig, then compi -> minimise the
 * surrogate -> R <- dR R, t <- t + dt -> Euler-angle convergence ->
 * update_sig, and denormalise.  Layouts as dugma_init; R0, R column-major. */
int dugma_register(const double *X, const double *X_C, const double *Y,
                   const double *Y_C, int N, int M, int D, const double *R0,
                   const double *t0, const dugma_register_opts *opts,
                   double *R, double *t, int *iterations, double *sig);

/* ---- convolved-uncertainty energy (no reference counterpart) ----------- */

/* Evaluation modes for the energy entry points. */
enum {
  DUGMA_COV_AUTO = 0, /* isotropic kernels when every covariance is s^2 I */
  DUGMA_COV_FULL = 1, /* full-covariance kernels always */
  DUGMA_COV_ISO = 2   /* isotropic kernels; error unless all are s^2 I */
};

/* E(R,t) = sum_i sum_j N(X_i - (R Y_j + t); CX_i + R CY_j R^T), per-pair
 * FP32 with FP64 accumulation.  Points whose coordinates or covariance are
 * not finite, or whose covariance is not positive definite, are left out. */
int dugma_energy(dugma_ctx *ctx, const double *R, const double *t, int cov_mode,
                 double *E);

/* Energy and its gradient with respect to the pose from one fused
 * forward+adjoint pass (the optimiser's gradient callback).
 * gpose (3-D): [dE/dw1 dE/dw2 dE/dw3 dE/dt1 dE/dt2 dE/dt3] for the left
 * perturbation R <- exp([w]x) R, t <- t + dt; (2-D): [dE/dtheta dE/dt1 dE/dt2]. */
int dugma_energy_grad(dugma_ctx *ctx, const double *R, const double *t,
                      int cov_mode, double *E, double *gpose);

/* dugma_energy_grad without any allocation: no reference counterpart; for
 * callers that cannot allocate (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, kernels and
 * results (bit for bit) as dugma_energy_grad; the pair kernel's
 * per-workgroup partials go to a file-scope device array of 2^18 records of
 * 8 doubles (16 MiB, held by each GPU the library runs on until the process
 * ends) instead of the context's grow-only buffer.  A launch that needs more
 * partial records (beyond about 1.7e11 pairs at the default wave counts)
 * fails with DUGMA_ERR_SIZE.  All calls on one GPU share that GPU's array:
 * calls that overlap on one GPU (several threads or streams) race. */
int dugma_energy_grad_noalloc(dugma_ctx *ctx, const double *R, const double *t,
                              int cov_mode, double *E, double *gpose);

/* Registration on the convolved energy itself: quasi-Newton (BFGS) descent
 * of -log E(R, t) over left pose increments R <- exp([w]x) R (2-D: a
 * rotation by theta on the left), t <- t + dt, re-centred on every accepted
 * pose, each trial pose evaluated by one fused energy + pose-gradient pass
 * (dugma_energy_grad); backtracking (Armijo) line search.  This is the
 * "SE(3) step on the convolved energy" beside the surrogate minimiser of
 * dugma_register; it has no reference counterpart (DUGMA.m minimises the
 * compi surrogate, DUGMA.m:57-123).  Rotation parameters are scaled by the
 * RMS radius of the source cloud so that both blocks are lengths; the first
 * step moves by about one kernel width.  Stops when a step rotates by less
 * than tol_rot (rad) and translates by less than sqrt(tol_trans) -- an
 * accepted step, or the next quasi-Newton step once the inverse Hessian has
 * been scaled (then without evaluating it) -- when no decrease is found
 * along the search direction, or after max_iter accepted steps.
 * iterations = accepted steps; E = energy at the returned pose.
 * opts may be NULL (max_iter 200, tol_rot 1e-7, tol_trans 1e-14, AUTO). */
typedef struct {
  int max_iter;
  double tol_rot;
  double tol_trans;
  int cov_mode;
} dugma_register_energy_opts;
int dugma_register_energy(dugma_ctx *ctx, const double *R0, const double *t0,
                          const dugma_register_energy_opts *opts, double *R, double *t,
                          int *iterations, double *E);

/* Energy with per-point gradients (no N x M storage: one target-major and
 * one source-major pass).  Any output may be NULL.
 *  gX  : N x D      dE/dX_i            gY  : M x D      dE/dY_j
 *  gXC : N x D x D  dE/dCX_i           gYC : M x D x D  dE/dCY_j
 * (covariance gradients column-major per point, entries treated as
 * independent variables; zero for left-out points).
 *  gdT : 12 (3-D) or 4 (2-D) doubles, dE with respect to the fmincon
 *        variables dT of energy3D.m / energy2D.m at dT = identity, where the
 *        pose is updated as R <- dR R, t <- t + dt (DUGMA.m:67-78).
 * These passes always run the full-covariance kernels (cov_mode is only
 * checked): E and gdT describe source covariances R CY R^T also when R is
 * not a rotation, as for dugma_energy with DUGMA_COV_FULL.  The per-point
 * sums come back through a pinned buffer that the context keeps, and up to
 * 8 host threads turn them into the outputs. */
int dugma_energy_grad_points(dugma_ctx *ctx, const double *R, const double *t,
                             int cov_mode, double *E, double *gpose,
                             double *gX, double *gXC, double *gY, double *gYC,
                             double *gdT);

/* 3-D: dugma_energy_grad plus the 6 x 6 Hessian H of E in the same pose
 * parameters (row-major, symmetric), from one pass that also sums, per
 * transformed source x', J^T B J + the second-order turn of x' - t (J =
 * [-[x'-t]x | I], B = sum_i e (S^-1 d d^T S^-1 - S^-1)).  Exact for isotropic
 * covariances; with full ones the turn of the sources' covariances is left
 * out of the second order (the translation block stays exact).  For Newton
 * steps (dugma_register_energy).  No reference counterpart. */
int dugma_energy_grad_hess(dugma_ctx *ctx, const double *R, const double *t,
                           int cov_mode, double *E, double *gpose, double *H);

/* "Plain" gradient: central finite differences of dugma_energy over the pose
 * (2*dofs + 1 forward launches), step h.  Same gpose layout. */
int dugma_energy_grad_fd(dugma_ctx *ctx, const double *R, const double *t,
                         int cov_mode, double h, double *E, double *gpose);

/* Choose the launch shape (points per lane, grid depth) of the energy-only
 * and fused pose passes for this context's sizes by timing candidates on the
 * device; later calls with the same covariance mode use the fastest.
 * dugma_ctx_plan reports the chosen shape of the fused pass (0 waves: not
 * tuned yet). */
int dugma_autotune(dugma_ctx *ctx, int cov_mode);
int dugma_ctx_plan(const dugma_ctx *ctx, int cov_mode, int *lane_points,
                   int *target_waves);

/* ---- k nearest neighbours ---------------------------------------------- */

enum {
  DUGMA_KNN_TARGET_SELF = 0, /* queries X, candidates X, self excluded */
  DUGMA_KNN_SOURCE_SELF = 1, /* queries Y, candidates Y, self excluded */
  DUGMA_KNN_TARGET_TO_SOURCE = 2 /* queries X, candidates fl32(R Y_j + t) */
};

/* k <= 32 nearest neighbours with FP32 squared distances computed as
 * fmaf(dz,dz,fmaf(dy,dy,dx*dx)) on the FP32 points of init, ranked by
 * (distance, index) so that ties go to the lowest index.  idx: nq x k int,
 * d2: nq x k float (either may be NULL); missing neighbours are -1 / +inf;
 * points with a non-finite coordinate are never neighbours.  R, t are used
 * by DUGMA_KNN_TARGET_TO_SOURCE only.  Small problems compare all pairs;
 * from 2^24 query x candidate pairs on, the candidates are binned into a
 * uniform grid and searched ring by ring with an exact stopping bound: the
 * results are the same, bit for bit.  The lists reach idx / d2 through a
 * pinned buffer that the context keeps. */
int dugma_knn(dugma_ctx *ctx, int which, int k, const double *R,
              const double *t, int *idx, float *d2);

/* ---- per-point sigma from the k nearest neighbours ----------------------- */

enum {
  DUGMA_SIGMA_TARGET = 1, /* bit mask of the clouds dugma_update_sigma_knn updates */
  DUGMA_SIGMA_SOURCE = 2
};

/* Per-point isotropic covariances from the k <= 32 nearest neighbours within
 * each selected cloud, on the device (extends the nearest-distance pass of
 * update_sig.cu:181-269, which DUGMA.m:47,105 refreshes every iteration;
 * no reference function computes per-point sigma).  The cloud is the points
 * the energy keeps (finite coordinates and a positive-definite covariance at
 * init); neighbours are ranked as dugma_knn ranks them (FP32 squared
 * distances of init's FP32 points, (distance, index) order, self excluded).
 *   sigma_i = (sqrtf(d2_1) + sqrtf(d2_2) + ... + sqrtf(d2_k)) / k
 * in FP32, nearest first, and the point's covariance becomes fl32(sigma_i^2) I
 * in every device buffer the context holds, exactly as dugma_init would
 * have stored it: init's inverse and g (the reference handles), and the
 * energy state.  No host round trip of point data.  A point whose
 * fl32(sigma_i^2) is zero or not finite (coincident neighbours, fewer than
 * k other points) keeps its covariance; *unchanged (may be NULL) counts
 * them.  sigma_target (N floats) / sigma_source (M floats), either may be
 * NULL, receive sigma_i of the kept points and NaN for the others.  A
 * context whose covariances all become isotropic switches to the isotropic
 * kernels (DUGMA_COV_AUTO). */
int dugma_update_sigma_knn(dugma_ctx *ctx, int which, int k, float *sigma_target,
                           float *sigma_source, int *unchanged);

/* Per-point surface covariances from the same k (2..32) nearest neighbours,
 * on the device (the sample-covariance update; extends update_sig.cu:181-269
 * like dugma_update_sigma_knn, no reference function computes it): with
 * sigma_i as above and, in FP32, nearest first,
 *   mu = (p_1 + ... + p_k) / k,
 *   c_ab = ((p_1a - mu_a)(p_1b - mu_b) + ... + (p_ka - mu_a)(p_kb - mu_b)) / k,
 *   C_ab = fl32(scale) c_ab + (a == b ? fl32(floor) (sigma_i sigma_i) : 0),
 * written everywhere the context holds the point's covariance, exactly as
 * dugma_init would have stored that FP32 covariance (init's inverse and g,
 * and the energy state), with no host round trip.  Flat along the surface,
 * these covariances make the energy of two independently sampled clouds peak
 * at the true pose.  A point whose C is not positive definite, or with fewer
 * than k other points, keeps its covariance; *unchanged (may be NULL) counts
 * them.  sigma_target / sigma_source (may be NULL) as dugma_update_sigma_knn.
 * The context then uses the full-covariance kernels. */
int dugma_update_cov_knn(dugma_ctx *ctx, int which, int k, double scale, double floor,
                         float *sigma_target, float *sigma_source, int *unchanged);

/* ---- batched registration (no reference counterpart) -------------------- */

/* Many independent 3-D rigid registrations of uneven sizes at once (the
 * DUGMA.m:57-123 loop for each trial, with a closed-form pose step).  Each
 * maximises the convolved energy with isotropic covariances (every X_C / Y_C
 * block must be s I) by majorise-minimise steps (weighted Procrustes, FP64).
 * The pair variance is g (s_i + s_j) + a: with the sigma update, g = sig /
 * sbar where sig is DUGMA's mean squared nearest distance / 3 measured in
 * each pass (update_sig.cu:238-260, applied as in compi.cu:642-661) and sbar
 * the trial's mean target plus mean source variance; a is an optional
 * annealed extra variance, a -> 0. */
typedef struct dugma_batch dugma_batch;

typedef struct {
  int max_iter;              /* iteration limit per trial */
  int anneal;                /* 1: start at a = mean squared target radius / 4 */
  double anneal_decay;       /* a <- a * decay per iteration (0: 0.85) */
  double tol_rot;            /* converged when a = 0 and the step is below these */
  double tol_trans;
  int use_graph;             /* replay one captured iteration (hipGraph) */
  int one_trial_per_launch;  /* baseline: trials one after another, host check per step */
  int check_every;           /* graph mode: host checks for completion every n steps */
  int grid;                  /* workgroups of the persistent pair kernel (0: from the launch shape) */
  int sigma_update;          /* 1: DUGMA sigma update (one extra measuring pass first) */
  double relax;              /* > 1: over-relax the pose step once annealing has ended
                                (rotation angle and centroid move times relax) */
  int coarse_iters;          /* first iterations on every 4th target and source point
                                (no convergence test there); then the full clouds */
  int refine_iters;          /* > 0: then at most this many accepted steps per trial on
                                -log E with the full covariances of
                                dugma_batch_set_covariances (one fused energy + pose-gradient
                                pass per trial pose, which also sums an approximate pose
                                Hessian: Newton directions where it is positive definite, else
                                BFGS; stops when an accepted step, or the next step once the
                                inverse Hessian is scaled, is below tol_rot / tol_trans) */
  int refine_coarse_iters;   /* with refine_iters > 0: first at most this many steps on
                                every 4th target and source point, with their covariances
                                times 4 (the sampling density is a quarter, so a surface
                                covariance spans twice the distance), then the full clouds
                                from the pose reached and the inverse Hessian / 4 */
  int refine_mid_iters;      /* with refine_iters > 0: after the coarse level (if any), at most
                                this many steps on every 2nd point, covariances times 2 */
} dugma_batch_opts;

/* Arrays of ntrials pointers, each in the layout of dugma_init. */
int dugma_batch_create(int ntrials, int D, const int *N, const int *M,
                       const double *const *X, const double *const *X_C,
                       const double *const *Y, const double *const *Y_C,
                       dugma_batch **out);
/* The same with one variance per point (covariance s I): sX[k] has N[k]
 * entries, sY[k] has M[k]. */
int dugma_batch_create_iso(int ntrials, const int *N, const int *M,
                           const double *const *X, const double *const *sX,
                           const double *const *Y, const double *const *sY,
                           dugma_batch **out);
/* Full per-point covariances for the refinement phase (refine_iters), in
 * the layout of dugma_init (X_C[k]: 3 x (3 N[k]) column-major blocks); only
 * the points the batch keeps are read, and each of those must be finite and
 * positive definite in FP32.  For independently sampled surfaces, covariances
 * flat along the surface (e.g. the sample covariance of each point's nearest
 * neighbours plus a small isotropic floor) make the energy peak at the true
 * pose where isotropic ones do not.  No reference counterpart (the
 * reference's covariances are the sensor's, DUGMA.m:44). */
int dugma_batch_set_covariances(dugma_batch *b, const double *const *X_C,
                                const double *const *Y_C);
/* R0 / R: ntrials x 9 column-major, t0 / t: ntrials x 3; iters, status
 * (1 converged, 2 iteration limit, 3 no weight), E: ntrials each (any output
 * may be NULL).  E is the energy at the returned pose (isotropic: with the
 * variances of the last iteration, from one extra pass after the
 * iterations).  With refine_iters > 0, R, t, status and E are those of the
 * refinement and iters counts both phases' steps. */
int dugma_batch_register(dugma_batch *b, const double *R0, const double *t0,
                         const dugma_batch_opts *opts, double *R, double *t,
                         int *iters, int *status, double *E);
int dugma_batch_info(const dugma_batch *b, int *ntrials, int *nitems,
                     long long *pairs_per_iteration);
/* Launch shape of the persistent pair kernels: waves per workgroup (1, 2, 4;
 * the target tile is 64 x waves points) and source chunks per target tile
 * for each of 4 size classes of trials (source count < 1000, < 2000, < 4000,
 * more), and resident waves, per level (0: full clouds, 1: every 4th point).
 * dugma_batch_autotune times whole isotropic iterations -- plus, once
 * dugma_batch_set_covariances has run, whole refinement passes, their sum
 * being the figure minimised -- from the poses R0, t0 on the batch's own
 * data, on the full clouds first (waves, 4 kept unless fewer are faster by
 * more than 3%; then each present size class; then the resident waves),
 * then on the coarse level (waves, size classes, resident waves), keeps the
 * fastest shapes
 * and reports the full clouds' one with its time per isotropic iteration
 * (ms) and per refinement pass (refine_ms, 0 without covariances); any
 * output may be NULL.  dugma_batch_set_shape gives every level the same
 * shape, dugma_batch_set_level_shape one level (0, 1, or 2: every 2nd
 * point, refinement only) its own; dugma_batch_get_shape reports a level's
 * (chunks: 4 ints).  The shapes change how the pair sums are grouped, so
 * two batches give bit-identical results only with the same shapes. */
int dugma_batch_autotune(dugma_batch *b, const double *R0, const double *t0,
                         int sigma_update, int *waves, int *chunks,
                         int *grid_waves, double *ms, double *refine_ms);
int dugma_batch_set_shape(dugma_batch *b, int waves, const int *chunks,
                          int grid_waves);
int dugma_batch_set_level_shape(dugma_batch *b, int level, int waves,
                                const int *chunks, int grid_waves);
int dugma_batch_get_shape(const dugma_batch *b, int level, int *waves,
                          int *chunks, int *grid_waves);
int dugma_batch_free(dugma_batch *b);

/* ---- the six reference handles ------------------------------------------ */
/* Device pointers of the state the reference keeps behind its six MEX handles
 * (init.cu:326-331, in that order): X, inverse X_C, Y, inverse Y_C, g_X, g_Y
 * (FP32; sizes N*D, N*D*D, M*D, M*D*D, N, M).  ptrs must hold 6 entries. */
int dugma_ctx_handles(const dugma_ctx *ctx, void **ptrs);
/* Copy handle `which` (0..5, order as above) to host memory `out`. */
int dugma_read_handle(dugma_ctx *ctx, int which, float *out);

/* ---- context queries ---------------------------------------------------- */
int dugma_ctx_dims(const dugma_ctx *ctx, int *N, int *M, int *D);
/* Numbers of points kept for the energy (after leaving out invalid ones) and
 * whether every kept covariance is isotropic. */
int dugma_ctx_energy_info(const dugma_ctx *ctx, int *n_target, int *n_source,
                          int *isotropic);
/* The HIP stream (hipStream_t) the context launches on, for event timing. */
void *dugma_ctx_stream(const dugma_ctx *ctx);
/* Kernel timing: when on, HIP events are recorded on the context's stream
 * around every launch of the all-pairs kernels (energy, pose-gradient and
 * per-point passes; a per-point call counts two launches, and each launch
 * then waits for its kernel); dugma_ctx_pair_time returns the summed
 * milliseconds and launch count since timing was switched on. */
int dugma_ctx_set_timing(dugma_ctx *ctx, int on);
int dugma_ctx_pair_time(const dugma_ctx *ctx, double *ms, long long *launches);
const char *dugma_last_error(void);
const char *dugma_version(void);

#ifdef __cplusplus
}
#endif
#endif
