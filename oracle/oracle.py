"""ctypes bindings of the CPU oracle (liboracle built from dugma_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the dugma_amd package.

Pinning of the functions written here (those of the C library are listed in
dugma_oracle.c):
  dugma_loop   end to end against the reference's own outputs: on whole
               Kinect scenes with the reference driver's settings it lands on
               the pose published in Ours_k_result.mat (scenes 8 and 20 in
               the CPU tests, within 0.01 degrees); one iteration against an
               independent assembly (numpy normalisation, SLSQP under the
               fcontr3D.m constraints, R <- dR R, t <- t + dt, DUGMA.m:122
               denormalisation); noise-free clouds end at the truth up to
               the sigma stopping rule, and started at the truth it returns
               the truth through the denormalisation.  fmincon itself is
               absent: scipy's BFGS on SO(3) stands in for it
  rotm2eul_zyx the textbook Rz(z) Ry(y) Rx(x) construction
  normalize    through dugma_loop (normalize.m:3-26)
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_d = ctypes.POINTER(ctypes.c_double)
_f = ctypes.POINTER(ctypes.c_float)
_i = ctypes.POINTER(ctypes.c_int)
_u = ctypes.POINTER(ctypes.c_uint)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libdugma_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle library not built: run make -C oracle")
        L = ctypes.CDLL(path)
        c_int, c_double = ctypes.c_int, ctypes.c_double
        L.dor_prep_inverse.argtypes = [c_int, c_int, _d, _f, _f]
        L.dor_prep_inverse.restype = c_int
        L.dor_transform_points.argtypes = [c_int, c_int, _f, _d, _d, _f]
        L.dor_update_sig.argtypes = [c_int, c_int, _f, c_int, _f, _d, _d, _u]
        L.dor_update_sig.restype = ctypes.c_float
        L.dor_tree_sum_f32.argtypes = [c_int, _f]
        L.dor_tree_sum_f32.restype = ctypes.c_float
        L.dor_knn.argtypes = [c_int, c_int, _f, c_int, _f, c_int, c_int, _i, _f]
        L.dor_knn_sigma.argtypes = [c_int, c_int, _i, c_int, _f, c_int, _f]
        L.dor_knn_cov.argtypes = [c_int, c_int, _i, c_int, _f, c_int, ctypes.c_float, ctypes.c_float, _f, _f]
        L.dor_energy.argtypes = [c_int, c_int, _d, _d, c_int, _d, _d, _d, _d,
                                 _d, _d, _d, _d, _d, _d]
        L.dor_energy.restype = c_double
        L.dor_energy_fd_pose.argtypes = [c_int, c_int, _d, _d, c_int, _d, _d,
                                         _d, _d, c_double, _d]
        L.dor_energy_fd_pose.restype = c_double
        L.dor_compi.argtypes = [c_int, c_int, _f, _f, _f, c_int, _f, _f, _f,
                                _d, _d, c_double, _d]
        L.dor_compi.restype = c_int
        L.dor_compi_moments.argtypes = [c_int, c_int, _f, _f, _f, c_int, _f, _f, _f,
                                        _d, _d, c_double, _d]
        L.dor_compi_moments.restype = c_int
        L.dor_surrogate_eval.argtypes = [c_int, c_int, _f, _f, _f, c_int, _f,
                                         _f, _f, _d, _d, c_double, _d]
        L.dor_surrogate_eval.restype = c_double
        L.dor_poly_eval.argtypes = [c_int, _d, _d]
        L.dor_poly_eval.restype = c_double
        L.dor_constraints.argtypes = [c_int, _d, _d]
        L.dor_constraints.restype = c_int
        L.dor_pair_weight.argtypes = [c_int, _f, _f, ctypes.c_float, _f, _f,
                                      ctypes.c_float, _d, _d, c_double]
        L.dor_pair_weight.restype = c_double
        L.dor_num_threads.argtypes = []
        L.dor_num_threads.restype = c_int
        _LIB = L
    return _LIB


def _p(a, ptr):
    return None if a is None else a.ctypes.data_as(ptr)


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _c32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _colmajor(R):
    """MATLAB column-major copy of a D-by-D matrix given as a numpy 2-D array."""
    return np.ascontiguousarray(np.asarray(R, dtype=np.float64).T).ravel()


def num_threads():
    return lib().dor_num_threads()


def prep_inverse(C, D):
    """C: (n, D, D) covariances (numpy row-major; symmetric)."""
    C = np.asarray(C, dtype=np.float64)
    n = C.shape[0]
    cm = _c64(np.transpose(C, (0, 2, 1)).reshape(n, D * D))  # column-major per point
    inv = np.empty((n, D * D), np.float32)
    g = np.empty(n, np.float32)
    bad = lib().dor_prep_inverse(D, n, _p(cm, _d), _p(inv, _f), _p(g, _f))
    return inv, g, bad


def transform_points(Yf, R, t):
    Yf = _c32(Yf)
    M, D = Yf.shape
    out = np.empty_like(Yf)
    lib().dor_transform_points(D, M, _p(Yf, _f), _p(_colmajor(R), _d), _p(_c64(t), _d), _p(out, _f))
    return out


def update_sig(Xf, Yf, R, t, with_q=False):
    Xf, Yf = _c32(Xf), _c32(Yf)
    N, D = Xf.shape
    q = np.empty(N, np.uint32)
    s = lib().dor_update_sig(D, N, _p(Xf, _f), Yf.shape[0], _p(Yf, _f),
                             _p(_colmajor(R), _d), _p(_c64(t), _d), _p(q, _u))
    return (float(s), q) if with_q else float(s)


def tree_sum_f32(v):
    v = _c32(v)
    return float(lib().dor_tree_sum_f32(v.size, _p(v, _f)))


def knn(Q, P, k, self_exclude=False):
    Q, P = _c32(Q), _c32(P)
    nq, D = Q.shape
    idx = np.empty((nq, k), np.int32)
    d2 = np.empty((nq, k), np.float32)
    lib().dor_knn(D, nq, _p(Q, _f), P.shape[0], _p(P, _f), k, int(bool(self_exclude)),
                  _p(idx, _i), _p(d2, _f))
    return idx, d2


def knn_sigma(P, k=8, queries=None):
    """sigma of points `queries` (indices into P; None: all) from their k
    nearest other points of P (see dugma_oracle.h)."""
    P = _c32(P)
    n, D = P.shape
    q = None if queries is None else np.ascontiguousarray(queries, np.int32)
    nq = n if q is None else q.size
    out = np.empty(nq, np.float32)
    lib().dor_knn_sigma(D, nq, _p(q, _i), n, _p(P, _f), int(k), _p(out, _f))
    return out


def knn_cov(P, k=8, scale=0.5, floor=0.002, queries=None):
    """(sigma, C) of points `queries` of P (None: all): sigma as knn_sigma
    and the surface covariance scale * (sample covariance of the k nearest
    other points) + floor * sigma^2 I, C as (nq, D, D) (see dugma_oracle.h)."""
    P = _c32(P)
    n, D = P.shape
    q = None if queries is None else np.ascontiguousarray(queries, np.int32)
    nq = n if q is None else q.size
    sig = np.empty(nq, np.float32)
    cov = np.empty((nq, D * D), np.float32)
    lib().dor_knn_cov(D, nq, _p(q, _i), n, _p(P, _f), int(k), float(scale), float(floor), _p(sig, _f),
                      _p(cov, _f))
    return sig, np.transpose(cov.reshape(nq, D, D), (0, 2, 1))


def energy(Y, SY, X, SX, R, t, grads=False, per_target=False):
    """Convolved energy; Y (N,D), SY (N,D,D), X (M,D), SX (M,D,D)."""
    Y, X = _c64(Y), _c64(X)
    N, D = Y.shape
    M = X.shape[0]
    SY = _c64(np.transpose(np.asarray(SY, np.float64), (0, 2, 1)))
    SX = _c64(np.transpose(np.asarray(SX, np.float64), (0, 2, 1)))
    npose = 6 if D == 3 else 3
    gp = np.zeros(npose) if grads else None
    gY = np.zeros((N, D)) if grads else None
    gSY = np.zeros((N, D, D)) if grads else None
    gX = np.zeros((M, D)) if grads else None
    gSX = np.zeros((M, D, D)) if grads else None
    eT = np.zeros(N) if per_target else None
    E = lib().dor_energy(D, N, _p(Y, _d), _p(SY, _d), M, _p(X, _d), _p(SX, _d),
                         _p(_colmajor(R), _d), _p(_c64(t), _d), _p(gp, _d), _p(gY, _d),
                         _p(gSY, _d), _p(gX, _d), _p(gSX, _d), _p(eT, _d))
    out = {"E": E}
    if grads:
        # covariance gradients were written column-major per point: transpose back
        out.update(gpose=gp, gY=gY, gSY=np.transpose(gSY, (0, 2, 1)), gX=gX,
                   gSX=np.transpose(gSX, (0, 2, 1)))
    if per_target:
        out["eT"] = eT
    return out


def energy_fd_pose(Y, SY, X, SX, R, t, h=1e-4):
    Y, X = _c64(Y), _c64(X)
    N, D = Y.shape
    SY = _c64(np.transpose(np.asarray(SY, np.float64), (0, 2, 1)))
    SX = _c64(np.transpose(np.asarray(SX, np.float64), (0, 2, 1)))
    gp = np.zeros(6 if D == 3 else 3)
    E = lib().dor_energy_fd_pose(D, N, _p(Y, _d), _p(SY, _d), X.shape[0], _p(X, _d), _p(SX, _d),
                                 _p(_colmajor(R), _d), _p(_c64(t), _d), h, _p(gp, _d))
    return E, gp


def _compi_args(Xf, XCinv, gX, Yf, YCinv, gY):
    Xf, Yf = _c32(Xf), _c32(Yf)
    return (Xf, _c32(XCinv), _c32(gX), Yf, _c32(YCinv), _c32(gY))


def compi(Xf, XCinv, gX, Yf, YCinv, gY, R, t, sig):
    """XCinv/YCinv: (n, D*D) column-major prepped inverses (as prep_inverse returns)."""
    Xf, XCinv, gX, Yf, YCinv, gY = _compi_args(Xf, XCinv, gX, Yf, YCinv, gY)
    N, D = Xf.shape
    coef = np.zeros(637 if D == 3 else 45)
    missing = lib().dor_compi(D, N, _p(Xf, _f), _p(XCinv, _f), _p(gX, _f), Yf.shape[0],
                              _p(Yf, _f), _p(YCinv, _f), _p(gY, _f), _p(_colmajor(R), _d),
                              _p(_c64(t), _d), float(sig), _p(coef, _d))
    return coef, missing


def compi_moments(Xf, XCinv, gX, Yf, YCinv, gY, R, t, sig):
    """compi's coefficients from FP64 per-source moments (same arguments and
    result as compi)."""
    Xf, XCinv, gX, Yf, YCinv, gY = _compi_args(Xf, XCinv, gX, Yf, YCinv, gY)
    N, D = Xf.shape
    coef = np.zeros(637 if D == 3 else 45)
    missing = lib().dor_compi_moments(D, N, _p(Xf, _f), _p(XCinv, _f), _p(gX, _f), Yf.shape[0],
                                      _p(Yf, _f), _p(YCinv, _f), _p(gY, _f), _p(_colmajor(R), _d),
                                      _p(_c64(t), _d), float(sig), _p(coef, _d))
    return coef, missing


def surrogate_eval(Xf, XCinv, gX, Yf, YCinv, gY, R, t, sig, dT):
    Xf, XCinv, gX, Yf, YCinv, gY = _compi_args(Xf, XCinv, gX, Yf, YCinv, gY)
    N, D = Xf.shape
    return lib().dor_surrogate_eval(D, N, _p(Xf, _f), _p(XCinv, _f), _p(gX, _f), Yf.shape[0],
                                    _p(Yf, _f), _p(YCinv, _f), _p(gY, _f), _p(_colmajor(R), _d),
                                    _p(_c64(t), _d), float(sig), _p(_c64(dT), _d))


def poly_eval(coef, dT):
    coef = _c64(coef)
    D = 3 if coef.size == 637 else 2
    return lib().dor_poly_eval(D, _p(coef, _d), _p(_c64(dT), _d))


def constraints(dT):
    dT = _c64(dT)
    D = 3 if dT.size == 12 else 2
    ceq = np.zeros(10 if D == 3 else 1)
    lib().dor_constraints(D, _p(dT, _d), _p(ceq, _d))
    return ceq


def pair_weight(n, binv, gx, m, ainv, gy, R, t, sig):
    n, m = _c32(n), _c32(m)
    D = n.size
    return lib().dor_pair_weight(D, _p(n, _f), _p(_c32(binv), _f), float(gx), _p(m, _f),
                                 _p(_c32(ainv), _f), float(gy), _p(_colmajor(R), _d),
                                 _p(_c64(t), _d), float(sig))


def rotm2eul_zyx(R):
    """ZYX Euler angles [z, y, x] (radians) of a rotation matrix, the
    convention of MATLAB's rotm2eul that DUGMA.m:80 calls:
    R = Rz(z) Ry(y) Rx(x)."""
    sy = np.hypot(R[0, 0], R[1, 0])
    return np.array([np.arctan2(R[1, 0], R[0, 0]), np.arctan2(-R[2, 0], sy), np.arctan2(R[2, 1], R[2, 2])])


def normalize(X, Y):
    """normalize.m:3-26: centre each cloud on its mean, divide both by the
    larger RMS radius.  Returns (x, y, xd, yd, scale)."""
    xd, yd = X.mean(1), Y.mean(1)
    x, y = X - xd[:, None], Y - yd[:, None]
    sc = max(np.sqrt((x ** 2).sum() / x.shape[1]), np.sqrt((y ** 2).sum() / y.shape[1]))
    return x / sc, y / sc, xd, yd, sc


def dugma_loop(X, Y, X_C, Y_C, R0, t0, max_iter, acc_rot=1e-4, acc_trans=1e-8, acc_sig=1e-9,
               moments=False):
    """DUGMA.m:19-137 restated on the CPU: normalize.m, the oracle's
    update_sig / compi (moments=True: the moment form, for clouds of
    thousands of points), and scipy's BFGS over a left rotation perturbation
    and dt in place of fmincon (DUGMA.m:65,73).  Returns (R, t, iterations)."""
    from scipy.optimize import minimize
    D = X.shape[0]
    X = np.asarray(X, np.float64)
    Y = np.asarray(Y, np.float64)
    x, y, xd, yd, sc = normalize(X, Y)
    Xf, Yf = x.T.astype(np.float32), y.T.astype(np.float32)
    compi_fn = compi_moments if moments else compi

    def blocks_to_pts(C):
        n = C.shape[1] // D
        return np.transpose(C.reshape(D, n, D), (1, 0, 2))

    XCi, gX, _ = prep_inverse(blocks_to_pts(X_C), D)
    YCi, gY, _ = prep_inverse(blocks_to_pts(Y_C), D)
    R, t = np.array(R0, np.float64), np.array(t0, np.float64).ravel()
    sig = update_sig(Xf, Yf, R, t)

    def rot(p):
        if D == 2:
            c, s = np.cos(p[0]), np.sin(p[0])
            return np.array([[c, -s], [s, c]])
        th = np.linalg.norm(p[:3])
        K = np.array([[0, -p[2], p[1]], [p[2], 0, -p[0]], [-p[1], p[0], 0]])
        if th < 1e-12:
            return np.eye(3) + K
        return np.eye(3) + np.sin(th) / th * K + (1 - np.cos(th)) / th ** 2 * K @ K

    def dT_of(p):
        dR = rot(p)
        if D == 2:
            return np.array([p[1], p[2], dR[0, 0], dR[1, 0]])
        return np.concatenate([p[3:], dR.ravel()])

    angle = np.full(3, 10000.0)
    tprev = t.copy()
    it = 1
    con = True
    while it <= max_iter and (con or ((tprev - t) ** 2).sum() > acc_trans) and sig > acc_sig:
        coef, _ = compi_fn(Xf, XCi, gX, Yf, YCi, gY, R, t, sig)
        np_ = 6 if D == 3 else 3
        res = minimize(lambda p: poly_eval(coef, dT_of(p)), np.zeros(np_), method="BFGS",
                       options={"gtol": 1e-12 * max(1.0, abs(poly_eval(coef, dT_of(np.zeros(np_))))),
                                "maxiter": 500})
        dR = rot(res.x)
        dt = res.x[3:] if D == 3 else res.x[1:]
        R = dR @ R
        tprev = t.copy()
        t = t + dt
        a0 = angle.copy()
        if D == 3:
            angle = rotm2eul_zyx(R)
            con = bool(np.any(np.abs(a0 - angle) > acc_rot))
        else:
            angle = np.array([np.degrees(np.arcsin(np.clip(R[1, 0], -1, 1))), 0, 0])
            con = bool(abs(a0[0] - angle[0]) > acc_rot)
        sig = update_sig(Xf, Yf, R, t)
        it += 1
    return R, xd + t * sc - R @ yd, it - 1
