"""dugma_amd -- MI355X (gfx950) implementation of DUGMA's registration hot path.

Python mirror of the reference's MEX entry points (DUGMA.m:44-126):

    ok, h = init(X, X_C, Y, Y_C)              # init.cpp:36-83
    sig = update_sig(h, R, t, N, M, D)        # update_sig.cpp:8-66
    coef = compi(h, R, t, N, M, D, sig)       # compi.cpp:8-74
    ok = fr_m(h)                              # fr_m.cpp:8-35

X is the fixed target cloud (D x N, MATLAB layout), Y the moving source cloud
(D x M), X_C / Y_C the D x (D*N) covariance blocks; R, t the current pose.
The MEX functions also hand back six device handles; here the one context
object ``h`` stands for all six.

The convolved-uncertainty energy  E = sum_ij N(X_i - R Y_j - t; CX_i + R CY_j R^T)
and its adjoints are exposed through ``energy``, ``energy_grad``,
``energy_grad_points`` and ``energy_grad_fd``; exact k nearest neighbours
through ``knn``.  Everything runs in libdugma_hip.so; nothing here computes.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import DugmaError

COV_AUTO, COV_FULL, COV_ISO = 0, 1, 2
KNN_TARGET_SELF, KNN_SOURCE_SELF, KNN_TARGET_TO_SOURCE = 0, 1, 2
SIGMA_TARGET, SIGMA_SOURCE = 1, 2
_COV = {"auto": COV_AUTO, "full": COV_FULL, "iso": COV_ISO}


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _cm(a, shape=None):
    """Column-major FP64 copy (MATLAB memory order)."""
    a = np.asfortranarray(np.asarray(a, dtype=np.float64))
    if shape is not None and a.shape != shape:
        raise ValueError("expected shape %s, got %s" % (shape, a.shape))
    return a


def _cov_mode(mode):
    return _COV[mode] if isinstance(mode, str) else int(mode)


class _PoseBuffers:
    """Per-context argument and result buffers of the per-pose calls, with
    their ctypes pointers made once (building a pointer from a numpy array
    costs more per call than the copy of R and t).  A context is not
    re-entrant, so one set serves every call on it."""

    def __init__(self, D):
        self.D = D
        self.R = np.zeros((D, D), order="F")
        self.t = np.zeros(D)
        self.g = np.zeros(6 if D == 3 else 3)
        self.E = ctypes.c_double()
        self.Rp, self.tp, self.gp = _dp(self.R), _dp(self.t), _dp(self.g)
        self.Ep = ctypes.byref(self.E)

    def pose(self, R, t):
        R = np.asarray(R, dtype=np.float64)
        if R.shape != (self.D, self.D):
            raise ValueError("expected shape %s, got %s" % ((self.D, self.D), R.shape))
        t = np.ravel(np.asarray(t, dtype=np.float64))
        if t.shape != (self.D,):
            raise ValueError("expected %d translation entries, got %d" % (self.D, t.size))
        self.R[...] = R
        self.t[...] = t


class Context:
    """The device state created by ``init`` (the reference's six handles)."""

    def __init__(self, ptr, N, M, D):
        self._ptr = ctypes.c_void_p(ptr)
        self.N, self.M, self.D = N, M, D
        self._pb = _PoseBuffers(D)

    @property
    def ptr(self):
        if not self._ptr or not self._ptr.value:
            raise DugmaError(-5, "context already freed")
        return self._ptr

    def stream(self):
        """hipStream_t the context launches on (as an integer)."""
        return _lib.load().dugma_ctx_stream(self.ptr)

    def set_timing(self, on=True):
        _lib.check(_lib.load().dugma_ctx_set_timing(self.ptr, int(bool(on))))

    def pair_time(self):
        """(milliseconds, launches) of the all-pairs kernels since set_timing(True)."""
        ms, n = ctypes.c_double(), ctypes.c_longlong()
        _lib.check(_lib.load().dugma_ctx_pair_time(self.ptr, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def autotune(self, cov="auto"):
        """Time candidate launch shapes of the energy passes; keep the fastest."""
        _lib.check(_lib.load().dugma_autotune(self.ptr, _cov_mode(cov)))
        T, w = ctypes.c_int(), ctypes.c_int()
        _lib.check(_lib.load().dugma_ctx_plan(self.ptr, _cov_mode(cov), ctypes.byref(T), ctypes.byref(w)))
        return T.value, w.value

    def energy_info(self):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _lib.check(_lib.load().dugma_ctx_energy_info(self.ptr, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, bool(c.value)

    def read_handle(self, which):
        """Host copy of reference handle ``which`` (0..5: X, inv X_C, Y, inv Y_C, g_X, g_Y)."""
        N, M, D = self.N, self.M, self.D
        n = [N * D, N * D * D, M * D, M * D * D, N, M][which]
        out = np.empty(n, np.float32)
        _lib.check(_lib.load().dugma_read_handle(self.ptr, int(which),
                                                 out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def free(self):
        if self._ptr and self._ptr.value:
            ok = ctypes.c_double()
            _lib.check(_lib.load().dugma_fr_m(self._ptr, ctypes.byref(ok)))
            self._ptr = ctypes.c_void_p(0)
            return ok.value
        return 1.0


def init(X, X_C, Y, Y_C):
    """[ok, h] = init(X, X_C, Y, Y_C): X (D,N), X_C (D,D*N), Y (D,M), Y_C (D,D*M)."""
    X = _cm(X)
    D, N = X.shape
    Y = _cm(Y)
    M = Y.shape[1]
    if Y.shape[0] != D:
        raise ValueError("X and Y must have the same dimension")
    X_C = _cm(X_C, (D, D * N))
    Y_C = _cm(Y_C, (D, D * M))
    out = ctypes.c_void_p()
    ok = ctypes.c_double()
    _lib.check(_lib.load().dugma_init(_dp(X), _dp(X_C), _dp(Y), _dp(Y_C), N, M, D,
                                      ctypes.byref(out), ctypes.byref(ok)))
    return ok.value, Context(out.value, N, M, D)


def update_sig(h, R, t, N=None, M=None, D=None):
    """sig = update_sig(h, R, t, N, M, D) (FP32 value returned as a float)."""
    b = h._pb
    b.pose(R, t)
    _lib.check(_lib.load().dugma_update_sig(h.ptr, b.Rp, b.tp, float(N or h.N),
                                            float(M or h.M), float(D or h.D), b.Ep))
    return b.E.value


def compi(h, R, t, N=None, M=None, D=None, sig=1.0):
    """coef = compi(h, R, t, N, M, D, sig): 637 (3-D) or 45 (2-D) coefficients."""
    b = h._pb
    b.pose(R, t)
    coef = np.zeros(637 if h.D == 3 else 45)
    _lib.check(_lib.load().dugma_compi(h.ptr, b.Rp, b.tp, float(N or h.N), float(M or h.M),
                                       float(D or h.D), float(sig), _dp(coef)))
    return coef


def fr_m(h):
    """ok = fr_m(h)."""
    return h.free()


def poly_eval(coef, dT, grad=False):
    """F = energy3D(dT) / energy2D(dT) (and dF/ddT) for compi's coefficients."""
    coef = np.ascontiguousarray(coef, np.float64)
    dT = np.ascontiguousarray(dT, np.float64)
    D = 3 if coef.size == 637 else 2
    g = np.zeros(dT.size) if grad else None
    F = _lib.load().dugma_poly_eval(D, _dp(coef), _dp(dT), _dp(g) if grad else None)
    return (F, g) if grad else F


def poly_minimize(coef):
    """dT minimising the surrogate over rotations dR and translations dt
    (what fmincon('energy3D', ..., 'fcontr3D') returns in DUGMA.m:73)."""
    coef = np.ascontiguousarray(coef, np.float64)
    D = 3 if coef.size == 637 else 2
    dT = np.zeros(12 if D == 3 else 4)
    F = ctypes.c_double()
    _lib.check(_lib.load().dugma_poly_minimize(D, _dp(coef), _dp(dT), ctypes.byref(F)))
    return dT, F.value


def DUGMA(X, Y, X_C, Y_C, R0, t0, Max_Iteration=300, accuracy_rotation=1e-4,
          accuracy_translation=1e-8, accuracy_sig=1e-9, verbose=False):
    """[R, t] = DUGMA(X, Y, X_C, Y_C, R0, t0, ...) with the argument order of
    DUGMA.m:19; returns (R, t, iterations, sig)."""
    X = _cm(X)
    D, N = X.shape
    Y = _cm(Y)
    M = Y.shape[1]
    X_C = _cm(X_C, (D, D * N))
    Y_C = _cm(Y_C, (D, D * M))
    R0 = _cm(R0, (D, D))
    t0 = _cm(np.ravel(t0))
    o = _lib.RegisterOpts(int(Max_Iteration), float(accuracy_rotation), float(accuracy_translation),
                          float(accuracy_sig), int(bool(verbose)))
    R = np.zeros((D, D), order="F")
    t = np.zeros(D)
    it = ctypes.c_int()
    sig = ctypes.c_double()
    _lib.check(_lib.load().dugma_register(_dp(X), _dp(X_C), _dp(Y), _dp(Y_C), N, M, D, _dp(R0), _dp(t0),
                                          ctypes.byref(o), _dp(R), _dp(t), ctypes.byref(it), ctypes.byref(sig)))
    return np.ascontiguousarray(R), t, it.value, sig.value


def energy(h, R, t, cov="auto"):
    b = h._pb
    b.pose(R, t)
    _lib.check(_lib.load().dugma_energy(h.ptr, b.Rp, b.tp, _cov_mode(cov), b.Ep))
    return b.E.value


def energy_grad(h, R, t, cov="auto"):
    """(E, gpose): gpose = [dE/dw (3), dE/dt (3)] (3-D) or [dE/dtheta, dE/dt (2)]."""
    b = h._pb
    b.pose(R, t)
    _lib.check(_lib.load().dugma_energy_grad(h.ptr, b.Rp, b.tp, _cov_mode(cov), b.Ep, b.gp))
    return b.E.value, b.g.copy()


def energy_grad_hess(h, R, t, cov="auto"):
    """(E, gpose (6), H (6, 6)): the fused gradient and the pose Hessian of
    E from one pass (3-D; see include/dugma.h)."""
    R = _cm(R, (3, 3))
    t = _cm(np.ravel(t))
    E = ctypes.c_double()
    g = np.zeros(6)
    H = np.zeros(36)
    _lib.check(_lib.load().dugma_energy_grad_hess(h.ptr, _dp(R), _dp(t), _cov_mode(cov), ctypes.byref(E),
                                                  _dp(g), _dp(H)))
    return E.value, g, H.reshape(6, 6)


def energy_grad_fd(h, R, t, step=1e-4, cov="auto"):
    R = _cm(R, (h.D, h.D))
    t = _cm(np.ravel(t))
    E = ctypes.c_double()
    g = np.zeros(6 if h.D == 3 else 3)
    _lib.check(_lib.load().dugma_energy_grad_fd(h.ptr, _dp(R), _dp(t), _cov_mode(cov), float(step),
                                                ctypes.byref(E), _dp(g)))
    return E.value, g


def register_energy(h, R0, t0, max_iter=200, tol_rot=1e-7, tol_trans=1e-14, cov="auto"):
    """Registration on the convolved energy (dugma_register_energy): descent
    of -log E over left pose increments from (R0, t0), Newton directions
    from each pass's pose Hessian where positive definite (3-D), else BFGS.
    Returns (R, t, iterations, E)."""
    D = h.D
    R0 = _cm(R0, (D, D))
    t0 = _cm(np.ravel(t0))
    o = _lib.RegisterEnergyOpts(int(max_iter), float(tol_rot), float(tol_trans), _cov_mode(cov))
    R = np.zeros((D, D), order="F")
    t = np.zeros(D)
    it = ctypes.c_int()
    E = ctypes.c_double()
    _lib.check(_lib.load().dugma_register_energy(h.ptr, _dp(R0), _dp(t0), ctypes.byref(o), _dp(R), _dp(t),
                                                 ctypes.byref(it), ctypes.byref(E)))
    return np.ascontiguousarray(R), t, it.value, E.value


def energy_grad_points(h, R, t, cov="auto"):
    """Energy with per-point gradients.  Returns a dict of numpy arrays:
    gX (N,D), gXC (N,D,D), gY (M,D), gYC (M,D,D), gpose, gdT (12 or 4)."""
    D, N, M = h.D, h.N, h.M
    R = _cm(R, (D, D))
    t = _cm(np.ravel(t))
    E = ctypes.c_double()
    gpose = np.zeros(6 if D == 3 else 3)
    gX = np.zeros((N, D))
    gXC = np.zeros((N, D, D))
    gY = np.zeros((M, D))
    gYC = np.zeros((M, D, D))
    gdT = np.zeros(12 if D == 3 else 4)
    _lib.check(_lib.load().dugma_energy_grad_points(h.ptr, _dp(R), _dp(t), _cov_mode(cov),
                                                    ctypes.byref(E), _dp(gpose), _dp(gX), _dp(gXC),
                                                    _dp(gY), _dp(gYC), _dp(gdT)))
    # covariance gradients are column-major per point
    return dict(E=E.value, gpose=gpose, gX=gX, gXC=np.transpose(gXC, (0, 2, 1)), gY=gY,
                gYC=np.transpose(gYC, (0, 2, 1)), gdT=gdT)


def knn(h, which, k, R=None, t=None):
    """(idx, d2) of the k nearest neighbours; see include/dugma.h for ``which``."""
    nq = h.M if which == KNN_SOURCE_SELF else h.N
    idx = np.empty((nq, k), np.int32)
    d2 = np.empty((nq, k), np.float32)
    Rp = tp = None
    if R is not None:
        R = _cm(R, (h.D, h.D))
        t = _cm(np.ravel(t))
        Rp, tp = _dp(R), _dp(t)
    _lib.check(_lib.load().dugma_knn(h.ptr, int(which), int(k), Rp, tp,
                                     idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                     d2.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
    return idx, d2


def update_sigma_knn(h, which=SIGMA_TARGET | SIGMA_SOURCE, k=8, want_sigma=True):
    """Per-point isotropic covariances sigma_i^2 I from the mean distance to
    the k nearest neighbours within each selected cloud, on the device (see
    include/dugma.h).  Returns (sigma_target, sigma_source, unchanged); a
    sigma array is None for a cloud not selected or when want_sigma is
    False."""
    fp = ctypes.POINTER(ctypes.c_float)
    sT = np.empty(h.N, np.float32) if (want_sigma and which & SIGMA_TARGET) else None
    sS = np.empty(h.M, np.float32) if (want_sigma and which & SIGMA_SOURCE) else None
    un = ctypes.c_int(0)
    _lib.check(_lib.load().dugma_update_sigma_knn(
        h.ptr, int(which), int(k), None if sT is None else sT.ctypes.data_as(fp),
        None if sS is None else sS.ctypes.data_as(fp), ctypes.byref(un)))
    return sT, sS, un.value


def update_cov_knn(h, which=SIGMA_TARGET | SIGMA_SOURCE, k=8, scale=0.5, floor=0.002, want_sigma=True):
    """Per-point surface covariances scale * (sample covariance of the k
    nearest neighbours) + floor * sigma_i^2 I within each selected cloud, on
    the device (see include/dugma.h).  Returns (sigma_target, sigma_source,
    unchanged) as update_sigma_knn."""
    fp = ctypes.POINTER(ctypes.c_float)
    sT = np.empty(h.N, np.float32) if (want_sigma and which & SIGMA_TARGET) else None
    sS = np.empty(h.M, np.float32) if (want_sigma and which & SIGMA_SOURCE) else None
    un = ctypes.c_int(0)
    _lib.check(_lib.load().dugma_update_cov_knn(
        h.ptr, int(which), int(k), float(scale), float(floor), None if sT is None else sT.ctypes.data_as(fp),
        None if sS is None else sS.ctypes.data_as(fp), ctypes.byref(un)))
    return sT, sS, un.value


def version():
    return _lib.load().dugma_version().decode()


class Batch:
    """Many independent 3-D registrations in one device context
    (dugma_batch_create / dugma_batch_register).  ``trials`` is a list of
    dicts with X, X_C, Y, Y_C in the layout of ``init`` (isotropic
    covariances), or X, sX, Y, sY (one variance per point).  Trials that
    also hold CX and CY (n x 3 x 3 full covariances per point) get them set
    for the refinement phase (dugma_batch_set_covariances)."""

    def __init__(self, trials):
        K = len(trials)
        iso = "sX" in trials[0]
        names = ("X", "sX", "Y", "sY") if iso else ("X", "X_C", "Y", "Y_C")
        ptrs = {k: (ctypes.POINTER(ctypes.c_double) * K)() for k in names}
        N = (ctypes.c_int * K)()
        M = (ctypes.c_int * K)()
        keep = []
        for i, tr in enumerate(trials):
            arrs = {k: _cm(tr[k]) for k in names}
            keep.append(arrs)
            N[i], M[i] = arrs["X"].shape[1], arrs["Y"].shape[1]
            for k in names:
                ptrs[k][i] = _dp(arrs[k])
        out = ctypes.c_void_p()
        L = _lib.load()
        if iso:
            _lib.check(L.dugma_batch_create_iso(K, N, M, ptrs["X"], ptrs["sX"], ptrs["Y"], ptrs["sY"],
                                                ctypes.byref(out)))
        else:
            _lib.check(L.dugma_batch_create(K, 3, N, M, ptrs["X"], ptrs["X_C"], ptrs["Y"], ptrs["Y_C"],
                                            ctypes.byref(out)))
        del keep  # the library holds its own copy
        self._ptr = out
        self.ntrials = K
        if "CX" in trials[0]:
            self.set_covariances([tr["CX"] for tr in trials], [tr["CY"] for tr in trials])

    def set_covariances(self, CX, CY):
        """Full per-point covariances (lists of (n, 3, 3) arrays) for the
        refinement phase (dugma_batch_set_covariances)."""
        K = self.ntrials
        px = (ctypes.POINTER(ctypes.c_double) * K)()
        py = (ctypes.POINTER(ctypes.c_double) * K)()
        keep = []
        for k in range(K):
            # (n, 3, 3) -> one column-major 3 x 3 block per point
            a = np.ascontiguousarray(np.transpose(np.asarray(CX[k], np.float64), (0, 2, 1)))
            b = np.ascontiguousarray(np.transpose(np.asarray(CY[k], np.float64), (0, 2, 1)))
            keep.append((a, b))
            px[k], py[k] = _dp(a), _dp(b)
        _lib.check(_lib.load().dugma_batch_set_covariances(self._ptr, px, py))
        del keep

    def info(self):
        k, n, p = ctypes.c_int(), ctypes.c_int(), ctypes.c_longlong()
        _lib.check(_lib.load().dugma_batch_info(self._ptr, ctypes.byref(k), ctypes.byref(n), ctypes.byref(p)))
        return dict(ntrials=k.value, items=n.value, pairs_per_iteration=p.value)

    def _poses(self, R0, t0):
        K = self.ntrials
        R0 = np.tile(np.eye(3), (K, 1, 1)) if R0 is None else np.asarray(R0, np.float64)
        t0 = np.zeros((K, 3)) if t0 is None else np.asarray(t0, np.float64)
        return np.ascontiguousarray(np.transpose(R0, (0, 2, 1))), np.ascontiguousarray(t0)  # column-major R

    def autotune(self, R0=None, t0=None, sigma_update=True):
        """dugma_batch_autotune: returns the chosen launch shape, its time per
        isotropic iteration and (with covariances set) per refinement pass."""
        R0c, t0c = self._poses(R0, t0)
        w, g, ms, rms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        ch = (ctypes.c_int * 4)()
        _lib.check(_lib.load().dugma_batch_autotune(self._ptr, _dp(R0c), _dp(t0c), int(bool(sigma_update)),
                                                    ctypes.byref(w), ch, ctypes.byref(g), ctypes.byref(ms),
                                                    ctypes.byref(rms)))
        levels = [self.get_shape(l) for l in range(3)]
        return dict(waves=w.value, chunks=list(ch), grid_waves=g.value, ms_per_iteration=ms.value,
                    ms_per_refine_pass=rms.value, levels=levels)

    def get_shape(self, level):
        """dugma_batch_get_shape: one level's launch shape (0 full clouds,
        1 every 4th point, 2 every 2nd point)."""
        w, g = ctypes.c_int(), ctypes.c_int()
        ch = (ctypes.c_int * 4)()
        _lib.check(_lib.load().dugma_batch_get_shape(self._ptr, int(level), ctypes.byref(w), ch, ctypes.byref(g)))
        return dict(waves=w.value, chunks=list(ch), grid_waves=g.value)

    def set_level_shape(self, level, waves, chunks, grid_waves):
        ch = (ctypes.c_int * 4)(*[int(c) for c in chunks])
        _lib.check(_lib.load().dugma_batch_set_level_shape(self._ptr, int(level), int(waves), ch, int(grid_waves)))

    def set_shape(self, waves, chunks, grid_waves):
        ch = (ctypes.c_int * 4)(*[int(c) for c in chunks])
        _lib.check(_lib.load().dugma_batch_set_shape(self._ptr, int(waves), ch, int(grid_waves)))

    def register(self, R0=None, t0=None, max_iter=80, anneal=True, anneal_decay=0.85,
                 tol_rot=1e-7, tol_trans=1e-7, use_graph=True, one_trial_per_launch=False,
                 check_every=10, grid=0, sigma_update=True, relax=1.0, coarse_iters=0, refine_iters=0,
                 refine_coarse_iters=0, refine_mid_iters=0):
        K = self.ntrials
        R0c, t0c = self._poses(R0, t0)
        o = _lib.BatchOpts(int(max_iter), int(bool(anneal)), float(anneal_decay), float(tol_rot),
                           float(tol_trans), int(bool(use_graph)), int(bool(one_trial_per_launch)),
                           int(check_every), int(grid), int(bool(sigma_update)), float(relax),
                           int(coarse_iters), int(refine_iters), int(refine_coarse_iters),
                           int(refine_mid_iters))
        R = np.zeros((K, 9))
        t = np.zeros((K, 3))
        iters = np.zeros(K, np.int32)
        status = np.zeros(K, np.int32)
        E = np.zeros(K)
        ip = ctypes.POINTER(ctypes.c_int)
        _lib.check(_lib.load().dugma_batch_register(self._ptr, _dp(R0c), _dp(t0c), ctypes.byref(o), _dp(R),
                                                    _dp(t), iters.ctypes.data_as(ip), status.ctypes.data_as(ip),
                                                    _dp(E)))
        return dict(R=np.transpose(R.reshape(K, 3, 3), (0, 2, 1)), t=t, iters=iters, status=status, E=E)

    def free(self):
        if self._ptr and self._ptr.value:
            _lib.check(_lib.load().dugma_batch_free(self._ptr))
            self._ptr = ctypes.c_void_p(0)
