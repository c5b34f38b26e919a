"""The DUGMA.m loop (dugma_register) end to end on the GPU."""
import os

import numpy as np
import pytest

import dugma_amd as dg
from dugma_amd import workloads as W

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def kinect_scale_case(seed, n, m, deg, var=2.25):
    """Anisotropic surface with covariances of the Kinect data's size
    relative to the normalised cloud (the regime DUGMA's Lambda/sig weights
    are made for)."""
    rng = np.random.default_rng(seed)
    surf = W.Surface(rng, nt=60, nph=120)
    A = np.diag([1.6, 1.0, 0.6])
    src = surf.sample(rng, m) @ A
    tgt0 = surf.sample(rng, n) @ A
    R = W.random_rotation(rng, deg)
    t = rng.uniform(-0.1, 0.1, 3)
    tgt = tgt0 @ R.T + t + rng.normal(scale=0.003, size=(n, 3))
    C = np.tile(np.eye(3) * var, (max(n, m), 1, 1))
    return dict(X=tgt.T.copy(), X_C=W.blocks(C[:n]), Y=src.T.copy(), Y_C=W.blocks(C[:m]), R_G=R, t_G=t)


def test_loop_iterates_match_cpu_restatement(oracle):
    """Three iterations of the GPU loop against the CPU restatement (oracle
    compi / update_sig, scipy BFGS instead of the library's Newton)."""
    d = kinect_scale_case(3, 260, 230, 30.0)
    R, t, it, sig = dg.DUGMA(d["X"], d["Y"], d["X_C"], d["Y_C"], np.eye(3), np.zeros(3), 3)
    Rc, tc, itc = oracle.dugma_loop(d["X"], d["Y"], d["X_C"], d["Y_C"], np.eye(3), np.zeros(3), 3)
    assert it == itc == 3
    assert W.rotation_error_deg(R, Rc) < 1e-3
    assert np.abs(t - tc).max() < 1e-5


def test_loop_converges_toward_ground_truth():
    errs = []
    for seed, deg in ((3, 35.0), (5, 25.0)):
        d = kinect_scale_case(seed, 1500, 1300, deg)
        start = W.rotation_error_deg(np.eye(3), d["R_G"])
        R, t, it, sig = dg.DUGMA(d["X"], d["Y"], d["X_C"], d["Y_C"], np.eye(3), np.zeros(3), 150)
        assert abs(np.linalg.det(R) - 1) < 1e-9
        errs.append((start, W.rotation_error_deg(R, d["R_G"])))
    assert all(e < 0.5 * s for s, e in errs), errs


def test_kinect_fixture_scenes_run():
    """Decimated Kinect scenes (every 25th point, NaN covariances kept): the
    loop runs to completion with finite, proper rotations
    (main_kinect_real_application.m:33-34 error measures are reported)."""
    g = np.load(os.path.join(HERE, "golden", "kinect_fixtures.npz"))
    for k in (2, 11, 20, 22):
        X, XC, Y, YC = g["s%d_X" % k], g["s%d_X_C" % k], g["s%d_Y" % k], g["s%d_Y_C" % k]
        R, t, it, sig = dg.DUGMA(X, Y, XC, YC, np.eye(3), np.zeros(3), 300, 1e-4, 1e-8, 1e-9)
        assert it >= 1 and np.all(np.isfinite(R)) and np.all(np.isfinite(t))
        assert abs(np.linalg.det(R) - 1) < 1e-9


FULL_SCENES = (2, 8, 20, 22)


def full_scene(k):
    """A whole Kinect scene of the reference (dataset_all/Kinect_Application/k.mat):
    covariance blocks v I rebuilt from one variance per point."""
    g = np.load(os.path.join(HERE, "golden", "kinect_full_scenes.npz"))
    out = {"X": g["s%d_X" % k], "Y": g["s%d_Y" % k], "R_G": g["s%d_R_G" % k], "t_G": g["s%d_t_G" % k]}
    for name in ("X", "Y"):
        v = g["s%d_v%s" % (k, name)]
        C = np.zeros((len(v), 3, 3))
        for a in range(3):
            C[:, a, a] = v
        out[name + "_C"] = W.blocks(C)
    return out


def published(k):
    g = np.load(os.path.join(HERE, "golden", "kinect_fixtures.npz"))
    return g["published_R"][k - 1], g["published_t"][k - 1]


def kinect_errors(R, t, R_G, t_G):
    """main_kinect_real_application.m:33-34 (error_R, error_t) and the
    rotation error in degrees."""
    err_R = float(np.sqrt(((np.eye(3) - R @ np.linalg.pinv(R_G)) ** 2).sum()))
    err_t = float(np.sqrt(((t - t_G) ** 2).sum()))
    return err_R, err_t, W.rotation_error_deg(R, R_G)


@pytest.mark.parametrize("k", FULL_SCENES)
def test_kinect_full_scene_matches_published(k):
    """Whole Kinect scenes with the reference's own driver settings
    (main_kinect_real_application.m:14-21: R0 = I, t0 = 0, 300 iterations,
    accuracies 1e-4 / 1e-8 / 1e-9): the registration lands on the pose the
    reference published for the scene (Ours_k_result.mat), and within 2
    degrees of the ground truth where the reference got within 1 degree."""
    d = full_scene(k)
    R, t, it, sig = dg.DUGMA(d["X"], d["Y"], d["X_C"], d["Y_C"], np.eye(3), np.zeros(3), 300, 1e-4, 1e-8, 1e-9)
    Rp, tp = published(k)
    ours = kinect_errors(R, t, d["R_G"], d["t_G"])
    pub = kinect_errors(Rp, tp, d["R_G"], d["t_G"])
    gap_deg = W.rotation_error_deg(R, Rp)
    print("scene %d: %d iterations; error_R %.4f error_t %.5f (%.4f deg); published error_R %.4f error_t %.5f "
          "(%.4f deg); ours vs published %.4f deg, %.6f" % ((k, it) + ours + pub + (gap_deg, np.linalg.norm(t - tp))))
    assert it < 300
    assert gap_deg < 0.05
    assert np.linalg.norm(t - tp) < 1e-3 * np.linalg.norm(d["t_G"])
    if pub[2] < 1.0:
        assert ours[2] < 2.0


def test_2d_loop_runs(oracle):
    d = W.config5_2d(n=300, m=250)
    R, t, it, sig = dg.DUGMA(d["X"], d["Y"], d["X_C"], d["Y_C"], np.eye(2), np.zeros(2), 3)
    Rc, tc, itc = oracle.dugma_loop(d["X"], d["Y"], d["X_C"], d["Y_C"], np.eye(2), np.zeros(2), 3)
    assert it == itc
    assert abs(np.arctan2(R[1, 0], R[0, 0]) - np.arctan2(Rc[1, 0], Rc[0, 0])) < 1e-5
    assert np.abs(t - tc).max() < 1e-5


# ---------------------------------------------------------------- registration on the convolved energy

def _f32(a):
    return np.asarray(a, np.float64).astype(np.float32).astype(np.float64)


def _skew(w):
    return np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])


def _exp_so3(w):
    th = np.linalg.norm(w)
    K = _skew(w)
    if th < 1e-12:
        return np.eye(3) + K
    return np.eye(3) + np.sin(th) / th * K + (1 - np.cos(th)) / th ** 2 * K @ K


def _left_jacobian(w):
    """J with exp([w + dw]x) = exp([J dw]x) exp([w]x) to first order."""
    th = np.linalg.norm(w)
    K = _skew(w)
    if th < 1e-8:
        return np.eye(3) + 0.5 * K
    return np.eye(3) + (1 - np.cos(th)) / th ** 2 * K + (th - np.sin(th)) / th ** 3 * K @ K


def _oracle_minimise(oracle, Y, SY, X, SX, R0, t0):
    """-log E minimised in FP64 by scipy's BFGS over p = (w, dt) with
    R = exp([w]x) R0 (2-D: Rot(theta) R0), t = t0 + dt, on the oracle's
    energy and pose gradient: a minimiser independent of the library's."""
    from scipy.optimize import minimize
    D = Y.shape[1]

    def pose(p):
        if D == 3:
            return _exp_so3(p[:3]) @ R0, t0 + p[3:]
        c, s = np.cos(p[0]), np.sin(p[0])
        return np.array([[c, -s], [s, c]]) @ R0, t0 + p[1:]

    def fg(p):
        R, t = pose(p)
        o = oracle.energy(Y, SY, X, SX, R, t, grads=True)
        g = np.array(o["gpose"], float)
        if D == 3:
            g[:3] = _left_jacobian(p[:3]).T @ g[:3]
        return -np.log(o["E"]), -g / o["E"]

    res = minimize(fg, np.zeros(6 if D == 3 else 3), jac=True, method="BFGS",
                   options={"gtol": 1e-10, "maxiter": 400})
    return pose(res.x)


@pytest.mark.parametrize("D", [3, 2])
def test_register_energy_matches_independent_minimiser(oracle, D):
    """dugma_register_energy against scipy's BFGS on the FP64 oracle from
    the same start: the same pose, stationary under the oracle's gradient,
    and a higher energy than at the start.  Isotropic clouds with sigma from
    the 8-NN spacing (and 2-D contours): with covariances of a sensor's mm
    noise on independently sampled surfaces the energy is a field of spikes
    at near-coincident pairs, and local minimisers stop on different ones."""
    if D == 3:
        d = W.config3(n=1500, m=1400)
        R0 = W.random_rotation(np.random.default_rng(5), 4.0) @ d["R_G"]
        t0 = np.ravel(d["t_G"]) + np.array([0.02, -0.015, 0.02])
    else:
        d = W.config5_2d(n=500, m=420)
        R0 = W.rot2(np.radians(4.0)) @ d["R_G"]
        t0 = np.ravel(d["t_G"]) + np.array([0.01, -0.008])
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    try:
        E0 = dg.energy(h, R0, t0)
        R, t, it, E = dg.register_energy(h, R0, t0)
    finally:
        dg.fr_m(h)
    assert 1 <= it < 200
    assert E > E0
    Y, SY = _f32(d["X"].T), _f32(W.unblocks(d["X_C"]))
    X, SX = _f32(d["Y"].T), _f32(W.unblocks(d["Y_C"]))
    Rc, tc = _oracle_minimise(oracle, Y, SY, X, SX, R0, np.ravel(t0))
    L = np.sqrt(((X - X.mean(0)) ** 2).sum(1).mean())
    if D == 3:
        dang = W.rotation_error_deg(R, Rc)
    else:
        dang = abs(np.degrees(np.arctan2(R[1, 0], R[0, 0]) - np.arctan2(Rc[1, 0], Rc[0, 0])))
    assert dang < 5e-3, dang
    assert np.linalg.norm(t - tc) < 5e-5 * L, (t, tc)
    # stationary under the FP64 gradient of log E (rotation part times L)
    o0 = oracle.energy(Y, SY, X, SX, R0, t0, grads=True)
    o1 = oracle.energy(Y, SY, X, SX, R, t, grads=True)
    nr = 3 if D == 3 else 1

    def scaled(o):
        g = np.array(o["gpose"], float) / o["E"]
        g[:nr] /= L
        return np.linalg.norm(g)
    assert scaled(o1) <= 1e-3 * scaled(o0), (scaled(o1), scaled(o0))


def _left(R, t, p):
    """The pose after the left increment p = [w, dt]: exp([w]x) R, t + dt."""
    w = np.asarray(p[:3], float)
    th = np.linalg.norm(w)
    K = np.array([[0.0, -w[2], w[1]], [w[2], 0.0, -w[0]], [-w[1], w[0], 0.0]])
    A = np.sin(th) / th if th > 1e-12 else 1.0
    B = (1.0 - np.cos(th)) / th ** 2 if th > 1e-12 else 0.5
    return (np.eye(3) + A * K + B * K @ K) @ R, np.asarray(t, float) + np.asarray(p[3:], float)


@pytest.mark.parametrize("cov", ["iso", "full"])
def test_pose_hessian_matches_gradient_differences(cov):
    """dugma_energy_grad_hess at a converged pose (gradient ~ 0: there the
    Hessian does not depend on how pose increments compose) against central
    differences of dugma_energy_grad's own gradient, translation parameters
    times the clouds' RMS radius: every entry for isotropic covariances (the
    Hessian is exact there), the translation block for surface covariances
    (the turn of the covariances is left out of the rest), 2e-3 of the
    largest entry."""
    d = W.config3(n=1500, m=1400)
    R0 = W.random_rotation(np.random.default_rng(5), 4.0) @ d["R_G"]
    t0 = np.ravel(d["t_G"]) + np.array([0.02, -0.015, 0.02])
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    try:
        if cov == "full":
            dg.update_cov_knn(h)  # surface covariances on both clouds
        # the same energy and gradient as the pose pass (summed in another order)
        E1, g1, _ = dg.energy_grad_hess(h, R0, t0, cov=cov)
        E2, g2 = dg.energy_grad(h, R0, t0, cov=cov)
        assert abs(E1 / E2 - 1) < 1e-10
        assert np.abs(g1 - g2).max() < 1e-8 * np.abs(g2).max()
        R, t, it, E = dg.register_energy(h, R0, t0, cov=cov)
        E1, g, H = dg.energy_grad_hess(h, R, t, cov=cov)
        Y = np.asarray(d["Y"]).T
        L = np.sqrt(((Y - Y.mean(0)) ** 2).sum(1).mean())
        steps = np.array([1e-4, 1e-4, 1e-4, 1e-4 * L, 1e-4 * L, 1e-4 * L])
        Hfd = np.zeros((6, 6))
        for a in range(6):
            p = np.zeros(6)
            p[a] = steps[a]
            gp = dg.energy_grad(h, *_left(R, t, p), cov=cov)[1]
            gm = dg.energy_grad(h, *_left(R, t, -p), cov=cov)[1]
            Hfd[:, a] = (gp - gm) / (2 * steps[a])
    finally:
        dg.fr_m(h)
    S = np.diag([1.0, 1.0, 1.0, L, L, L])
    Hs, Fs = S @ H @ S, S @ (0.5 * (Hfd + Hfd.T)) @ S
    assert np.allclose(H, H.T)
    if cov == "iso":
        err = np.abs(Hs - Fs).max() / np.abs(Fs).max()
    else:
        err = np.abs(Hs[3:, 3:] - Fs[3:, 3:]).max() / np.abs(Fs[3:, 3:]).max()
    print("pose Hessian (%s) vs gradient differences: %.2e of the largest entry" % (cov, err))
    assert err < 2e-3, (Hs, Fs)
    assert np.all(np.linalg.eigvalsh(Hs) < 0)  # a maximum of E


def test_pose_hessian_is_3d_only():
    d = W.config5_2d(n=200, m=180)
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    try:
        with pytest.raises(dg.DugmaError, match="3-D"):
            dg.energy_grad_hess(h, np.eye(3), np.zeros(3))
    finally:
        dg.fr_m(h)


def test_register_energy_newton_and_bfgs_agree(monkeypatch):
    """The registration with Newton directions (the default in 3-D) and with
    BFGS alone (DUGMA_REG_NEWTON=0) end at the same pose, the Newton run in
    no more steps."""
    d = W.config3(n=1500, m=1400)
    R0 = W.random_rotation(np.random.default_rng(7), 6.0) @ d["R_G"]
    t0 = np.ravel(d["t_G"]) + np.array([-0.02, 0.01, 0.015])
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    try:
        Rn, tn, itn, En = dg.register_energy(h, R0, t0)
        monkeypatch.setenv("DUGMA_REG_NEWTON", "0")
        Rb, tb, itb, Eb = dg.register_energy(h, R0, t0)
    finally:
        dg.fr_m(h)
    Y = np.asarray(d["Y"]).T
    L = np.sqrt(((Y - Y.mean(0)) ** 2).sum(1).mean())
    print("register_energy: Newton %d steps, BFGS %d steps" % (itn, itb))
    assert W.rotation_error_deg(Rn, Rb) < 5e-3
    assert np.linalg.norm(tn - tb) < 5e-5 * L
    assert abs(En / Eb - 1) < 1e-7
    assert itn <= itb
