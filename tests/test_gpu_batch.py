"""Batched registration on the GPU.

One majorise-minimise step is checked against a numpy restatement (FP64
weights e_ij / s_ij, weighted Procrustes by SVD -- a different solver from
the kernel's Horn quaternion), with and without the sigma update (s_ij =
g (s_i + s_j), g = sig / sbar, sig = mean_i min_j |d_ij|^2 / 3 as in
update_sig.cu:238-260); graph replay, plain launches and one trial per
launch must give identical poses; easy trials must converge; the autotuned
launch shape must not change the result beyond summation order."""
import math

import numpy as np
import pytest

import dugma_amd as dg
from dugma_amd import workloads as W

pytestmark = pytest.mark.gpu


def small_trial(seed, n, m, deg=20.0, noise=0.003, outl=0.0, stretch=(1.0, 1.0, 1.0)):
    rng = np.random.default_rng(seed)
    surf = W.Surface(rng, nt=60, nph=120)
    A = np.diag(stretch)
    src = surf.sample(rng, m) @ A
    tgt0 = surf.sample(rng, n) @ A
    R = W.random_rotation(rng, deg)
    t = rng.uniform(-0.1, 0.1, size=3)
    tgt = tgt0 @ R.T + t + rng.normal(scale=noise, size=(n, 3))
    k = int(round(outl * n))
    if k:
        idx = rng.choice(n, size=k, replace=False)
        tgt[idx] = rng.uniform(tgt.min(0), tgt.max(0), size=(k, 3))
    return dict(X=tgt.T.copy(), sX=W.knn_sigma(tgt) ** 2, Y=src.T.copy(), sY=W.knn_sigma(src) ** 2,
                R_G=R, t_G=t)


def numpy_step(tr, R0, t0, sigma=False, block=1024):
    """One majorise-minimise step in FP64, over blocks of targets: weights
    w_ij = s_ij^-5/2 exp(-|d_ij|^2 / 2 s_ij), weighted Procrustes by SVD."""
    f32 = lambda a: np.asarray(a, np.float64).astype(np.float32).astype(np.float64)  # noqa: E731
    y, x = f32(tr["X"].T), f32(tr["Y"].T)
    sy, sx = f32(tr["sX"]), f32(tr["sY"])
    xp = x @ R0.T + t0
    g = 1.0
    if sigma:  # DUGMA's sig at the start pose, scaled by the mean variances
        mins = np.concatenate([((y[a:a + block, None, :] - xp[None]) ** 2).sum(-1).min(1)
                               for a in range(0, len(y), block)])
        g = max(mins.mean() / 3.0 / (sy.mean() + sx.mean()), 1e-6)
    Wt, wy, wx, Sxy = 0.0, np.zeros(3), np.zeros(3), np.zeros((3, 3))
    for a in range(0, len(y), block):
        yb = y[a:a + block]
        d2 = ((yb[:, None, :] - xp[None]) ** 2).sum(-1)
        s = g * (sy[a:a + block, None] + sx[None, :])
        w = s ** -2.5 * np.exp(-0.5 * d2 / s)
        r, c = w.sum(1), w.sum(0)
        Wt += r.sum()
        wy += r @ yb
        wx += c @ xp
        Sxy += xp.T @ (w.T @ yb)  # sum_ij w_ij x'_j y_i^T
    my, mx = wy / Wt, wx / Wt
    H = Sxy - Wt * np.outer(mx, my)  # sum_ij w_ij (x'_j - mx)(y_i - my)^T
    U, S, Vt = np.linalg.svd(H)
    Dg = np.diag([1.0, 1.0, np.sign(np.linalg.det(Vt.T @ U.T))])
    dR = Vt.T @ Dg @ U.T
    dt = my - dR @ mx
    return dR @ R0, dR @ t0 + dt


@pytest.mark.parametrize("sigma", [False, True])
def test_one_step_matches_procrustes(sigma):
    trials = [small_trial(s, n, m) for s, n, m in ((1, 700, 500), (2, 300, 900), (3, 1500, 1200))]
    b = dg.Batch(trials)
    try:
        rng = np.random.default_rng(0)
        R0 = np.array([W.random_rotation(rng, 5.0) for _ in trials])
        t0 = rng.uniform(-0.02, 0.02, size=(3, 3))
        out = b.register(R0, t0, max_iter=1, anneal=False, use_graph=False, sigma_update=sigma)
    finally:
        b.free()
    for k, tr in enumerate(trials):
        Rr, tt = numpy_step(tr, R0[k], t0[k], sigma)
        assert W.rotation_error_deg(out["R"][k], Rr) < 1e-4
        assert np.abs(out["t"][k] - tt).max() < 1e-6
        assert out["iters"][k] == 1


def test_isotropic_energy_at_returned_pose(oracle):
    """Without a refinement, E is the isotropic energy at the returned pose
    (one pass after the iterations): against the FP64 oracle's with the
    trials' own variances (no sigma update, no annealing: the variance scale
    stays 1)."""
    trials = [small_trial(s, n, m) for s, n, m in ((11, 700, 500), (12, 300, 900), (13, 1500, 1200))]
    b = dg.Batch(trials)
    try:
        out = b.register(max_iter=25, anneal=False, sigma_update=False, coarse_iters=8)
    finally:
        b.free()
    f32 = lambda a: np.asarray(a, np.float64).astype(np.float32).astype(np.float64)  # noqa: E731
    for k, tr in enumerate(trials):
        Y, X = f32(tr["X"].T), f32(tr["Y"].T)
        SY = f32(tr["sX"])[:, None, None] * np.eye(3)
        SX = f32(tr["sY"])[:, None, None] * np.eye(3)
        ref = oracle.energy(Y, SY, X, SX, out["R"][k], out["t"][k])["E"]
        assert abs(out["E"][k] / ref - 1) <= 1e-5, (k, out["E"][k], ref)


@pytest.mark.parametrize("coarse", [0, 9])
def test_modes_give_identical_poses(coarse):
    """Graph replay, plain launches and one trial per launch run the same
    kernels in the same order per trial: bit-identical poses, also across
    the switch from the every-4th-point level to the full clouds."""
    trials = [small_trial(10 + s, 400 + 97 * s, 350 + 55 * s, deg=30.0, outl=0.1) for s in range(6)]
    b = dg.Batch(trials)
    try:
        kw = dict(max_iter=25, anneal=True, coarse_iters=coarse)
        g = b.register(use_graph=True, check_every=5, **kw)
        p = b.register(use_graph=False, **kw)
        o = b.register(one_trial_per_launch=True, **kw)
    finally:
        b.free()
    for a in (p, o):
        np.testing.assert_array_equal(g["R"], a["R"])
        np.testing.assert_array_equal(g["t"], a["t"])
        np.testing.assert_array_equal(g["iters"], a["iters"])


def test_easy_trials_converge():
    """Anisotropic shapes (well-conditioned rotation), 25 degrees: the error
    must drop from the start to a few degrees (the energy maximum of two
    independently sampled clouds is not exactly the ground truth)."""
    trials = [small_trial(20 + s, 1000, 900, deg=25.0, stretch=(1.6, 1.0, 0.6)) for s in range(8)]
    b = dg.Batch(trials)
    try:
        out = b.register(max_iter=100)
    finally:
        b.free()
    start = [W.rotation_error_deg(np.eye(3), tr["R_G"]) for tr in trials]
    errs = [W.rotation_error_deg(out["R"][k], tr["R_G"]) for k, tr in enumerate(trials)]
    assert max(errs) < 5.0, (start, errs)
    assert np.median(errs) < 3.0, errs
    for k, tr in enumerate(trials):
        assert np.linalg.norm(out["t"][k] - tr["t_G"]) < 0.05


def test_coarse_phase_then_full_clouds():
    """The first coarse_iters steps run on every 4th point and never stop a
    trial; the full clouds then reach the same accuracy as without them."""
    trials = [small_trial(40 + s, 1000, 900, deg=25.0, stretch=(1.6, 1.0, 0.6)) for s in range(8)]
    b = dg.Batch(trials)
    try:
        with pytest.raises(dg.DugmaError, match="coarse_iters"):
            b.register(max_iter=10, coarse_iters=-1)
        loose = dict(max_iter=60, anneal=False, tol_rot=1e-2, tol_trans=1e-2)
        quick = b.register(coarse_iters=0, **loose)
        held = b.register(coarse_iters=30, **loose)
        out = b.register(max_iter=130, coarse_iters=30)
        ref = b.register(max_iter=100)
    finally:
        b.free()
    # loose tolerances stop every trial early on the full clouds, never
    # during the coarse phase
    assert (quick["status"] == 1).all() and (quick["iters"] < 30).all(), quick["iters"]
    assert (held["iters"] > 30).all(), held["iters"]
    errs = [W.rotation_error_deg(out["R"][k], tr["R_G"]) for k, tr in enumerate(trials)]
    errs_ref = [W.rotation_error_deg(ref["R"][k], tr["R_G"]) for k, tr in enumerate(trials)]
    assert max(errs) < 5.0, (errs, errs_ref)
    for k, tr in enumerate(trials):
        assert np.linalg.norm(out["t"][k] - tr["t_G"]) < 0.05


def test_no_overlap_is_reported():
    tr = small_trial(30, 300, 300)
    tr["Y"] = tr["Y"] + 1e6  # every weight underflows
    b = dg.Batch([tr, small_trial(31, 300, 300)])
    try:
        out = b.register(max_iter=5, anneal=False, sigma_update=False)
    finally:
        b.free()
    assert out["status"][0] == 3


def test_autotuned_shape_keeps_results():
    """The launch shape only regroups the FP64 sums: poses after 20
    iterations agree to near FP64 round-off of the Horn step."""
    trials = [small_trial(40 + s, 300 + 450 * s, 2600 - 300 * s, deg=20.0) for s in range(7)]
    b = dg.Batch(trials)
    try:
        b.set_shape(4, [1, 1, 2, 4], 8192)
        ref = b.register(max_iter=20)
        shape = b.autotune()
        assert shape["waves"] in (1, 2, 4) and all(c in (1, 2, 4, 8) for c in shape["chunks"])
        assert shape["ms_per_iteration"] > 0
        b.set_shape(1, [8, 8, 8, 8], 4096)  # a different grouping for sure
        alt = b.register(max_iter=20)
    finally:
        b.free()
    for k in range(len(trials)):
        assert W.rotation_error_deg(ref["R"][k], alt["R"][k]) < 1e-3
        assert np.abs(ref["t"][k] - alt["t"][k]).max() < 1e-5


def _rot_about(axis, deg):
    a = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    th = math.radians(deg)
    return np.eye(3) + math.sin(th) * K + (1 - math.cos(th)) * K @ K


def test_config1_energy_maximum(oracle):
    """Config 1 (1,000 x 1,000, isotropic 8-NN sigma, angle up to 45 deg):
    one full 6-DoF maximisation of the convolved energy (a batch of one
    trial without the sigma update, so the objective is exactly E).  At the
    result the FP64 oracle's pose gradient must vanish next to its size one
    degree / 0.01 units away; the rotation and translation errors against
    the ground truth are printed."""
    d = W.config1()
    b = dg.Batch([d])
    try:
        out = b.register(max_iter=2000, anneal=True, sigma_update=False, tol_rot=1e-8, tol_trans=1e-8)
    finally:
        b.free()
    # FP32 per-pair weights leave a noise floor under the step size, so the
    # iteration limit may end the run; stationarity is checked below
    assert out["status"][0] in (1, 2), out
    R, t = out["R"][0], out["t"][0]
    f32 = lambda a: np.asarray(a, np.float64).astype(np.float32).astype(np.float64)  # noqa: E731
    Y, X = f32(d["X"].T), f32(d["Y"].T)
    SY, SX = f32(W.unblocks(d["X_C"])), f32(W.unblocks(d["Y_C"]))
    g = oracle.energy(Y, SY, X, SX, R, t, grads=True)["gpose"]
    g_rot = oracle.energy(Y, SY, X, SX, _rot_about([1, 2, 3], 1.0) @ R, t, grads=True)["gpose"]
    g_tr = oracle.energy(Y, SY, X, SX, R, t + np.array([0.01, -0.01, 0.01]), grads=True)["gpose"]
    assert np.linalg.norm(g[:3]) < 1e-2 * np.linalg.norm(g_rot[:3]), (g, g_rot)
    assert np.linalg.norm(g[3:]) < 1e-2 * np.linalg.norm(g_tr[3:]), (g, g_tr)
    print("config 1: rotation error %.4f deg, translation error %.5f, %d iterations"
          % (W.rotation_error_deg(R, d["R_G"]), np.linalg.norm(t - d["t_G"]), out["iters"][0]))


# A slice of the batched workload itself (4,096 trials, one seed and one
# surface per trial): the smallest and largest clouds, the heaviest outlier
# fractions, and every 40th trial.
C4_SLICE = (0, 5, 10, 40, 68, 80, 81, 84, 110, 111, 120, 124, 127, 160, 163, 168, 181, 196, 200, 207, 210,
            215, 233, 240, 246, 258, 280, 283, 285, 292, 297, 320, 324, 353, 358, 360, 370, 400, 401, 409,
            440, 480, 504, 507, 509)
# first-step check: every slice trial below 4e6 pairs, and the largest in
# both clouds (5: 7,727 x 7,308), the most uneven (163: 7,864 x 1,534;
# 297: 7,912 x 604)
C4_STEP = (0, 5, 68, 80, 81, 110, 111, 120, 124, 163, 181, 207, 210, 215, 233, 240, 258, 280, 297, 353, 358,
           370, 401, 480, 504, 509)
# stationarity: the five smallest (500-930 points per cloud)
C4_SMALL = (210, 215, 258, 280, 111)


@pytest.fixture(scope="module")
def c4_trials():
    return {k: W.config4_trial(k) for k in C4_SLICE}


@pytest.mark.parametrize("sigma", [False, True])
def test_config4_first_step_matches_procrustes(c4_trials, sigma):
    """The first step of every trial from identity (the workload's 0-90
    degree starts, up to 30% outliers), with and without the sigma update,
    against the FP64 blockwise Procrustes step."""
    trials = [c4_trials[k] for k in C4_STEP]
    print("config 4 first step: %d trials, %.3g pairs, sigma=%s"
          % (len(trials), sum(tr["n"] * tr["m"] for tr in trials), sigma), flush=True)
    b = dg.Batch(trials)
    try:
        out = b.register(max_iter=1, anneal=False, use_graph=False, sigma_update=sigma)
    finally:
        b.free()
    for k, tr in enumerate(trials):
        Rr, tt = numpy_step(tr, np.eye(3), np.zeros(3), sigma)
        assert W.rotation_error_deg(out["R"][k], Rr) < 1e-4, (C4_STEP[k], tr["n"], tr["m"])
        assert np.abs(out["t"][k] - tt).max() < 1e-6, (C4_STEP[k], tr["n"], tr["m"])


def test_config4_graph_equals_one_trial_per_launch(c4_trials):
    """All 45 trials of the slice, bench settings (annealing, sigma update,
    over-relaxation, coarse phase): graph replay and one trial per launch
    give bit-identical poses, iteration counts and status."""
    trials = [c4_trials[k] for k in C4_SLICE]
    b = dg.Batch(trials)
    try:
        kw = dict(max_iter=30, anneal=True, anneal_decay=0.95, sigma_update=True, relax=1.9, coarse_iters=12,
                  tol_rot=1e-4, tol_trans=1e-4)
        g = b.register(use_graph=True, check_every=10, **kw)
        o = b.register(one_trial_per_launch=True, **kw)
    finally:
        b.free()
    np.testing.assert_array_equal(g["R"], o["R"])
    np.testing.assert_array_equal(g["t"], o["t"])
    np.testing.assert_array_equal(g["iters"], o["iters"])
    np.testing.assert_array_equal(g["status"], o["status"])


def test_config4_small_trials_stationary(c4_trials, oracle):
    """The five smallest trials maximised to the end without the sigma
    update (the objective is then exactly the convolved energy): the FP64
    oracle's pose gradient at the result vanishes next to its size one
    degree / 0.01 units away."""
    trials = [c4_trials[k] for k in C4_SMALL]
    b = dg.Batch(trials)
    try:
        out = b.register(max_iter=2000, anneal=True, sigma_update=False, tol_rot=1e-8, tol_trans=1e-8)
    finally:
        b.free()
    f32 = lambda a: np.asarray(a, np.float64).astype(np.float32).astype(np.float64)  # noqa: E731
    for k, tr in enumerate(trials):
        assert out["status"][k] in (1, 2), (C4_SMALL[k], out["status"][k])
        R, t = out["R"][k], out["t"][k]
        Y, X = f32(tr["X"].T), f32(tr["Y"].T)
        SY = f32(tr["sX"])[:, None, None] * np.eye(3)
        SX = f32(tr["sY"])[:, None, None] * np.eye(3)
        g = oracle.energy(Y, SY, X, SX, R, t, grads=True)["gpose"]
        g_rot = oracle.energy(Y, SY, X, SX, _rot_about([1, 2, 3], 1.0) @ R, t, grads=True)["gpose"]
        g_tr = oracle.energy(Y, SY, X, SX, R, t + np.array([0.01, -0.01, 0.01]), grads=True)["gpose"]
        print("config 4 trial %d (%d x %d, outliers %.2f): rotation error %.3f deg, %d iterations"
              % (C4_SMALL[k], tr["n"], tr["m"], tr["outliers"], W.rotation_error_deg(R, tr["R_G"]), out["iters"][k]))
        assert np.linalg.norm(g[:3]) < 1e-2 * np.linalg.norm(g_rot[:3]), (C4_SMALL[k], g, g_rot)
        assert np.linalg.norm(g[3:]) < 1e-2 * np.linalg.norm(g_tr[3:]), (C4_SMALL[k], g, g_tr)


# ---- refinement with surface covariances --------------------------------
C4_REFINE = (0, 68, 80, 81, 110, 111, 120, 181, 207, 210, 215, 233, 258, 280, 353, 401, 480, 509)


@pytest.fixture(scope="module")
def c4_cov_trials():
    return {k: W.config4_trial(k, covariances=True) for k in C4_REFINE}


def _oracle_arrays(tr):
    f32 = lambda a: np.asarray(a, np.float64).astype(np.float32).astype(np.float64)  # noqa: E731
    sym = lambda C: f32(0.5 * (C + np.transpose(C, (0, 2, 1))))  # noqa: E731
    return f32(tr["X"].T), sym(tr["CX"]), f32(tr["Y"].T), sym(tr["CY"])


def test_refine_energy_matches_oracle(c4_cov_trials, oracle):
    """The refinement's energy at the pose it returns (full covariances, one
    fused pass per trial pose) against the FP64 oracle's at that pose."""
    trials = [c4_cov_trials[k] for k in C4_REFINE[:8]]
    b = dg.Batch(trials)
    try:
        out = b.register(max_iter=60, anneal=True, sigma_update=True, refine_iters=3)
    finally:
        b.free()
    for k, tr in enumerate(trials):
        Y, SY, X, SX = _oracle_arrays(tr)
        ref = oracle.energy(Y, SY, X, SX, out["R"][k], out["t"][k])["E"]
        assert abs(out["E"][k] / ref - 1) <= 1e-5, (C4_REFINE[k], out["E"][k], ref)


@pytest.mark.parametrize("refine_coarse, refine_mid", [(0, 0), (8, 0), (8, 6), (0, 6)])
def test_refine_graph_equals_one_trial_per_launch(c4_cov_trials, refine_coarse, refine_mid):
    trials = [c4_cov_trials[k] for k in C4_REFINE]
    b = dg.Batch(trials)
    try:
        kw = dict(max_iter=40, anneal=True, sigma_update=True, coarse_iters=10, refine_iters=20,
                  refine_coarse_iters=refine_coarse, refine_mid_iters=refine_mid, tol_rot=1e-4, tol_trans=1e-4)
        g = b.register(use_graph=True, check_every=7, **kw)
        p = b.register(use_graph=False, **kw)
        o = b.register(one_trial_per_launch=True, **kw)
    finally:
        b.free()
    for a in (p, o):
        np.testing.assert_array_equal(g["R"], a["R"])
        np.testing.assert_array_equal(g["t"], a["t"])
        np.testing.assert_array_equal(g["iters"], a["iters"])
        np.testing.assert_array_equal(g["status"], a["status"])


@pytest.mark.parametrize("schedule", ["full", "coarse"])
def test_refine_reaches_stationary_accurate_poses(c4_cov_trials, oracle, schedule):
    """Then the refinement -- after 300 full-cloud isotropic iterations, or
    (the bench's schedule) straight from the coarse isotropic phase through
    60 coarse and 20 middle-level refinement steps: every trial ends where
    the FP64 oracle's pose gradient (surface covariances, all points)
    vanishes next to its size one degree / 0.01 units away, and within 2
    degrees of the ground truth in at least 80% of these trials (500-8,000
    points, up to 30% outliers, 0-90 degree starts)."""
    trials = [c4_cov_trials[k] for k in C4_REFINE]
    sched = (dict(max_iter=460, refine_coarse_iters=0) if schedule == "full"
             else dict(max_iter=120, refine_coarse_iters=60, refine_mid_iters=20))
    b = dg.Batch(trials)
    try:
        out = b.register(anneal=True, anneal_decay=0.95, sigma_update=True, relax=1.9, coarse_iters=160,
                         tol_rot=1e-5, tol_trans=1e-5, refine_iters=200, **sched)
    finally:
        b.free()
    errs = []
    for k, tr in enumerate(trials):
        R, t = out["R"][k], out["t"][k]
        e = W.rotation_error_deg(R, tr["R_G"])
        errs.append(e)
        print("config 4 trial %d (%d x %d, outliers %.2f): refined rotation error %.3f deg, status %d, %d steps"
              % (C4_REFINE[k], tr["n"], tr["m"], tr["outliers"], e, out["status"][k], out["iters"][k]))
        assert out["status"][k] in (1, 2)
        Y, SY, X, SX = _oracle_arrays(tr)
        g = oracle.energy(Y, SY, X, SX, R, t, grads=True)["gpose"]
        g_rot = oracle.energy(Y, SY, X, SX, _rot_about([1, 2, 3], 1.0) @ R, t, grads=True)["gpose"]
        g_tr = oracle.energy(Y, SY, X, SX, R, t + np.array([0.01, -0.01, 0.01]), grads=True)["gpose"]
        assert np.linalg.norm(g[:3]) < 1e-2 * np.linalg.norm(g_rot[:3]), (C4_REFINE[k], g, g_rot)
        assert np.linalg.norm(g[3:]) < 1e-2 * np.linalg.norm(g_tr[3:]), (C4_REFINE[k], g, g_tr)
    assert np.mean(np.array(errs) < 2.0) >= 0.8, errs


def test_far_field_skip_changes_nothing_measurable(c4_cov_trials, monkeypatch):
    """The refinement's pair passes let a wave skip a block of 64 sources
    when a bound (box distance squared over the largest eigenvalues of the
    two covariances) puts every term below 7e-13 of its own peak.  The same
    registrations with the skip off (DUGMA_FAR_SKIP=0 when the covariances
    are set) take the same steps to the same poses and energies, to 1e-9."""
    trials = [c4_cov_trials[k] for k in C4_REFINE]
    kw = dict(max_iter=160, anneal=True, anneal_decay=0.95, sigma_update=True, relax=1.9, coarse_iters=160,
              refine_iters=100, refine_coarse_iters=60, refine_mid_iters=20, tol_rot=1e-4, tol_trans=1e-4)
    out = {}
    for skip in ("1", "0"):
        monkeypatch.setenv("DUGMA_FAR_SKIP", skip)
        b = dg.Batch(trials)  # the covariances are set here
        try:
            out[skip] = b.register(**kw)
        finally:
            b.free()
    a, o = out["1"], out["0"]
    np.testing.assert_array_equal(a["iters"], o["iters"])
    np.testing.assert_array_equal(a["status"], o["status"])
    for k in range(len(trials)):
        assert np.abs(a["R"][k] - o["R"][k]).max() < 1e-9, k
        assert np.abs(a["t"][k] - o["t"][k]).max() < 1e-9, k
        assert abs(a["E"][k] / o["E"][k] - 1) < 1e-9, (k, a["E"][k], o["E"][k])


def test_far_field_skip_edge_sizes(monkeypatch, oracle):
    """Trials smaller than one 64-source block, at block boundaries (64,
    65, 128, 129 points) and with duplicated points (equal spatial keys),
    isotropic covariances: with and without the skip, the same steps,
    poses and energies (1e-9), and the energy at the returned pose against
    the FP64 oracle."""
    rng = np.random.default_rng(11)
    sizes = [(40, 50), (64, 64), (65, 129), (128, 65), (129, 300), (300, 128)]
    trials = []
    for s0, (n, m) in enumerate(sizes):
        tr = small_trial(300 + s0, n, m, deg=float(rng.uniform(5.0, 30.0)))
        if s0 == 5:  # every source twice
            tr["Y"] = np.concatenate([tr["Y"], tr["Y"]], axis=1)
            tr["sY"] = np.concatenate([tr["sY"], tr["sY"]])
        trials.append(tr)
    kw = dict(max_iter=60, anneal=True, anneal_decay=0.9, sigma_update=True, coarse_iters=30, refine_iters=40,
              refine_coarse_iters=10, refine_mid_iters=5, tol_rot=1e-6, tol_trans=1e-6)
    out = {}
    for skip in ("1", "0"):
        monkeypatch.setenv("DUGMA_FAR_SKIP", skip)
        b = dg.Batch(trials)
        try:
            b.set_covariances([np.eye(3) * tr["sX"][:, None, None] for tr in trials],
                              [np.eye(3) * tr["sY"][:, None, None] for tr in trials])
            out[skip] = b.register(**kw)
        finally:
            b.free()
    a, o = out["1"], out["0"]
    np.testing.assert_array_equal(a["iters"], o["iters"])
    np.testing.assert_array_equal(a["status"], o["status"])
    np.testing.assert_allclose(a["R"], o["R"], rtol=0, atol=1e-9)
    np.testing.assert_allclose(a["t"], o["t"], rtol=0, atol=1e-9)
    np.testing.assert_allclose(a["E"], o["E"], rtol=1e-9, atol=0)
    f32 = lambda v: np.asarray(v, np.float64).astype(np.float32).astype(np.float64)  # noqa: E731
    for k, tr in enumerate(trials):
        Y, X = f32(tr["X"].T), f32(tr["Y"].T)
        SY = f32(tr["sX"])[:, None, None] * np.eye(3)
        SX = f32(tr["sY"])[:, None, None] * np.eye(3)
        ref = oracle.energy(Y, SY, X, SX, a["R"][k], a["t"][k])["E"]
        assert abs(a["E"][k] / ref - 1) <= 1e-5, (sizes[k], a["E"][k], ref)


def test_level_shapes_copied_give_identical_results(c4_cov_trials):
    """Each level has its own launch shape; a second batch of the same trials
    given the same per-level shapes (set_level_shape) reproduces the first
    one's results bit for bit, one trial per launch against the graph."""
    trials = [c4_cov_trials[k] for k in C4_REFINE[:10]]
    kw = dict(max_iter=30, anneal=True, sigma_update=True, coarse_iters=20, refine_iters=15,
              refine_coarse_iters=6, refine_mid_iters=4, tol_rot=1e-4, tol_trans=1e-4)
    b = dg.Batch(trials)
    try:
        b.set_level_shape(1, 2, [1, 2, 2, 4], 4096)
        b.set_level_shape(2, 1, [2, 1, 1, 1], 8192)
        assert b.get_shape(1) == dict(waves=2, chunks=[1, 2, 2, 4], grid_waves=4096)
        assert b.get_shape(2) == dict(waves=1, chunks=[2, 1, 1, 1], grid_waves=8192)
        shapes = [b.get_shape(l) for l in range(3)]
        g = b.register(use_graph=True, check_every=4, **kw)
    finally:
        b.free()
    b1 = dg.Batch(trials)
    try:
        for l, sh in enumerate(shapes):
            b1.set_level_shape(l, sh["waves"], sh["chunks"], sh["grid_waves"])
        o = b1.register(one_trial_per_launch=True, **kw)
    finally:
        b1.free()
    for key in ("R", "t", "iters", "status"):
        np.testing.assert_array_equal(g[key], o[key])


def test_tail_queue_keeps_results():
    """Once at most a sixteenth of the trials run, passes use a queue of the
    running trials' items only (rebuilt as that count halves); 48 trials of
    uneven sizes that finish at different passes must give exactly the
    poses, step counts and status of one trial per launch, in the isotropic
    phase (full clouds) and in both refinement levels."""
    rng = np.random.default_rng(7)
    trials = []
    for s in range(48):
        tr = small_trial(200 + s, int(rng.integers(150, 700)), int(rng.integers(150, 700)),
                         deg=float(rng.uniform(2.0, 40.0)))
        trials.append(tr)
    b = dg.Batch(trials)
    try:
        b.set_covariances([np.tile(np.eye(3), (tr["X"].shape[1], 1, 1)) * tr["sX"][:, None, None] for tr in trials],
                          [np.tile(np.eye(3), (tr["Y"].shape[1], 1, 1)) * tr["sY"][:, None, None] for tr in trials])
        kw = dict(max_iter=150, anneal=False, sigma_update=True, tol_rot=1e-6, tol_trans=1e-6,
                  refine_iters=40, refine_coarse_iters=10)
        g = b.register(use_graph=True, check_every=3, **kw)
        o = b.register(one_trial_per_launch=True, **kw)
    finally:
        b.free()
    assert len(set(g["iters"].tolist())) > 8  # the trials finish at many different passes
    np.testing.assert_array_equal(g["R"], o["R"])
    np.testing.assert_array_equal(g["t"], o["t"])
    np.testing.assert_array_equal(g["iters"], o["iters"])
    np.testing.assert_array_equal(g["status"], o["status"])


def test_refine_argument_errors():
    tr = small_trial(50, 300, 300)
    b = dg.Batch([tr])
    try:
        with pytest.raises(dg.DugmaError, match="set_covariances"):
            b.register(max_iter=3, refine_iters=2)
        C = np.tile(np.eye(3), (300, 1, 1)) * 1e-3
        bad = C.copy()
        bad[7] = -bad[7]
        with pytest.raises(dg.DugmaError, match="positive definite"):
            b.set_covariances([bad], [C])
        b.set_covariances([C], [C])
        with pytest.raises(dg.DugmaError, match="refine_iters"):
            b.register(max_iter=3, refine_iters=-1)
        with pytest.raises(dg.DugmaError, match="refine_coarse_iters"):
            b.register(max_iter=3, refine_iters=2, refine_coarse_iters=-1)
        with pytest.raises(dg.DugmaError, match="refine_mid_iters"):
            b.register(max_iter=3, refine_iters=2, refine_mid_iters=-1)
    finally:
        b.free()
