"""Batched registration on the GPU.

One majorise-minimise step is checked against a numpy restatement (FP64
weights e_ij / s_ij, weighted Procrustes by SVD -- a different solver from
the kernel's Horn quaternion); graph replay, plain launches and one trial
per launch must give identical poses; easy trials must converge."""
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


def numpy_step(tr, R0, t0):
    f32 = lambda a: np.asarray(a, np.float64).astype(np.float32).astype(np.float64)  # noqa: E731
    y, x = f32(tr["X"].T), f32(tr["Y"].T)
    sy, sx = f32(tr["sX"]), f32(tr["sY"])
    xp = x @ R0.T + t0
    d2 = ((y[:, None, :] - xp[None, :, :]) ** 2).sum(-1)
    s = sy[:, None] + sx[None, :]
    e = s ** -1.5 * np.exp(-0.5 * d2 / s)
    w = e / s
    Wt = w.sum()
    my = (w.sum(1) @ y) / Wt
    mx = (w.sum(0) @ xp) / Wt
    H = (xp - mx).T @ w.T @ (y - my)  # sum_ij w_ij (x'_j - mx)(y_i - my)^T
    U, S, Vt = np.linalg.svd(H)
    Dg = np.diag([1.0, 1.0, np.sign(np.linalg.det(Vt.T @ U.T))])
    dR = Vt.T @ Dg @ U.T
    dt = my - dR @ mx
    return dR @ R0, dR @ t0 + dt


def test_one_step_matches_procrustes():
    trials = [small_trial(s, n, m) for s, n, m in ((1, 700, 500), (2, 300, 900), (3, 1500, 1200))]
    b = dg.Batch(trials)
    try:
        rng = np.random.default_rng(0)
        R0 = np.array([W.random_rotation(rng, 5.0) for _ in trials])
        t0 = rng.uniform(-0.02, 0.02, size=(3, 3))
        out = b.register(R0, t0, max_iter=1, anneal=False, use_graph=False)
    finally:
        b.free()
    for k, tr in enumerate(trials):
        Rr, tt = numpy_step(tr, R0[k], t0[k])
        assert W.rotation_error_deg(out["R"][k], Rr) < 1e-4
        assert np.abs(out["t"][k] - tt).max() < 1e-6
        assert out["iters"][k] == 1


def test_modes_give_identical_poses():
    trials = [small_trial(10 + s, 400 + 97 * s, 350 + 55 * s, deg=30.0, outl=0.1) for s in range(6)]
    b = dg.Batch(trials)
    try:
        kw = dict(max_iter=25, anneal=True)
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


def test_no_overlap_is_reported():
    tr = small_trial(30, 300, 300)
    tr["Y"] = tr["Y"] + 1e6  # every weight underflows
    b = dg.Batch([tr, small_trial(31, 300, 300)])
    try:
        out = b.register(max_iter=5, anneal=False)
    finally:
        b.free()
    assert out["status"][0] == 3
