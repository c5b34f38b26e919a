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


def test_2d_loop_runs(oracle):
    d = W.config5_2d(n=300, m=250)
    R, t, it, sig = dg.DUGMA(d["X"], d["Y"], d["X_C"], d["Y_C"], np.eye(2), np.zeros(2), 3)
    Rc, tc, itc = oracle.dugma_loop(d["X"], d["Y"], d["X_C"], d["Y_C"], np.eye(2), np.zeros(2), 3)
    assert it == itc
    assert abs(np.arctan2(R[1, 0], R[0, 0]) - np.arctan2(Rc[1, 0], Rc[0, 0])) < 1e-5
    assert np.abs(t - tc).max() < 1e-5
