"""GPU parity of compi (the 637 / 45 surrogate coefficients) against the
oracle's pair-by-pair FP64 expansion of the same surrogate.

Per-pair weights are FP32 on the GPU (two v_exp each) and the moments FP64,
so the bar is 1e-5 relative to the largest coefficient, and the same for the
polynomial's value F(dT) at random updates."""
import os

import numpy as np
import pytest

import dugma_amd as dg
from dugma_amd import workloads as W

pytestmark = pytest.mark.gpu
TOL = 1e-5
HERE = os.path.dirname(os.path.abspath(__file__))


def device_state(oracle, d):
    D = d["X"].shape[0]
    Xf, Yf = d["X"].T.astype(np.float32), d["Y"].T.astype(np.float32)
    XCi, gX, _ = oracle.prep_inverse(W.unblocks(d["X_C"]), D)
    YCi, gY, _ = oracle.prep_inverse(W.unblocks(d["Y_C"]), D)
    return Xf, XCi, gX, Yf, YCi, gY


def check(oracle, d, R, t, sig, seed=0):
    D = d["X"].shape[0]
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    try:
        coef = dg.compi(h, R, t, h.N, h.M, D, sig)
    finally:
        dg.fr_m(h)
    st = device_state(oracle, d)
    ref, missing = oracle.compi(*st, R, t, sig)
    assert missing == 0
    scale = np.abs(ref).max()
    err = np.abs(coef - ref).max() / scale
    assert err <= TOL, err
    rng = np.random.default_rng(seed)
    nv = 12 if D == 3 else 4
    ident = np.concatenate([np.zeros(3), np.eye(3).ravel()]) if D == 3 else np.array([0, 0, 1.0, 0])
    for _ in range(5):
        dT = ident + rng.normal(size=nv) * 0.05
        F1 = oracle.poly_eval(coef, dT)
        F2 = oracle.surrogate_eval(*st, R, t, sig, dT)
        # F is a sum of large terms of both signs: measure against their size
        mag = oracle.poly_eval(np.abs(ref), np.abs(dT))
        assert abs(F1 - F2) <= TOL * mag
    return coef


def test_compi_3d_random(oracle):
    d = W.config1(n=220, m=180)
    R = W.random_rotation(np.random.default_rng(1), 10.0) @ d["R_G"]
    check(oracle, d, R, d["t_G"], 0.8)


def test_compi_3d_kinect_with_fallback_points(oracle):
    g = np.load(os.path.join(HERE, "golden", "kinect_fixtures.npz"))
    for k in (2, 22):
        d = dict(X=g["s%d_X" % k], X_C=g["s%d_X_C" % k], Y=g["s%d_Y" % k], Y_C=g["s%d_Y_C" % k])
        # DUGMA.m:32 normalises the points (not the covariances) before init
        X, Y = d["X"], d["Y"]
        xd, yd = X.mean(axis=1, keepdims=True), Y.mean(axis=1, keepdims=True)
        sc = max(np.sqrt(((X - xd) ** 2).sum() / X.shape[1]), np.sqrt(((Y - yd) ** 2).sum() / Y.shape[1]))
        d["X"], d["Y"] = (X - xd) / sc, (Y - yd) / sc
        check(oracle, d, np.eye(3), np.zeros(3), 0.05)


def test_compi_2d(oracle):
    d = W.config5_2d(n=300, m=260)
    check(oracle, d, d["R_G"], d["t_G"], 0.01)
    check(oracle, d, np.eye(2), np.zeros(2), 0.3)


def test_compi_dims_mismatch():
    d = W.config1(n=50, m=40)
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    try:
        with pytest.raises(dg.DugmaError):
            dg.compi(h, np.eye(3), np.zeros(3), h.N, h.M + 1, 3, 1.0)
        with pytest.raises(dg.DugmaError):
            dg.compi(h, np.eye(3), np.zeros(3), h.N, h.M, 3, -1.0)
    finally:
        dg.fr_m(h)
