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


# per-coefficient bar at full size: P_ij is FP32 on the GPU (two exp2 of
# FP32 quadratic forms of the hi/lo pair difference), FP64 in the reference
# (compi.cu:747-772); coefficients are P-weighted sums, so each carries about
# the per-pair relative error of P (a few FP32 ulps times the exponent).
FULL_REL = 2e-5


@pytest.mark.parametrize("pose", ["truth", "start"])
def test_compi_full_size_per_coefficient(oracle, pose):
    """6,000 x 6,000 structured-light clouds (the size bench.py times compi
    on), sigma from update_sig at the pose: every coefficient above 1e-8 of
    the largest within FULL_REL of the oracle's FP64 moment form (pinned
    against the pair-by-pair expansion and compi.cu's own expressions), and
    the polynomial's value at random updates near identity within FULL_REL
    of the size of its terms."""
    d = W.config2(n=6000, m=6000)
    if pose == "truth":
        R, t = d["R_G"], np.ravel(d["t_G"])
    else:
        rng = np.random.default_rng(3)
        R, t = W.random_rotation(rng, 2.0) @ d["R_G"], np.ravel(d["t_G"]) + rng.uniform(-0.003, 0.003, size=3)
    st = device_state(oracle, d)
    sig = oracle.update_sig(st[0], st[3], R, t)
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    try:
        assert dg.update_sig(h, R, t, h.N, h.M, 3) == sig
        coef = dg.compi(h, R, t, h.N, h.M, 3, sig)
    finally:
        dg.fr_m(h)
    ref, missing = oracle.compi_moments(*st, R, t, sig)
    assert missing == 0
    big = np.abs(ref) > 1e-8 * np.abs(ref).max()
    rel = np.abs(coef[big] - ref[big]) / np.abs(ref[big])
    print("compi 6K (%s pose): %d of 637 coefficients above the floor; per-coefficient relative error "
          "max %.3g, median %.3g, 99th percentile %.3g"
          % (pose, big.sum(), rel.max(), np.median(rel), np.percentile(rel, 99)))
    assert rel.max() <= FULL_REL, (np.argmax(rel), rel.max())
    rng = np.random.default_rng(11)
    ident = np.concatenate([np.zeros(3), np.eye(3).ravel()])
    for _ in range(5):
        dT = ident + rng.normal(size=12) * np.r_[np.full(3, 1e-3), np.full(9, 1e-3)]
        F1 = oracle.poly_eval(coef, dT)
        F2 = oracle.poly_eval(ref, dT)
        mag = oracle.poly_eval(np.abs(ref), np.abs(dT))
        assert abs(F1 - F2) <= FULL_REL * mag


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
