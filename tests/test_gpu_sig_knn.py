"""GPU parity of init (the six reference handles), update_sig and k-NN:
bit-exact against the oracle.  For init and update_sig the oracle restates
the reference's FP32 operations in source order with IEEE semantics (no
contraction, correctly rounded division and sqrtf); the reference's own
binaries are built with -use_fast_math (compile.m:2), so equality with them
is not established here."""
import os

import math

import numpy as np
import pytest

import dugma_amd as dg
from dugma_amd import workloads as W

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def kinect(k):
    g = np.load(os.path.join(HERE, "golden", "kinect_fixtures.npz"))
    return dict(X=g["s%d_X" % k], X_C=g["s%d_X_C" % k], Y=g["s%d_Y" % k], Y_C=g["s%d_Y_C" % k],
                R_G=g["s%d_R_G" % k], t_G=g["s%d_t_G" % k].ravel())


@pytest.mark.parametrize("k", [2, 20])
def test_init_handles_bit_exact(oracle, k):
    d = kinect(k)
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    try:
        assert ok == 1.0
        np.testing.assert_array_equal(h.read_handle(0), d["X"].T.astype(np.float32).ravel())
        np.testing.assert_array_equal(h.read_handle(2), d["Y"].T.astype(np.float32).ravel())
        for which, gwhich, P in ((1, 4, d["X_C"]), (3, 5, d["Y_C"])):
            inv, g, _ = oracle.prep_inverse(W.unblocks(P), 3)
            np.testing.assert_array_equal(h.read_handle(which), inv.ravel())
            np.testing.assert_array_equal(h.read_handle(gwhich), g)
    finally:
        dg.fr_m(h)


def test_init_handles_2d(oracle):
    d = W.config5_2d(n=500, m=400)
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    try:
        inv, g, _ = oracle.prep_inverse(W.unblocks(d["X_C"]), 2)
        np.testing.assert_array_equal(h.read_handle(1), inv.ravel())
        np.testing.assert_array_equal(h.read_handle(4), g)
    finally:
        dg.fr_m(h)


@pytest.mark.parametrize("k", [2, 11, 20, 22])
def test_update_sig_kinect_bit_exact(oracle, k):
    d = kinect(k)
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    try:
        for R, t in ((np.eye(3), np.zeros(3)), (d["R_G"], d["t_G"])):
            s = dg.update_sig(h, R, t, h.N, h.M, 3)
            ref = oracle.update_sig(d["X"].T, d["Y"].T, R, t)
            assert s == ref
    finally:
        dg.fr_m(h)


def test_update_sig_2d_and_mismatch(oracle):
    d = W.config5_2d(n=2000, m=1700)
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    try:
        s = dg.update_sig(h, d["R_G"], d["t_G"], h.N, h.M, 2)
        assert s == oracle.update_sig(d["X"].T, d["Y"].T, d["R_G"], d["t_G"])
        with pytest.raises(dg.DugmaError):
            dg.update_sig(h, d["R_G"], d["t_G"], h.N + 1, h.M, 2)
    finally:
        dg.fr_m(h)


def test_update_sig_larger(oracle):
    d = W.config3(n=30000, m=20000)
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    try:
        s = dg.update_sig(h, d["R_G"], d["t_G"], h.N, h.M, 3)
        assert s == oracle.update_sig(d["X"].T, d["Y"].T, d["R_G"], d["t_G"])
    finally:
        dg.fr_m(h)


@pytest.mark.parametrize("which", [dg.KNN_TARGET_SELF, dg.KNN_SOURCE_SELF, dg.KNN_TARGET_TO_SOURCE])
def test_knn_bit_exact(oracle, which):
    d = W.config3(n=12000, m=9000)
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    try:
        R, t = d["R_G"], d["t_G"]
        idx, d2 = dg.knn(h, which, 8, R, t)
        Xf, Yf = d["X"].T.astype(np.float32), d["Y"].T.astype(np.float32)
        if which == dg.KNN_TARGET_SELF:
            ri, rd = oracle.knn(Xf, Xf, 8, True)
        elif which == dg.KNN_SOURCE_SELF:
            ri, rd = oracle.knn(Yf, Yf, 8, True)
        else:
            ri, rd = oracle.knn(Xf, oracle.transform_points(Yf, R, t), 8, False)
        np.testing.assert_array_equal(idx, ri)
        np.testing.assert_array_equal(d2, rd)
    finally:
        dg.fr_m(h)


@pytest.mark.parametrize("k", [1, 8, 13, 32])
def test_knn_ties_lowest_index(oracle, k):
    rng = np.random.default_rng(k)
    X = rng.integers(0, 5, size=(3, 3000)).astype(float)  # many equal distances
    C = W.iso_blocks(np.full(3000, 0.01))
    ok, h = dg.init(X, C, X[:, :1000], C[:, :3000])
    try:
        idx, d2 = dg.knn(h, dg.KNN_TARGET_SELF, k)
        ri, rd = oracle.knn(X.T.astype(np.float32), X.T.astype(np.float32), k, True)
        np.testing.assert_array_equal(idx, ri)
        np.testing.assert_array_equal(d2, rd)
    finally:
        dg.fr_m(h)


# The cases below hold at least 2^24 query x candidate pairs, where the
# library searches a uniform grid instead of all pairs; the results must be
# the all-pairs ones, bit for bit.

@pytest.mark.parametrize("k", [1, 8, 13, 32])
def test_knn_grid_ties_lowest_index(oracle, k):
    rng = np.random.default_rng(100 + k)
    X = rng.integers(0, 10, size=(3, 5000)).astype(float)  # ~5 points per lattice site
    C = W.iso_blocks(np.full(5000, 0.01))
    ok, h = dg.init(X, C, X, C)
    try:
        idx, d2 = dg.knn(h, dg.KNN_TARGET_SELF, k)
        Xf = X.T.astype(np.float32)
        ri, rd = oracle.knn(Xf, Xf, k, True)
        np.testing.assert_array_equal(idx, ri)
        np.testing.assert_array_equal(d2, rd)
    finally:
        dg.fr_m(h)


def test_knn_grid_disjoint_clouds(oracle):
    """Every query far outside the candidates' box."""
    d = W.config3(n=6000, m=5000)
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    try:
        R = d["R_G"]
        t = np.ravel(d["t_G"]) + np.array([4.0, -1.5, 0.5])
        idx, d2 = dg.knn(h, dg.KNN_TARGET_TO_SOURCE, 8, R, t)
        Xf, Yf = d["X"].T.astype(np.float32), d["Y"].T.astype(np.float32)
        ri, rd = oracle.knn(Xf, oracle.transform_points(Yf, R, t), 8, False)
        np.testing.assert_array_equal(idx, ri)
        np.testing.assert_array_equal(d2, rd)
    finally:
        dg.fr_m(h)


def test_knn_grid_non_finite_and_flat(oracle):
    """Non-finite points (never neighbours; a non-finite query gets -1 /
    +inf) and a flat cloud (zero extent along z)."""
    rng = np.random.default_rng(7)
    X = rng.random((3, 6000))
    X[2] = 0.25
    X[:, 10] = np.nan
    X[1, 20] = np.inf
    X[:, 30] = 1e6
    C = W.iso_blocks(np.full(6000, 0.01))
    ok, h = dg.init(X, C, X, C)
    try:
        idx, d2 = dg.knn(h, dg.KNN_TARGET_SELF, 8)
        Xf = X.T.astype(np.float32)
        ri, rd = oracle.knn(Xf, Xf, 8, True)
        np.testing.assert_array_equal(idx, ri)
        np.testing.assert_array_equal(d2, rd)
        assert (idx[10] == -1).all() and np.isinf(d2[10]).all()
    finally:
        dg.fr_m(h)


def test_knn_grid_2d(oracle):
    rng = np.random.default_rng(8)
    X = rng.random((2, 5000))
    Y = rng.random((2, 4500)) * 0.8 + 0.1
    CX = W.blocks(np.tile(0.01 * np.eye(2), (5000, 1, 1)))
    CY = W.blocks(np.tile(0.01 * np.eye(2), (4500, 1, 1)))
    ok, h = dg.init(X, CX, Y, CY)
    try:
        th = 0.3
        R = np.array([[math.cos(th), -math.sin(th)], [math.sin(th), math.cos(th)]])
        t = np.array([0.05, -0.02])
        idx, d2 = dg.knn(h, dg.KNN_TARGET_TO_SOURCE, 16, R, t)
        ri, rd = oracle.knn(X.T.astype(np.float32), oracle.transform_points(Y.T.astype(np.float32), R, t), 16,
                            False)
        np.testing.assert_array_equal(idx, ri)
        np.testing.assert_array_equal(d2, rd)
        idx, d2 = dg.knn(h, dg.KNN_SOURCE_SELF, 8)
        ri, rd = oracle.knn(Y.T.astype(np.float32), Y.T.astype(np.float32), 8, True)
        np.testing.assert_array_equal(idx, ri)
        np.testing.assert_array_equal(d2, rd)
    finally:
        dg.fr_m(h)


def test_knn_full_size_config3(oracle):
    """Config 3 at its full size (200,000 x 200,000, k = 8): every row holds
    valid indices sorted by (distance, index), self mode never returns the
    query itself, and 256 sampled queries are bit-exact against the oracle's
    brute force (self mode through k + 1 neighbours with the query dropped)."""
    d = W.config3()
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    try:
        R, t = d["R_G"], d["t_G"]
        idx, d2 = dg.knn(h, dg.KNN_TARGET_TO_SOURCE, 8, R, t)
        idx_s, d2_s = dg.knn(h, dg.KNN_TARGET_SELF, 8)
    finally:
        dg.fr_m(h)
    Xf = d["X"].T.astype(np.float32)
    Yt = oracle.transform_points(d["Y"].T.astype(np.float32), R, t)
    N, M = Xf.shape[0], Yt.shape[0]
    for I, D2, ncand in ((idx, d2, M), (idx_s, d2_s, N)):
        assert I.min() >= 0 and I.max() < ncand
        ordered = (D2[:, 1:] > D2[:, :-1]) | ((D2[:, 1:] == D2[:, :-1]) & (I[:, 1:] > I[:, :-1]))
        assert ordered.all()
    assert not (idx_s == np.arange(N)[:, None]).any()
    q = np.random.default_rng(0).choice(N, 256, replace=False)
    ri, rd = oracle.knn(Xf[q], Yt, 8, False)
    np.testing.assert_array_equal(idx[q], ri)
    np.testing.assert_array_equal(d2[q], rd)
    ri, rd = oracle.knn(Xf[q], Xf, 9, False)
    for r, qi in enumerate(q):
        keep = ri[r] != qi
        assert (~keep).sum() == 1
        np.testing.assert_array_equal(idx_s[qi], ri[r][keep][:8])
        np.testing.assert_array_equal(d2_s[qi], rd[r][keep][:8])


@pytest.mark.parametrize("n", [1, 63, 64, 65, 4097, 70000])
def test_update_sig_tree_sizes(oracle, n):
    """The pairwise tree of update_sig at sizes around its 64-element groups
    (one, a partial group, whole groups, a carried element, three levels of
    groups), bit-exact against the oracle."""
    d = W.config3(n=max(n, 2), m=300)
    X = d["X"][:, :n].copy()
    XC = W.blocks(W.unblocks(d["X_C"])[:n])
    rng = np.random.default_rng(n)
    R = W.random_rotation(rng, 3.0) @ d["R_G"]
    t = np.ravel(d["t_G"]) + rng.uniform(-0.01, 0.01, 3)
    ok, h = dg.init(X, XC, d["Y"], d["Y_C"])
    try:
        s = dg.update_sig(h, R, t, h.N, h.M, 3)
    finally:
        dg.fr_m(h)
    assert s == oracle.update_sig(X.T, d["Y"].T, R, t)
