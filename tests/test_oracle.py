"""Pins the CPU oracle (what the HIP path is tested against) against closed forms,
library routines, brute force and the reference's own fixtures."""
import math
import os

import numpy as np
import pytest
from scipy.stats import multivariate_normal

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def rot3(w):
    th = np.linalg.norm(w)
    K = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    return np.eye(3) + math.sin(th) / th * K + (1 - math.cos(th)) / th ** 2 * K @ K


def rot2(a):
    return np.array([[math.cos(a), -math.sin(a)], [math.sin(a), math.cos(a)]])


def spd(rng, n, D, s=0.2, floor=0.02):
    A = rng.normal(size=(n, D, D)) * s
    return A @ np.transpose(A, (0, 2, 1)) + floor * np.eye(D)


# ---------------------------------------------------------------- init
@pytest.mark.parametrize("D", [2, 3])
def test_prep_inverse_closed_forms(oracle, D):
    rng = np.random.default_rng(3)
    C = np.concatenate([np.eye(D)[None], (np.diag(np.arange(1.0, D + 1)))[None], spd(rng, 20, D, 1.0, 0.5)])
    inv, g, bad = oracle.prep_inverse(C, D)
    assert bad == 0
    Ci = np.transpose(inv.reshape(-1, D, D), (0, 2, 1))  # column-major per point
    np.testing.assert_array_equal(Ci[0], np.eye(D, dtype=np.float32))
    np.testing.assert_allclose(np.diag(Ci[1]), 1.0 / np.arange(1.0, D + 1), rtol=1e-7)
    for k in range(len(C)):
        np.testing.assert_allclose(Ci[k] @ C[k], np.eye(D), atol=5e-6)
    np.testing.assert_allclose(g, 1.0 / np.sqrt(np.linalg.det(C)), rtol=5e-6)
    assert g[0] == 1.0


@pytest.mark.parametrize("D", [2, 3])
def test_prep_inverse_fallback(oracle, D):
    # det > 0 is the only test (init.cu:39,73): -I passes in 2-D, diag(-1,1,..) does not
    C = np.stack([np.eye(D), np.diag([-1.0] + [1.0] * (D - 1)), np.full((D, D), np.nan), np.zeros((D, D))])
    inv, g, bad = oracle.prep_inverse(C, D)
    assert bad == 3
    assert (inv[1:] == 100000.0).all() and (g[1:] == 0).all()


def test_prep_inverse_kinect_fixtures(oracle):
    g = np.load(os.path.join(GOLD, "kinect_fixtures.npz"))
    nan_points = 0
    for k in (2, 11, 22):
        C = g["s%d_X_C" % k]
        n = C.shape[1] // 3
        B = np.transpose(C.reshape(3, n, 3), (1, 0, 2))
        inv, gg, bad = oracle.prep_inverse(B, 3)
        CY = g["s%d_Y_C" % k]
        BY = np.transpose(CY.reshape(3, CY.shape[1] // 3, 3), (1, 0, 2))
        inv2, gg2, bad2 = oracle.prep_inverse(BY, 3)
        nan_points += bad + bad2
        isnan = ~np.isfinite(B).all(axis=(1, 2))
        assert (gg[isnan] == 0).all()
    assert nan_points == 3  # one NaN covariance in each of scenes 2, 11, 22


# ---------------------------------------------------------------- update_sig
def _numpy_sig(Xf, Yf, R, t):
    """update_sig in numpy with the same FP32 operation order."""
    D = Xf.shape[1]
    m = np.empty_like(Yf)
    for a in range(D):
        s = R[a, 0] * Yf[:, 0].astype(np.float64)
        for b in range(1, D):
            s = s + R[a, b] * Yf[:, b].astype(np.float64)
        m[:, a] = (s + t[a]).astype(np.float32)
    e = Xf[:, None, :] - m[None, :, :]
    s2 = e[..., 0] * e[..., 0] + e[..., 1] * e[..., 1]
    if D == 3:
        s2 = (s2 + e[..., 2] * e[..., 2]).astype(np.float32)
    s = (s2 / np.float32(D)).astype(np.float32) * np.float32(1e6)
    s = s.astype(np.float32)
    u = np.where(s > 0, np.minimum(np.rint(s.astype(np.float64)), 4294967295.0), 0).astype(np.uint64)
    q = np.minimum(u.min(axis=1), 4000000000).astype(np.uint32)
    f = q.astype(np.float32) / np.float32(1e6)
    return q, f


@pytest.mark.parametrize("D", [2, 3])
def test_update_sig_brute_force(oracle, D):
    rng = np.random.default_rng(5 + D)
    Xf = rng.normal(size=(333, D)).astype(np.float32)
    Yf = rng.normal(size=(257, D)).astype(np.float32)
    R = rot3([0.3, 0.1, -0.2]) if D == 3 else rot2(0.4)
    t = rng.normal(size=D) * 0.1
    sig, q = oracle.update_sig(Xf, Yf, R, t, with_q=True)
    q2, f2 = _numpy_sig(Xf, Yf, R, t)
    np.testing.assert_array_equal(q, q2)
    assert sig == np.float32(oracle.tree_sum_f32(f2) / np.float32(333))


def test_update_sig_closed_forms(oracle):
    rng = np.random.default_rng(9)
    X = rng.normal(size=(100, 3)).astype(np.float32)
    assert oracle.update_sig(X, X, np.eye(3), np.zeros(3)) == 0.0
    # one point, shifted by (0.001, 0.002, -0.002): |d|^2/3 = 3e-6 exactly after quantisation
    a = np.array([[0.5, 0.25, 0.125]], np.float32)
    b = a - np.array([[0.001, 0.002, -0.002]], np.float32)
    s = oracle.update_sig(a, b, np.eye(3), np.zeros(3))
    assert abs(s - 3e-6) < 1e-12
    # far points saturate at the 4e9 cap -> 4000
    far = oracle.update_sig(a, b + 1e4, np.eye(3), np.zeros(3))
    assert far == np.float32(4000.0)


def test_tree_sum_order(oracle):
    # the pairwise tree differs from a left-to-right sum on this input
    v = np.array([1.0, 1e8, -1e8, 1.0, 3.0], np.float32)
    # level 1: (1+1e8, -1e8+1, 3) ; level 2: (1e8 + -99999999 , 3) ; level 3: sum
    l1 = [np.float32(v[0] + v[1]), np.float32(v[2] + v[3]), v[4]]
    l2 = [np.float32(l1[0] + l1[1]), l1[2]]
    expect = np.float32(l2[0] + l2[1])
    assert oracle.tree_sum_f32(v) == expect


# ---------------------------------------------------------------- k-NN
@pytest.mark.parametrize("D", [2, 3])
def test_knn_brute_force_with_ties(oracle, D):
    rng = np.random.default_rng(11)
    P = rng.integers(0, 6, size=(400, D)).astype(np.float32)  # many exact ties
    Q = rng.integers(0, 6, size=(150, D)).astype(np.float32)
    idx, d2 = oracle.knn(Q, P, 8)
    for i in range(len(Q)):
        # integer coordinates: every product and sum below is exact in FP32,
        # so the fused and unfused orders give the same distances
        dx = (Q[i] - P).astype(np.float64)
        dd = (dx ** 2).sum(axis=1).astype(np.float32)
        order = np.lexsort((np.arange(len(P)), dd))[:8]
        np.testing.assert_array_equal(idx[i], order)
        np.testing.assert_array_equal(d2[i], dd[order])


def test_knn_self_exclusion(oracle):
    rng = np.random.default_rng(12)
    P = rng.normal(size=(50, 3)).astype(np.float32)
    idx, _ = oracle.knn(P, P, 4, self_exclude=True)
    assert not (idx == np.arange(50)[:, None]).any()
    idx2, d2 = oracle.knn(P[:3], P[:3], 4, self_exclude=True)
    assert (idx2[:, 2:] == -1).all() and np.isinf(d2[:, 2:]).all()


# ---------------------------------------------------------------- convolved energy
@pytest.mark.parametrize("D", [2, 3])
def test_energy_matches_library_pdf(oracle, D):
    rng = np.random.default_rng(20 + D)
    N, M = 6, 5
    Y = rng.normal(size=(N, D)) * 0.3
    X = rng.normal(size=(M, D)) * 0.3
    SY, SX = spd(rng, N, D), spd(rng, M, D)
    R = rot3([0.3, -0.2, 0.5]) if D == 3 else rot2(0.7)
    t = rng.normal(size=D) * 0.1
    out = oracle.energy(Y, SY, X, SX, R, t, per_target=True)
    ref = np.array([[multivariate_normal.pdf(Y[i], mean=R @ X[j] + t, cov=SY[i] + R @ SX[j] @ R.T)
                     for j in range(M)] for i in range(N)])
    assert abs(out["E"] / ref.sum() - 1) < 1e-13
    np.testing.assert_allclose(out["eT"], ref.sum(axis=1), rtol=1e-13)


def test_energy_isotropic_closed_form(oracle):
    rng = np.random.default_rng(30)
    Y = rng.normal(size=(7, 3))
    X = rng.normal(size=(9, 3))
    sy, sx = rng.uniform(0.1, 0.5, 7), rng.uniform(0.1, 0.5, 9)
    E = oracle.energy(Y, sy[:, None, None] * np.eye(3), X, sx[:, None, None] * np.eye(3), np.eye(3), np.zeros(3))["E"]
    s = sy[:, None] + sx[None, :]
    d2 = ((Y[:, None, :] - X[None, :, :]) ** 2).sum(-1)
    ref = ((2 * math.pi * s) ** -1.5 * np.exp(-d2 / (2 * s))).sum()
    assert abs(E / ref - 1) < 1e-13


def test_energy_invariances(oracle):
    rng = np.random.default_rng(31)
    Y, X = rng.normal(size=(8, 3)), rng.normal(size=(6, 3))
    SY, SX = spd(rng, 8, 3), spd(rng, 6, 3)
    R, t = rot3([0.2, 0.4, -0.1]), rng.normal(size=3)
    E = oracle.energy(Y, SY, X, SX, R, t)["E"]
    # one rigid motion applied to both clouds (and the covariances)
    G, g = rot3([-0.5, 0.3, 0.9]), rng.normal(size=3)
    Y2 = Y @ G.T + g
    SY2 = np.einsum("ab,nbc,dc->nad", G, SY, G)
    E2 = oracle.energy(Y2, SY2, X, SX, G @ R, G @ t + g)["E"]
    assert abs(E2 / E - 1) < 1e-12
    p, q = rng.permutation(8), rng.permutation(6)
    E3 = oracle.energy(Y[p], SY[p], X[q], SX[q], R, t)["E"]
    assert abs(E3 / E - 1) < 1e-13


def test_energy_is_gmm_product_integral(oracle):
    """N(a-b; 0, A+B) = integral N(z; a, A) N(z; b, B) dz, by 2-D quadrature."""
    rng = np.random.default_rng(32)
    a, b = rng.normal(size=2) * 0.3, rng.normal(size=2) * 0.3
    A, B = spd(rng, 1, 2, 0.3, 0.05)[0], spd(rng, 1, 2, 0.3, 0.05)[0]
    E = oracle.energy(a[None], A[None], b[None], B[None], np.eye(2), np.zeros(2))["E"]
    g = np.linspace(-4, 4, 801)
    Z = np.stack(np.meshgrid(g, g, indexing="ij"), -1).reshape(-1, 2)
    f = multivariate_normal.pdf(Z, a, A) * multivariate_normal.pdf(Z, b, B)
    quad = f.sum() * (g[1] - g[0]) ** 2
    assert abs(quad / E - 1) < 1e-6


@pytest.mark.parametrize("D", [2, 3])
def test_energy_gradients_central_differences(oracle, D):
    rng = np.random.default_rng(40 + D)
    N, M = 7, 6
    Y, X = rng.normal(size=(N, D)) * 0.4, rng.normal(size=(M, D)) * 0.4
    SY, SX = spd(rng, N, D), spd(rng, M, D)
    R = rot3([0.3, -0.2, 0.5]) if D == 3 else rot2(0.4)
    t = rng.normal(size=D) * 0.1
    out = oracle.energy(Y, SY, X, SX, R, t, grads=True)
    _, gfd = oracle.energy_fd_pose(Y, SY, X, SX, R, t, h=1e-5)
    np.testing.assert_allclose(out["gpose"], gfd, rtol=1e-7, atol=1e-9 * np.abs(gfd).max())
    h = 1e-6

    def fd(arr_name, arr, idx):
        a1, a2 = arr.copy(), arr.copy()
        a1[idx] += h
        a2[idx] -= h
        kw = dict(Y=Y, SY=SY, X=X, SX=SX)
        kw[arr_name] = a1
        e1 = oracle.energy(kw["Y"], kw["SY"], kw["X"], kw["SX"], R, t)["E"]
        kw[arr_name] = a2
        e2 = oracle.energy(kw["Y"], kw["SY"], kw["X"], kw["SX"], R, t)["E"]
        return (e1 - e2) / (2 * h)

    for i in range(N):
        for a in range(D):
            assert abs(fd("Y", Y, (i, a)) - out["gY"][i, a]) < 1e-7 * np.abs(out["gY"]).max()
    for j in range(M):
        for a in range(D):
            assert abs(fd("X", X, (j, a)) - out["gX"][j, a]) < 1e-7 * np.abs(out["gX"]).max()
        for a in range(D):
            for b in range(D):
                assert abs(fd("SX", SX, (j, a, b)) - out["gSX"][j, a, b]) < 1e-7 * np.abs(out["gSX"]).max()
    for i in range(N):
        for a in range(D):
            for b in range(D):
                assert abs(fd("SY", SY, (i, a, b)) - out["gSY"][i, a, b]) < 1e-7 * np.abs(out["gSY"]).max()


def test_energy_gradients_torch_autograd(oracle):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(44)
    N, M = 5, 4
    Y, X = rng.normal(size=(N, 3)) * 0.4, rng.normal(size=(M, 3)) * 0.4
    SY, SX = spd(rng, N, 3), spd(rng, M, 3)
    R, t = rot3([0.1, 0.2, 0.3]), rng.normal(size=3) * 0.1
    out = oracle.energy(Y, SY, X, SX, R, t, grads=True)
    tY = torch.tensor(Y, requires_grad=True)
    tSY = torch.tensor(SY, requires_grad=True)
    tX = torch.tensor(X, requires_grad=True)
    tSX = torch.tensor(SX, requires_grad=True)
    tR, tt = torch.tensor(R), torch.tensor(t)
    xp = tX @ tR.T + tt
    Sxp = tR @ tSX @ tR.T
    S = tSY[:, None] + Sxp[None]
    d = tY[:, None] - xp[None]
    q = (d[..., None, :] @ torch.linalg.solve(S, d[..., None]))[..., 0, 0]
    E = ((2 * math.pi) ** -1.5 * torch.linalg.det(S) ** -0.5 * torch.exp(-0.5 * q)).sum()
    E.backward()
    assert abs(E.item() / out["E"] - 1) < 1e-13
    np.testing.assert_allclose(tY.grad.numpy(), out["gY"], rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(tX.grad.numpy(), out["gX"], rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(tSY.grad.numpy(), out["gSY"], rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(tSX.grad.numpy(), out["gSX"], rtol=1e-10, atol=1e-12)


# ---------------------------------------------------------------- compi
@pytest.mark.parametrize("D", [3, 2])
def test_compi_golden_from_reference_expressions(oracle, D):
    g = np.load(os.path.join(GOLD, "compi_golden.npz"))
    p = "d%d_" % D
    coef, missing = oracle.compi(g[p + "Xf"], g[p + "XCinv"], g[p + "gX"], g[p + "Yf"], g[p + "YCinv"],
                                 g[p + "gY"], g[p + "R"], g[p + "t"], float(g[p + "sig"]))
    assert missing == 0
    ref = g[p + "coef"]
    assert np.abs(coef - ref).max() <= 1e-12 * np.abs(ref).max()


def _compi_problem(rng, D, N=6, M=5):
    Xf = (rng.normal(size=(N, D)) * 0.5).astype(np.float32)
    Yf = (rng.normal(size=(M, D)) * 0.5).astype(np.float32)
    XCi, gX, _ = oracle_mod().prep_inverse(spd(rng, N, D, 0.3), D)
    YCi, gY, _ = oracle_mod().prep_inverse(spd(rng, M, D, 0.3), D)
    R = rot3([0.1, 0.2, -0.1]) if D == 3 else rot2(0.2)
    t = rng.normal(size=D) * 0.1
    return Xf, XCi, gX, Yf, YCi, gY, R, t


def oracle_mod():
    from oracle import oracle as O
    return O


@pytest.mark.parametrize("D", [3, 2])
def test_compi_polynomial_identity(oracle, D):
    rng = np.random.default_rng(50 + D)
    args = _compi_problem(rng, D)
    coef, missing = oracle.compi(*args, 0.7)
    assert missing == 0
    for _ in range(25):
        dT = rng.normal(size=12 if D == 3 else 4)
        F1 = oracle.poly_eval(coef, dT)
        F2 = oracle.surrogate_eval(*args, 0.7, dT)
        assert abs(F1 - F2) <= 1e-12 * abs(F2)


@pytest.mark.parametrize("D", [3, 2])
def test_compi_moments_equal_pairwise_expansion(oracle, D):
    """The moment form (FP64 moments per source, then one polynomial per
    source) gives the coefficients of the pair-by-pair expansion, which the
    golden vectors from compi.cu's own expressions pin, coefficient by
    coefficient, on the golden problem and on 300 x 250 random clouds with
    non-symmetric-rounded inverses."""
    g = np.load(os.path.join(GOLD, "compi_golden.npz"))
    p = "d%d_" % D
    args = [g[p + k] for k in ("Xf", "XCinv", "gX", "Yf", "YCinv", "gY", "R", "t")]
    cm, missing = oracle.compi_moments(*args, float(g[p + "sig"]))
    assert missing == 0
    ref = g[p + "coef"]
    big = np.abs(ref) > 1e-9 * np.abs(ref).max()
    np.testing.assert_allclose(cm[big], ref[big], rtol=1e-11)
    rng = np.random.default_rng(60 + D)
    Xf, XCi, gX, Yf, YCi, gY, R, t = _compi_problem(rng, D, N=300, M=250)
    c1, _ = oracle.compi(Xf, XCi, gX, Yf, YCi, gY, R, t, 0.05)
    c2, missing = oracle.compi_moments(Xf, XCi, gX, Yf, YCi, gY, R, t, 0.05)
    assert missing == 0
    big = np.abs(c1) > 1e-9 * np.abs(c1).max()
    np.testing.assert_allclose(c2[big], c1[big], rtol=1e-10)
    assert np.abs(c2 - c1).max() <= 1e-13 * np.abs(c1).max()


def test_compi_identity_pose_value(oracle):
    """At dT = identity, F = sum P (n - m')^T (A' + B) (n - m')."""
    rng = np.random.default_rng(55)
    Xf, XCi, gX, Yf, YCi, gY, R, t = _compi_problem(rng, 3)
    coef, _ = oracle.compi(Xf, XCi, gX, Yf, YCi, gY, R, t, 0.9)
    dT = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1.0])
    F = oracle.poly_eval(coef, dT)
    Isig = np.float32(1 / 0.9)
    s3 = np.sqrt(np.float32(Isig * Isig * Isig), dtype=np.float32)
    ref = 0.0
    for i in range(len(Xf)):
        B = np.transpose(XCi[i].reshape(3, 3)).astype(np.float32) * Isig
        for j in range(len(Yf)):
            A = np.transpose(YCi[j].reshape(3, 3)).astype(np.float32) * Isig
            P = oracle.pair_weight(Xf[i], XCi[i], gX[i], Yf[j], YCi[j], gY[j], R, t, 0.9)
            mp = (R @ Yf[j].astype(np.float64) + t).astype(np.float32).astype(np.float64)
            Ap = (R @ A.astype(np.float64).T @ R.T).astype(np.float32).astype(np.float64)
            dd = Xf[i].astype(np.float64) - mp
            ref += P * (dd @ (Ap + B.astype(np.float64)) @ dd)
    assert abs(F / ref - 1) < 1e-6  # A' and m' rounded independently here


def test_constraints(oracle):
    R = rot3([0.3, -0.4, 0.2])
    dT = np.concatenate([np.zeros(3), R.ravel()])
    assert np.abs(oracle.constraints(dT)).max() < 1e-14
    dT2 = dT.copy()
    dT2[3] += 0.1
    assert np.abs(oracle.constraints(dT2)).max() > 1e-3
    assert abs(oracle.constraints(np.array([0, 0, 0.6, 0.8]))[0]) < 1e-15


# ---------------------------------------------------------------- per-point sigma
@pytest.mark.parametrize("D", [3, 2])
def test_knn_sigma_lattice_closed_forms(oracle, D):
    """Interior points of a cubic (square) lattice of spacing h: the 8
    nearest other points are the 2D face neighbours at h and the rest at
    h sqrt 2, so sigma = (2D h + (8 - 2D) h sqrt 2) / 8; with k = 2D it is h.
    A point of a cloud with fewer than k other points gets +inf.  (Counting
    the point itself, using squared distances or dividing by k + 1 all fail
    these.)"""
    h = 0.37
    g = np.stack(np.meshgrid(*[np.arange(6)] * D, indexing="ij"), -1).reshape(-1, D) * h
    inner = np.all((g > 0.5 * h) & (g < 4.5 * h), axis=1)
    s8 = oracle.knn_sigma(g, 8)
    ref = (2 * D * h + (8 - 2 * D) * h * math.sqrt(2)) / 8
    np.testing.assert_allclose(s8[inner], ref, rtol=1e-6)
    np.testing.assert_allclose(oracle.knn_sigma(g, 2 * D)[inner], h, rtol=1e-6)
    q = np.nonzero(inner)[0][:5]
    np.testing.assert_array_equal(oracle.knn_sigma(g, 8, queries=q), s8[q])
    assert np.isinf(oracle.knn_sigma(g[:5], 8)).all()


def test_knn_sigma_follows_knn_lists(oracle):
    """On a random cloud with planted duplicates the sigma of every point is
    the FP32 mean of the square roots of its dor_knn distances, nearest
    first."""
    rng = np.random.default_rng(71)
    P = rng.normal(size=(400, 3)).astype(np.float32)
    P[10:14] = P[3]
    idx, d2 = oracle.knn(P, P, 8, True)
    acc = np.sqrt(d2[:, 0])
    for r in range(1, 8):
        acc = (acc + np.sqrt(d2[:, r])).astype(np.float32)
    np.testing.assert_array_equal(oracle.knn_sigma(P, 8), (acc / np.float32(8)).astype(np.float32))


# ---------------------------------------------------------------- the DUGMA.m loop
def _rz(a):
    return np.array([[math.cos(a), -math.sin(a), 0], [math.sin(a), math.cos(a), 0], [0, 0, 1.0]])


def _ry(a):
    return np.array([[math.cos(a), 0, math.sin(a)], [0, 1.0, 0], [-math.sin(a), 0, math.cos(a)]])


def _rx(a):
    return np.array([[1.0, 0, 0], [0, math.cos(a), -math.sin(a)], [0, math.sin(a), math.cos(a)]])


def test_rotm2eul_zyx_textbook(oracle):
    """The convergence test's angles (DUGMA.m:80, rotm2eul's default ZYX):
    R = Rz(z) Ry(y) Rx(x) gives back (z, y, x) for |y| < 90 degrees."""
    rng = np.random.default_rng(70)
    for _ in range(200):
        z, x = rng.uniform(-math.pi, math.pi, 2)
        y = rng.uniform(-1.5, 1.5)
        ang = oracle.rotm2eul_zyx(_rz(z) @ _ry(y) @ _rx(x))
        np.testing.assert_allclose(ang, [z, y, x], atol=1e-12)


def _surface_points(rng, n):
    from dugma_amd import workloads as W
    s = W.Surface(rng, nt=40, nph=80)
    return s.sample(rng, n) @ np.diag([1.6, 1.0, 0.6])


def _noise_free(D, rng, deg):
    from dugma_amd import workloads as W
    if D == 3:
        X = (_surface_points(rng, 400) * 2.0 + np.array([1.0, -2.0, 0.5])).T
        R_G = W.rot_axis_angle([1.0, 2.0, -1.0], math.radians(deg))
        t_G = np.array([0.4, -0.3, 0.25])
    else:
        X = (W.contour2d(rng, 300) * 2.0 + np.array([1.0, -2.0])).T
        R_G = W.rot2(math.radians(deg))
        t_G = np.array([0.4, -0.3])
    Y = R_G.T @ (X - t_G[:, None])
    C = W.blocks(np.tile(2.0 * np.eye(D), (X.shape[1], 1, 1)))
    return X, Y, C, R_G, t_G


@pytest.mark.parametrize("D", [3, 2])
def test_dugma_loop_noise_free_recovers_truth(oracle, D):
    """Identical points and covariances, Y = R_G^T (X - t_G), started 25
    degrees off: the loop (normalise, sigma, compi, surrogate minimiser,
    R <- dR R, t <- t + dt, denormalise) ends at (R_G, t_G) up to the point
    where DUGMA's sigma stops it: sigma is the mean of the nearest squared
    distances / D quantised to 1e-6 (update_sig.cu:256-259), which is 0 <
    accuracy_sig once every point is within sqrt(D * 5e-7) of its twin
    (DUGMA.m:57) -- about 1e-3 of the unit cloud."""
    X, Y, C, R_G, t_G = _noise_free(D, np.random.default_rng(80 + D), 25.0)
    R, t, it = oracle.dugma_loop(X, Y, C, C, np.eye(D), np.zeros(D), 300, 1e-10, 1e-18, 1e-12, moments=True)
    x, y, xd, yd, s = oracle.normalize(X, Y)
    Rn, tn = R, (t - xd + R @ yd) / s  # the loop's last pose in normalised units
    assert oracle.update_sig(x.T, y.T, Rn, tn) == 0.0
    assert it < 300
    from dugma_amd import workloads as W
    err = W.rotation_error_deg(R, R_G) if D == 3 else abs(math.degrees(math.atan2(R[1, 0], R[0, 0]) -
                                                                     math.atan2(R_G[1, 0], R_G[0, 0])))
    assert err < 0.05
    assert np.linalg.norm(t - t_G) < 1e-3 * s


@pytest.mark.parametrize("D", [3, 2])
def test_dugma_loop_at_truth_denormalises(oracle, D):
    """Started at the truth in normalised units, t_n = (R_G yd + t_G - xd) / s,
    sigma is 0 at once, so no iteration runs and the loop returns R0 and
    t = xd + t_n s - R yd (DUGMA.m:122) = t_G."""
    X, Y, C, R_G, t_G = _noise_free(D, np.random.default_rng(85 + D), 25.0)
    x, y, xd, yd, s = oracle.normalize(X, Y)
    tn = (R_G @ yd + t_G - xd) / s
    R, t, it = oracle.dugma_loop(X, Y, C, C, R_G, tn, 300, moments=True)
    assert it == 0
    np.testing.assert_array_equal(R, R_G)
    np.testing.assert_allclose(t, t_G, rtol=0, atol=1e-13 * np.abs(X).max())


def test_dugma_loop_one_iteration_composition(oracle):
    """One iteration of DUGMA.m:57-123 put together independently: numpy
    normalisation (normalize.m), the oracle's sigma and compi, and scipy's
    SLSQP over the twelve fmincon variables under fcontr3D.m's equality
    constraints (an fmincon-like constrained solver, not the loop's BFGS on
    SO(3)); then R <- dR R, t <- t + dt (DUGMA.m:76-78) and t = xd + t s -
    R yd (DUGMA.m:122).  The start has a large translation, so composing as
    t <- dR t + dt would miss by far more than the tolerance."""
    from scipy.optimize import minimize
    from dugma_amd import workloads as W
    rng = np.random.default_rng(90)
    P = _surface_points(rng, 500)
    R_G = W.rot_axis_angle([0.3, -1.0, 0.5], math.radians(20.0))
    X = (P[:250] @ R_G.T + np.array([0.2, 0.1, -0.3]) + rng.normal(scale=0.01, size=(250, 3))).T
    Y = P[250:].T
    var = 0.02
    XC = W.blocks(np.tile(var * np.eye(3), (250, 1, 1)))
    R0 = W.rot_axis_angle([1.0, 0.0, 1.0], math.radians(12.0))
    t0 = np.array([0.6, -0.4, 0.3])
    R1, t1, it = oracle.dugma_loop(X, Y, XC, XC, R0, t0, 1, moments=True)
    assert it == 1
    # the same iteration from its parts
    xd, yd = X.mean(1), Y.mean(1)
    x, y = X - xd[:, None], Y - yd[:, None]
    s = max(math.sqrt((x ** 2).sum() / 250), math.sqrt((y ** 2).sum() / 250))
    Xf, Yf = (x / s).T.astype(np.float32), (y / s).T.astype(np.float32)
    Ci, g, _ = oracle.prep_inverse(np.tile(var * np.eye(3), (250, 1, 1)), 3)
    sig = oracle.update_sig(Xf, Yf, R0, t0)
    coef, _ = oracle.compi_moments(Xf, Ci, g, Yf, Ci, g, R0, t0, sig)
    scale = abs(oracle.poly_eval(coef, np.r_[0, 0, 0, np.eye(3).ravel()]))
    # fcontr3D.m's ten equalities are redundant (R^T R - I is symmetric, and
    # det R = 1 follows near the identity); SLSQP needs independent ones:
    # the upper triangle of R^T R - I
    upper = [1, 4, 5, 7, 8, 9]
    res = minimize(lambda z: oracle.poly_eval(coef, z) / scale, np.r_[0, 0, 0, np.eye(3).ravel()],
                   method="SLSQP", constraints=[{"type": "eq", "fun": lambda z: oracle.constraints(z)[upper]}],
                   options={"ftol": 1e-16, "maxiter": 1000})
    assert res.success
    assert np.abs(oracle.constraints(res.x)).max() < 1e-9
    dT = res.x
    dR, dt = dT[3:].reshape(3, 3), dT[:3]
    R_ref = dR @ R0
    t_ref = xd + (t0 + dt) * s - R_ref @ yd
    np.testing.assert_allclose(R1, R_ref, atol=2e-6)
    np.testing.assert_allclose(t1, t_ref, atol=2e-6 * s)
    t_wrong = xd + (dR @ t0 + dt) * s - R_ref @ yd
    assert np.linalg.norm(t_wrong - t_ref) > 1e3 * 2e-6 * s


def _full_scene(k):
    g = np.load(os.path.join(GOLD, "kinect_full_scenes.npz"))
    out = {"X": g["s%d_X" % k], "Y": g["s%d_Y" % k], "R_G": g["s%d_R_G" % k], "t_G": g["s%d_t_G" % k]}
    for name in ("X", "Y"):
        v = g["s%d_v%s" % (k, name)]
        C = np.zeros((3, 3 * len(v)))
        for a in range(3):
            C[a, a::3] = v
        out[name + "_C"] = C
    return out


@pytest.mark.parametrize("k", [8, 20])
def test_dugma_loop_reproduces_published_kinect_outcomes(oracle, k):
    """The whole loop on whole Kinect scenes with the reference's driver
    settings (main_kinect_real_application.m:14-21) lands on the pose the
    reference published for the scene (Our_result_in_Kinect_Application/
    Ours_k_result.mat): within 0.01 degrees and 1e-4 of the translation.
    This pins the loop end to end against the reference's own outputs:
    normalisation, sigma, compi, the minimiser, the pose composition and the
    denormalisation (a wrong one of these moves the result by far more)."""
    from dugma_amd import workloads as W
    d = _full_scene(k)
    R, t, it = oracle.dugma_loop(d["X"], d["Y"], d["X_C"], d["Y_C"], np.eye(3), np.zeros(3), 300, 1e-4, 1e-8,
                                 1e-9, moments=True)
    g = np.load(os.path.join(GOLD, "kinect_fixtures.npz"))
    Rp, tp = g["published_R"][k - 1], g["published_t"][k - 1]
    assert W.rotation_error_deg(R, Rp) < 0.01
    assert np.linalg.norm(t - tp) < 1e-4 * np.linalg.norm(d["t_G"])



def test_knn_cov_against_numpy_and_plane(oracle):
    """dor_knn_cov: the same neighbours as dor_knn, their FP64 sample
    covariance (numpy) to FP32 rounding, and on a planar patch the normal
    direction holds exactly the floor: C_zz = floor sigma^2, C_xz = C_yz = 0."""
    rng = np.random.default_rng(7)
    P = rng.normal(size=(400, 3)).astype(np.float32)
    sig, C = oracle.knn_cov(P, 8, 0.5, 0.002)
    idx, d2 = oracle.knn(P, P, 8, self_exclude=True)
    np.testing.assert_array_equal(sig, oracle.knn_sigma(P, 8))
    nb = P.astype(np.float64)[idx]
    mu = nb.mean(1, keepdims=True)
    ref = 0.5 * np.einsum("nka,nkb->nab", nb - mu, nb - mu) / 8 + (0.002 * sig.astype(np.float64) ** 2)[:, None, None] * np.eye(3)
    np.testing.assert_allclose(C, ref, rtol=2e-5, atol=1e-6 * np.abs(ref).max())
    g = np.stack(np.meshgrid(np.arange(12), np.arange(12), indexing="ij"), -1).reshape(-1, 2) * 0.25
    Q = np.concatenate([g + rng.uniform(-0.05, 0.05, size=g.shape), np.full((len(g), 1), 2.0)], 1).astype(np.float32)
    s2, C2 = oracle.knn_cov(Q, 8, 0.5, 0.002)
    np.testing.assert_array_equal(C2[:, 2, 2], np.float32(0.002) * (s2 * s2))
    assert (C2[:, 0, 2] == 0).all() and (C2[:, 1, 2] == 0).all()
