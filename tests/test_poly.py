"""The surrogate polynomial on the host (no GPU): value and analytic
gradient against the oracle, and the fmincon replacement's minimiser."""
import math

import numpy as np
import pytest

import dugma_amd as dg


def rot3(w):
    th = np.linalg.norm(w)
    K = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    return np.eye(3) + math.sin(th) / th * K + (1 - math.cos(th)) / th ** 2 * K @ K


def surrogate(oracle, D, n=50, m=45, rot=None, t=None, noise=0.0, seed=0, sig=0.05):
    rng = np.random.default_rng(seed)
    Y = rng.normal(size=(m, D)) * 0.5
    R = rot if rot is not None else np.eye(D)
    tt = t if t is not None else np.zeros(D)
    X = (Y[:n] if n <= m else np.concatenate([Y, Y[: n - m]])) @ R.T + tt + rng.normal(size=(n, D)) * noise
    C = np.tile(np.eye(D) * 0.01, (max(n, m), 1, 1))
    XCi, gX, _ = oracle.prep_inverse(C[:n], D)
    YCi, gY, _ = oracle.prep_inverse(C[:m], D)
    coef, miss = oracle.compi(X.astype(np.float32), XCi, gX, Y.astype(np.float32), YCi, gY,
                              np.eye(D), np.zeros(D), sig)
    assert miss == 0
    return coef


@pytest.mark.parametrize("D", [3, 2])
def test_poly_value_and_gradient(oracle, D):
    coef = surrogate(oracle, D, rot=rot3([0.05, -0.02, 0.03]) if D == 3 else None)
    rng = np.random.default_rng(1)
    for _ in range(5):
        dT = rng.normal(size=12 if D == 3 else 4)
        F, g = dg.poly_eval(coef, dT, grad=True)
        assert abs(F - oracle.poly_eval(coef, dT)) <= 1e-12 * abs(F)
        h = 1e-6
        for v in range(dT.size):
            e = np.zeros(dT.size)
            e[v] = h
            fd = (oracle.poly_eval(coef, dT + e) - oracle.poly_eval(coef, dT - e)) / (2 * h)
            assert abs(fd - g[v]) <= 1e-6 * np.abs(g).max()


def test_minimizer_recovers_known_increment(oracle):
    Rt = rot3(np.array([0.04, -0.06, 0.05]))
    tt = np.array([0.03, -0.02, 0.01])
    coef = surrogate(oracle, 3, n=60, m=60, rot=Rt, t=tt, sig=0.02)
    dT, F = dg.poly_minimize(coef)
    dR = dT[3:].reshape(3, 3)
    np.testing.assert_allclose(dR @ dR.T, np.eye(3), atol=1e-10)
    assert abs(np.linalg.det(dR) - 1) < 1e-10
    assert F <= oracle.poly_eval(coef, np.concatenate([np.zeros(3), np.eye(3).ravel()]))
    ang = math.degrees(math.acos(min(1, (np.trace(dR @ Rt.T) - 1) / 2)))
    assert ang < 0.5, ang
    assert np.linalg.norm(dT[:3] - tt) < 0.01
    # first-order optimality on the constraint set: no small rotation or
    # translation lowers F
    for a in range(6):
        for sgn in (1, -1):
            p = np.zeros(6)
            p[a] = sgn * 1e-4
            dR2 = rot3(p[:3]) @ dR if a < 3 else dR
            dT2 = np.concatenate([dT[:3] + p[3:], dR2.ravel()])
            assert oracle.poly_eval(coef, dT2) >= F - 1e-9 * abs(F)


def test_minimizer_2d_matches_grid_search(oracle):
    """The 2-D surrogate is a partial (majorise-minimise) step: its minimum
    is found by brute force over (theta, dt) and must match."""
    c, s = math.cos(0.1), math.sin(0.1)
    coef = surrogate(oracle, 2, n=50, m=50, rot=np.array([[c, -s], [s, c]]), t=np.array([0.02, 0.01]), sig=0.02)
    dT, F = dg.poly_minimize(coef)
    assert abs(dT[2] ** 2 + dT[3] ** 2 - 1) < 1e-12
    th = math.atan2(dT[3], dT[2])
    best = None
    for a in np.linspace(th - 0.02, th + 0.02, 41):
        for tx in np.linspace(dT[0] - 0.01, dT[0] + 0.01, 21):
            for ty in np.linspace(dT[1] - 0.01, dT[1] + 0.01, 21):
                v = oracle.poly_eval(coef, [tx, ty, math.cos(a), math.sin(a)])
                if best is None or v < best[0]:
                    best = (v, a, tx, ty)
    assert F <= best[0] + 1e-9 * abs(best[0])
    assert abs(best[1] - th) <= 1e-3 and abs(best[2] - dT[0]) <= 1e-3 and abs(best[3] - dT[1]) <= 1e-3
