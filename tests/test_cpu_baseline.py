"""The CPU baseline (oracle/dugma_cpu_baseline.cpp, timed by bench.py) computes
what the oracle computes: energy and pose gradient (FP64 and FP32 per pair),
central differences, update_sig's nearest-distance sigma and compi's 637
coefficients in both forms."""
import math

import numpy as np
import pytest

from oracle import cpu_baseline as CB


def rot3(w):
    th = np.linalg.norm(w)
    K = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    return np.eye(3) + math.sin(th) / th * K + (1 - math.cos(th)) / th ** 2 * K @ K


def spd(rng, n, s=0.02, floor=1e-4):
    A = rng.normal(size=(n, 3, 3)) * s
    return A @ np.transpose(A, (0, 2, 1)) + floor * np.eye(3)


def clouds(rng, N=300, M=260):
    Y = rng.normal(size=(N, 3)) * 0.1
    X = rng.normal(size=(M, 3)) * 0.1
    return Y, spd(rng, N), X, spd(rng, M)


R0 = rot3([0.1, -0.2, 0.05])
T0 = np.array([0.01, -0.02, 0.015])


@pytest.mark.parametrize("prec,etol,gtol", [(64, 1e-12, 1e-10), (32, 2e-6, 2e-5)])
def test_energy_full_covariance(oracle, prec, etol, gtol):
    rng = np.random.default_rng(1)
    Y, SY, X, SX = clouds(rng)
    ref = oracle.energy(Y, SY, X, SX, R0, T0, grads=True)
    E, g, ts = CB.energy(Y, SY, X, SX, R0, T0, mode="fused", prec=prec, reps=1)
    assert ts > 0
    assert abs(E / ref["E"] - 1) <= etol
    assert np.linalg.norm(g - ref["gpose"]) <= gtol * np.linalg.norm(ref["gpose"])
    Ef, gf, _ = CB.energy(Y, SY, X, SX, R0, T0, mode="forward", prec=prec, reps=1)
    assert gf is None and abs(Ef / ref["E"] - 1) <= etol


@pytest.mark.parametrize("prec,etol,gtol", [(64, 1e-12, 1e-10), (32, 2e-6, 2e-5)])
def test_energy_isotropic(oracle, prec, etol, gtol):
    rng = np.random.default_rng(2)
    Y = rng.normal(size=(280, 3)) * 0.1
    X = rng.normal(size=(300, 3)) * 0.1
    vy, vx = rng.uniform(1e-4, 4e-4, 280), rng.uniform(1e-4, 4e-4, 300)
    ref = oracle.energy(Y, vy[:, None, None] * np.eye(3), X, vx[:, None, None] * np.eye(3), R0, T0, grads=True)
    E, g, _ = CB.energy(Y, vy, X, vx, R0, T0, mode="fused", prec=prec, reps=1)
    assert abs(E / ref["E"] - 1) <= etol
    assert np.linalg.norm(g - ref["gpose"]) <= gtol * np.linalg.norm(ref["gpose"])


def test_energy_central_differences(oracle):
    rng = np.random.default_rng(3)
    Y, SY, X, SX = clouds(rng, 120, 110)
    Eref, gref = oracle.energy_fd_pose(Y, SY, X, SX, R0, T0, 1e-4)
    E, g, _ = CB.energy(Y, SY, X, SX, R0, T0, mode="fd", prec=64, h=1e-4, reps=1)
    assert abs(E / Eref - 1) <= 1e-12
    np.testing.assert_allclose(g, gref, rtol=1e-7, atol=1e-9 * np.abs(gref).max())


def test_update_sig(oracle):
    rng = np.random.default_rng(4)
    Xf = rng.normal(size=(700, 3)).astype(np.float32)
    Yf = rng.normal(size=(650, 3)).astype(np.float32)
    ref = oracle.update_sig(Xf, Yf, R0, T0)
    s, _ = CB.update_sig(Xf, Yf, R0, T0, reps=1)
    assert abs(s / ref - 1) <= 1e-6


@pytest.mark.parametrize("form", [0, 1])
def test_compi_both_forms(oracle, form):
    """The per-pair expansion and the per-source moment form give the
    oracle's coefficients (pinned by compi.cu's own expressions)."""
    rng = np.random.default_rng(5)
    N, M = 90, 80
    Xf = (rng.normal(size=(N, 3)) * 0.5).astype(np.float32)
    Yf = (rng.normal(size=(M, 3)) * 0.5).astype(np.float32)
    XCi, gX, _ = oracle.prep_inverse(spd(rng, N, 0.3, 0.02), 3)
    YCi, gY, _ = oracle.prep_inverse(spd(rng, M, 0.3, 0.02), 3)
    ref, missing = oracle.compi(Xf, XCi, gX, Yf, YCi, gY, R0, T0, 0.7)
    assert missing == 0
    c, _ = CB.compi(Xf, XCi, gX, Yf, YCi, gY, R0, T0, 0.7, form=form, reps=1)
    assert np.abs(c - ref).max() <= 1e-11 * np.abs(ref).max()
    big = np.abs(ref) > 1e-8 * np.abs(ref).max()
    np.testing.assert_allclose(c[big], ref[big], rtol=1e-9)


def test_host_description():
    assert CB.num_threads() >= 1
    assert isinstance(CB.cpu_model(), str) and CB.cpu_model()
