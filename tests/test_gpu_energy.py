"""GPU parity of the convolved energy and its adjoints against the FP64 oracle.

Bars: energy relative error <= 1e-6; gradients (pose, per point, per
covariance) relative error <= 1e-5 in the norm of each output array.
"""
import math
import os

import numpy as np
import pytest

import dugma_amd as dg
from dugma_amd import workloads as W

pytestmark = pytest.mark.gpu

E_TOL = 1e-6
G_TOL = 1e-5
HERE = os.path.dirname(os.path.abspath(__file__))


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def f32(a):
    return np.asarray(a, np.float64).astype(np.float32).astype(np.float64)


def as_oracle(d):
    """MATLAB-layout workload -> oracle arrays (point-major, (n,D,D)).
    The library evaluates the energy of the FP32 points and covariances (the
    reference's device state after init); the pose stays FP64."""
    return f32(d["X"].T), f32(W.unblocks(d["X_C"])), f32(d["Y"].T), f32(W.unblocks(d["Y_C"]))


def perturbed_pose(d, deg=3.0, dt=0.01, seed=0):
    rng = np.random.default_rng(seed)
    D = d["X"].shape[0]
    if D == 3:
        R = W.random_rotation(rng, deg) @ d["R_G"]
    else:
        R = W.rot2(math.radians(deg)) @ d["R_G"]
    return R, np.ravel(d["t_G"]) + rng.uniform(-dt, dt, size=D)


def check_energy(h, d, R, t, oracle, cov="auto", e_tol=E_TOL, g_tol=G_TOL):
    Y, SY, X, SX = as_oracle(d)
    ref = oracle.energy(Y, SY, X, SX, R, t, grads=True)
    E = dg.energy(h, R, t, cov)
    assert abs(E / ref["E"] - 1) <= e_tol, (E, ref["E"])
    E2, g = dg.energy_grad(h, R, t, cov)
    assert abs(E2 / ref["E"] - 1) <= e_tol
    assert rel(g, ref["gpose"]) <= g_tol, (g, ref["gpose"])
    return ref


def test_energy_config2_full_covariance(oracle):
    d = W.config2(n=900, m=700)
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    assert ok == 1.0
    try:
        nT, nS, iso = h.energy_info()
        assert (nT, nS, iso) == (900, 700, False)
        for s in range(3):
            R, t = perturbed_pose(d, deg=1.0 + s, dt=0.003, seed=s)
            check_energy(h, d, R, t, oracle)
    finally:
        dg.fr_m(h)


def test_energy_config2_full_size(oracle):
    """The whole 6,000 x 6,000 structured-light workload (full covariances,
    10% outliers): energy and pose gradient against the FP64 oracle at the
    truth and at the bench's start pose (2 degrees, 3 mm off), to the bars
    above; the 13-launch finite-difference call returns the same energy."""
    d = W.config2(n=6000, m=6000)
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    try:
        rng = np.random.default_rng(0)
        poses = [(d["R_G"], np.ravel(d["t_G"])),
                 (W.random_rotation(rng, 2.0) @ d["R_G"], np.ravel(d["t_G"]) + rng.uniform(-0.003, 0.003, size=3))]
        for R, t in poses:
            ref = check_energy(h, d, R, t, oracle)
            E, _ = dg.energy_grad_fd(h, R, t, 1e-4)
            assert abs(E / ref["E"] - 1) <= E_TOL
    finally:
        dg.fr_m(h)


def test_energy_grad_points_config2(oracle):
    d = W.config2(n=500, m=400)
    R, t = perturbed_pose(d, deg=2.0, dt=0.002, seed=7)
    Y, SY, X, SX = as_oracle(d)
    ref = oracle.energy(Y, SY, X, SX, R, t, grads=True)
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    try:
        out = dg.energy_grad_points(h, R, t)
    finally:
        dg.fr_m(h)
    assert abs(out["E"] / ref["E"] - 1) <= E_TOL
    assert rel(out["gpose"], ref["gpose"]) <= G_TOL
    assert rel(out["gX"], ref["gY"]) <= G_TOL  # library X = target = oracle Y
    assert rel(out["gXC"], ref["gSY"]) <= G_TOL
    assert rel(out["gY"], ref["gX"]) <= G_TOL
    assert rel(out["gYC"], ref["gSX"]) <= G_TOL
    # dE/d(dT) of energy3D.m's variables: central differences of the FP64
    # oracle along R <- dR R, t <- t + dt (DUGMA.m:76-78)
    hstep = 1e-7
    gd = np.zeros(12)
    base = np.concatenate([np.zeros(3), np.eye(3).ravel()])
    for k in range(12):
        vals = []
        for sgn in (1, -1):
            v = base.copy()
            v[k] += sgn * hstep
            dR = v[3:].reshape(3, 3)
            vals.append(oracle.energy(Y, SY, X, SX, dR @ R, t + v[:3])["E"])
        gd[k] = (vals[0] - vals[1]) / (2 * hstep)
    assert rel(out["gdT"], gd) <= 1e-4  # FD truncation/rounding of the oracle itself


def test_energy_isotropic_config3_small(oracle):
    d = W.config3(n=3000, m=2500)
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    try:
        assert h.energy_info()[2]  # isotropic path
        R, t = perturbed_pose(d, deg=2.0, dt=0.005, seed=3)
        check_energy(h, d, R, t, oracle, "auto")
        check_energy(h, d, R, t, oracle, "full")  # full kernels on isotropic data
        check_energy(h, d, R, t, oracle, "iso")
    finally:
        dg.fr_m(h)


def test_energy_2d(oracle):
    d = W.config5_2d(n=1500, m=1300)
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    try:
        R, t = perturbed_pose(d, deg=4.0, dt=0.01, seed=5)
        check_energy(h, d, R, t, oracle, "auto")
        check_energy(h, d, R, t, oracle, "full")
    finally:
        dg.fr_m(h)
    # full 2-D covariances
    rng = np.random.default_rng(8)
    A = rng.normal(size=(800, 2, 2)) * 0.01
    CX = A @ np.transpose(A, (0, 2, 1)) + 1e-5 * np.eye(2)
    B = rng.normal(size=(600, 2, 2)) * 0.01
    CY = B @ np.transpose(B, (0, 2, 1)) + 1e-5 * np.eye(2)
    d2 = dict(X=d["X"][:, :800], X_C=W.blocks(CX), Y=d["Y"][:, :600], Y_C=W.blocks(CY),
              R_G=d["R_G"], t_G=d["t_G"])
    ok, h = dg.init(d2["X"], d2["X_C"], d2["Y"], d2["Y_C"])
    try:
        R, t = perturbed_pose(d2, deg=2.0, dt=0.005, seed=6)
        ref = check_energy(h, d2, R, t, oracle)
        out = dg.energy_grad_points(h, R, t)
        assert rel(out["gX"], ref["gY"]) <= G_TOL
        assert rel(out["gYC"], ref["gSX"]) <= G_TOL
        assert rel(out["gpose"], ref["gpose"]) <= G_TOL
    finally:
        dg.fr_m(h)


def test_energy_invalid_points_left_out(oracle):
    """NaN covariances (as in Kinect scenes 2, 11, 22) and non-PD ones."""
    g = np.load(os.path.join(HERE, "golden", "kinect_fixtures.npz"))
    for k in (2, 11, 22):
        d = dict(X=g["s%d_X" % k], X_C=g["s%d_X_C" % k], Y=g["s%d_Y" % k], Y_C=g["s%d_Y_C" % k],
                 R_G=g["s%d_R_G" % k], t_G=g["s%d_t_G" % k])
        XC = W.unblocks(d["X_C"]).copy()
        XC[1] = np.diag([1.0, -1.0, 1.0])  # not positive definite
        d["X_C"] = W.blocks(XC)
        ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
        try:
            nT, nS, _ = h.energy_info()
            assert nT < d["X"].shape[1] or nS < d["Y"].shape[1]
            check_energy(h, d, d["R_G"], d["t_G"].ravel(), oracle)
        finally:
            dg.fr_m(h)


@pytest.mark.parametrize("n,m", [(1, 1), (1, 777), (333, 1), (257, 65), (4099, 129)])
def test_energy_ragged_sizes(oracle, n, m):
    """Every pair counted once at sizes off the 64/256 grid.  With a handful
    of pairs nothing averages: the FP32 exp argument error (~q * 2^-24 for
    Mahalanobis distance q) is the whole error, so the bar here is 2e-5."""
    rng = np.random.default_rng(n * 7 + m)
    X = rng.normal(size=(3, n)) * 0.2
    Y = rng.normal(size=(3, m)) * 0.2
    A = rng.normal(size=(n, 3, 3)) * 0.05
    B = rng.normal(size=(m, 3, 3)) * 0.05
    CX = A @ np.transpose(A, (0, 2, 1)) + 1e-3 * np.eye(3)
    CY = B @ np.transpose(B, (0, 2, 1)) + 1e-3 * np.eye(3)
    d = dict(X=X, X_C=W.blocks(CX), Y=Y, Y_C=W.blocks(CY), R_G=np.eye(3), t_G=np.zeros(3))
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    try:
        check_energy(h, d, W.random_rotation(rng, 10), rng.normal(size=3) * 0.01, oracle,
                     e_tol=2e-5, g_tol=2e-5)
    finally:
        dg.fr_m(h)


def test_fd_gradient_is_consistent_with_fused(oracle):
    d = W.config3(n=2000, m=2000)
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    try:
        R, t = perturbed_pose(d, deg=3.0, dt=0.01, seed=2)
        E, g = dg.energy_grad(h, R, t)
        E2, gfd = dg.energy_grad_fd(h, R, t, step=1e-4)
        assert abs(E2 / E - 1) < 1e-12
        assert rel(gfd, g) < 1e-3  # FD truncation at h = 1e-4 and FP32 rounding
    finally:
        dg.fr_m(h)


def test_deterministic_repeats():
    d = W.config2(n=700, m=650)
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    try:
        R, t = perturbed_pose(d, seed=1)
        a = dg.energy_grad(h, R, t)
        b = dg.energy_grad(h, R, t)
        assert a[0] == b[0] and (a[1] == b[1]).all()
    finally:
        dg.fr_m(h)


def test_reduced_pass_result_paths_agree():
    """Above 256 workgroup partials the pose pass reduces on the device over
    several workgroups and the host polls the pinned result slots; with the
    kernel timing on it synchronises the stream instead.  Both paths and
    repeated calls give bit-identical results, and the energy-only entry
    point gives the same energy."""
    d = W.config2(n=3000, m=3000)
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    try:
        R, t = perturbed_pose(d, seed=2)
        a = dg.energy_grad(h, R, t)
        b = dg.energy_grad(h, R, t)
        h.set_timing(True)
        c = dg.energy_grad(h, R, t)
        ms, launches = h.pair_time()
        h.set_timing(False)
        e = dg.energy(h, R, t)
    finally:
        dg.fr_m(h)
    assert launches == 1 and ms > 0.0
    for x in (b, c):
        assert x[0] == a[0] and (x[1] == a[1]).all()
    assert math.isclose(e, a[0], rel_tol=1e-6)


def test_depth_views_full_size_adjoint_identities():
    """Configuration 5, 3-D: two 320 x 240 depth views (about 76,800 points
    each, full 3 x 3 covariances, invalid pixels left out) at full size,
    checked through identities that do not need an N x M reference:
      * moving every target by v and every source by R^T v leaves E alone,
        so sum_i dE/dy_i + R sum_j dE/dx_j = 0;
      * the pose pass's dE/dt is -sum_i dE/dy_i and its energy agrees;
      * scaling every covariance by lambda: dE/dlambda at 1 equals
        sum over points of tr(dE/dS S), against central differences of E."""
    d = W.config5_3d()
    R, t = perturbed_pose(d, deg=2.0, dt=0.005, seed=11)
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    try:
        out = dg.energy_grad_points(h, R, t)
        E1, gp = dg.energy_grad(h, R, t, "full")
    finally:
        dg.fr_m(h)
    gX, gY = out["gX"], out["gY"]
    assert abs(out["E"] / E1 - 1) <= 2 * E_TOL
    mass = np.abs(gX).sum() + np.abs(gY).sum()
    assert np.linalg.norm(gX.sum(0) + R @ gY.sum(0)) <= 1e-5 * mass
    assert rel(gp[3:], -gX.sum(0)) <= G_TOL
    SXc, SYc = W.unblocks(d["X_C"]), W.unblocks(d["Y_C"])
    tr = np.einsum("nab,nba->", out["gXC"], SXc) + np.einsum("nab,nba->", out["gYC"], SYc)
    lam = 1e-3
    Es = []
    for s in (1 + lam, 1 - lam):
        okk, hs = dg.init(d["X"], d["X_C"] * s, d["Y"], d["Y_C"] * s)
        try:
            Es.append(dg.energy(hs, R, t, "full"))
        finally:
            dg.fr_m(hs)
    fd = (Es[0] - Es[1]) / (2 * lam)
    assert abs(tr - fd) <= 1e-3 * abs(fd), (tr, fd)


def test_depth_views_subsample_matches_oracle(oracle):
    """The same depth views, 1,500 points of each, per-point gradients against
    the FP64 oracle."""
    d = W.config5_3d()
    rng = np.random.default_rng(3)
    n, m = d["X"].shape[1], d["Y"].shape[1]
    it, js = np.sort(rng.choice(n, 1500, replace=False)), np.sort(rng.choice(m, 1500, replace=False))
    sub = dict(X=d["X"][:, it].copy(), Y=d["Y"][:, js].copy(),
               X_C=W.blocks(W.unblocks(d["X_C"])[it]), Y_C=W.blocks(W.unblocks(d["Y_C"])[js]),
               R_G=d["R_G"], t_G=d["t_G"])
    R, t = perturbed_pose(sub, deg=2.0, dt=0.005, seed=12)
    Y, SY, X, SX = as_oracle(sub)
    ref = oracle.energy(Y, SY, X, SX, R, t, grads=True)
    ok, h = dg.init(sub["X"], sub["X_C"], sub["Y"], sub["Y_C"])
    try:
        out = dg.energy_grad_points(h, R, t)
    finally:
        dg.fr_m(h)
    assert abs(out["E"] / ref["E"] - 1) <= E_TOL
    assert rel(out["gX"], ref["gY"]) <= G_TOL
    assert rel(out["gXC"], ref["gSY"]) <= G_TOL
    assert rel(out["gY"], ref["gX"]) <= G_TOL
    assert rel(out["gYC"], ref["gSX"]) <= G_TOL


def test_points_repeatable_and_shape_independent():
    """Per-point gradients come from GPU sums that the host turns into
    gradients on several threads over fixed chunks: repeated calls give the
    same bits.  Another launch shape (fewer streamed chunks) regroups the
    FP32 sub-tile sums, and agrees within the parity bars."""
    d = W.config2(n=9000, m=8500)  # more than two host chunks on each side
    R, t = perturbed_pose(d, deg=2.0, dt=0.005, seed=21)
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    old = os.environ.get("DUGMA_POINT_WAVES")
    try:
        a = dg.energy_grad_points(h, R, t)
        b = dg.energy_grad_points(h, R, t)
        for key in ("gX", "gXC", "gY", "gYC", "gpose", "gdT"):
            assert np.array_equal(a[key], b[key]), key
        assert a["E"] == b["E"]
        os.environ["DUGMA_POINT_WAVES"] = "2048"
        c = dg.energy_grad_points(h, R, t)
    finally:
        if old is None:
            os.environ.pop("DUGMA_POINT_WAVES", None)
        else:
            os.environ["DUGMA_POINT_WAVES"] = old
        dg.fr_m(h)
    assert abs(c["E"] / a["E"] - 1) <= E_TOL
    for key in ("gX", "gXC", "gY", "gYC", "gpose", "gdT"):
        assert rel(c[key], a[key]) <= G_TOL, key


def test_energy_full_size_config3(oracle):
    """The fused pass at the large configuration's full size (200,000 x
    200,000, isotropic), through properties that do not need an all-pairs
    FP64 reference:
      * 512 sampled targets against all 200,000 sources match the FP64
        oracle (energy and pose gradient);
      * the energy and pose gradient are additive over a split of the
        targets (the two halves' contexts against the whole one's);
      * the fused pose gradient agrees with central differences of the
        GPU energy itself."""
    d = W.config3()
    R, t = perturbed_pose(d, deg=1.0, dt=0.003, seed=31)
    XC = W.unblocks(d["X_C"])
    N = d["X"].shape[1]

    def sub(idx):
        return dict(X=d["X"][:, idx].copy(), X_C=W.blocks(XC[idx]), Y=d["Y"], Y_C=d["Y_C"])

    def run(dd, fn):
        ok, h = dg.init(dd["X"], dd["X_C"], dd["Y"], dd["Y_C"])
        try:
            return fn(h)
        finally:
            dg.fr_m(h)

    q = np.sort(np.random.default_rng(4).choice(N, 512, replace=False))
    s = sub(q)
    E_s, g_s = run(s, lambda h: dg.energy_grad(h, R, t))
    Y, SY, X, SX = as_oracle(s)
    ref = oracle.energy(Y, SY, X, SX, R, t, grads=True)
    assert abs(E_s / ref["E"] - 1) <= E_TOL, (E_s, ref["E"])
    assert rel(g_s, ref["gpose"]) <= G_TOL

    def full(h):
        E, g = dg.energy_grad(h, R, t)
        fd = dg.energy_grad_fd(h, R, t, step=1e-4)[1]
        return E, g, fd
    E, g, gfd = run(d, full)
    half = np.arange(N) < N // 2
    E1, g1 = run(sub(np.flatnonzero(half)), lambda h: dg.energy_grad(h, R, t))
    E2, g2 = run(sub(np.flatnonzero(~half)), lambda h: dg.energy_grad(h, R, t))
    assert abs((E1 + E2) / E - 1) <= E_TOL
    assert rel(np.asarray(g1) + np.asarray(g2), g) <= G_TOL
    assert rel(gfd, g) < 1e-3  # FD truncation at h = 1e-4 and FP32 rounding


@pytest.mark.parametrize("case", ["full3", "iso3", "iso2"])
def test_stationary_set_does_not_change_results(case):
    """The pose and energy passes keep the transformed sources stationary
    (the gradient sums change sign, the torque is taken about the sources);
    with DUGMA_STATIONARY_SOURCES=0 the targets are stationary.  Both give
    the same energy and pose gradient within the parity bars."""
    if case == "full3":
        d = W.config2(n=1500, m=1300)
    elif case == "iso3":
        d = W.config3(n=3000, m=2600)
    else:
        d = W.config5_2d(n=900, m=800)
    R, t = perturbed_pose(d, deg=2.0, dt=0.005, seed=41)
    old = os.environ.get("DUGMA_STATIONARY_SOURCES")
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    try:
        out = {}
        for flag in ("1", "0"):
            os.environ["DUGMA_STATIONARY_SOURCES"] = flag
            out[flag] = (dg.energy(h, R, t),) + tuple(dg.energy_grad(h, R, t))
    finally:
        if old is None:
            os.environ.pop("DUGMA_STATIONARY_SOURCES", None)
        else:
            os.environ["DUGMA_STATIONARY_SOURCES"] = old
        dg.fr_m(h)
    (Ea, E2a, ga), (Eb, E2b, gb) = out["1"], out["0"]
    assert abs(Ea / Eb - 1) <= E_TOL and abs(E2a / E2b - 1) <= E_TOL
    assert rel(ga, gb) <= G_TOL, (ga, gb)


@pytest.mark.parametrize("case", ["full3", "iso3"])
def test_one_point_per_lane_matches_oracle(oracle, case):
    """The autotuner's usual pick at 6K and 200K (one stationary point per
    lane, many streamed chunks) against the FP64 oracle; the other tests run
    the default plan (two per lane)."""
    d = W.config2(n=1400, m=1200) if case == "full3" else W.config3(n=3000, m=2600)
    R, t = perturbed_pose(d, deg=2.0, dt=0.005, seed=43)
    saved = {k: os.environ.get(k) for k in ("DUGMA_LANE_POINTS", "DUGMA_TARGET_WAVES")}
    os.environ["DUGMA_LANE_POINTS"] = "1"
    os.environ["DUGMA_TARGET_WAVES"] = "16384"
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    try:
        check_energy(h, d, R, t, oracle)
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
        dg.fr_m(h)


@pytest.mark.parametrize("cov", ["iso", "full"])
def test_xcd_tile_order_gives_the_same_sums(monkeypatch, cov):
    """Clouds beyond 4 MiB of records run their energy and pose passes in the
    XCD-aware tile order.  Every tile writes its partial to its own slot and
    the partials are reduced in a fixed order, so the results equal the 2-D
    order's (DUGMA_XCD_ORDER=0) bit for bit, which also shows that every
    tile of the padded grid runs exactly once."""
    d = W.config3(n=90000, m=80000)  # 32-B isotropic / 64-B full records: beyond 4 MiB
    R, t = d["R_G"], np.ravel(d["t_G"])
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    try:
        out = {}
        for v in ("1", "0"):
            monkeypatch.setenv("DUGMA_XCD_ORDER", v)
            E0 = dg.energy(h, R, t, cov=cov)
            E1, g = dg.energy_grad(h, R, t, cov=cov)
            out[v] = (E0, E1, g)
    finally:
        dg.fr_m(h)
    assert out["1"][0] == out["0"][0] and out["1"][1] == out["0"][1]
    np.testing.assert_array_equal(out["1"][2], out["0"][2])
