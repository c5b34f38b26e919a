"""The MATLAB MEX gateways (dugma_amd/mex) driven through a minimal mxArray
layer: argument and handle checks on the CPU, and the DUGMA.m call
sequence init -> update_sig -> compi -> fr_m on the GPU."""
import math
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "mex_harness", "build.sh")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("mex") / "harness")
    subprocess.run(["bash", BUILD, exe], check=True, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    return exe


def test_gateway_argument_and_handle_checks(harness):
    r = subprocess.run([harness, "errors"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout


@pytest.mark.gpu
def test_gateway_call_sequence_matches_library(harness, tmp_path):
    import dugma_amd as dg
    from dugma_amd import workloads as W
    d = W.config1(n=300, m=250)
    R = W.random_rotation(np.random.default_rng(2), 10.0) @ d["R_G"]
    t = d["t_G"]
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        np.array([300, 250, 3], np.int32).tofile(f)
        for a in (d["X"], d["X_C"], d["Y"], d["Y_C"], R):
            np.asfortranarray(a, dtype=np.float64).ravel(order="F").tofile(f)
        np.asarray(t, np.float64).tofile(f)
    r = subprocess.run([harness, "run", str(inp), str(out)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    res = np.fromfile(out, np.float64)
    sig, coef, ok = res[0], res[1:-1], res[-1]
    assert coef.size == 637 and ok == 1.0
    okp, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    try:
        s2 = dg.update_sig(h, R, t, h.N, h.M, 3)
        c2 = dg.compi(h, R, t, h.N, h.M, 3, s2)
    finally:
        dg.fr_m(h)
    assert sig == s2
    np.testing.assert_array_equal(coef, c2)


def _problem(path, d, R, t, extra=()):
    D = d["X"].shape[0]
    with open(path, "wb") as f:
        np.array([d["X"].shape[1], d["Y"].shape[1], D], np.int32).tofile(f)
        for a in (d["X"], d["X_C"], d["Y"], d["Y_C"], R):
            np.asfortranarray(a, dtype=np.float64).ravel(order="F").tofile(f)
        np.asarray(t, np.float64).tofile(f)
        np.asarray(extra, np.float64).tofile(f)


@pytest.mark.gpu
@pytest.mark.parametrize("D", [3, 2])
def test_energy_grad_gateway_matches_central_differences(harness, tmp_path, D):
    """[negE, g] = energy_grad(ad0, R, t, dT): the gateway's gradient over
    fmincon's variables (energy3D.m / energy2D.m layout, dR entered as dR R)
    against central differences of the gateway's own value, at a dT that is
    not the identity so that the dR^-T chain rule is exercised."""
    import dugma_amd as dg
    from dugma_amd import workloads as W
    rng = np.random.default_rng(7)
    if D == 3:
        d = W.config1(n=300, m=250)
        R = W.random_rotation(rng, 10.0) @ d["R_G"]
        dR = W.random_rotation(rng, 3.0)
        dT = np.concatenate([[0.01, -0.02, 0.015], dR.ravel()])  # r11 r12 r13 r21 ... row-major
    else:
        d = W.config5_2d(n=400, m=350)
        th = 0.1
        R = np.array([[math.cos(th), -math.sin(th)], [math.sin(th), math.cos(th)]]) @ d["R_G"]
        dT = np.array([0.01, -0.015, math.cos(0.05), math.sin(0.05)])
    t = d["t_G"]
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    _problem(inp, d, R, t, list(dT) + [1e-4])
    r = subprocess.run([harness, "grad", str(inp), str(out)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    nv = dT.size
    res = np.fromfile(out, np.float64)
    negE, g, gfd = res[0], res[1:1 + nv], res[1 + nv:]
    # value: the library energy at the composed pose (dR R, t + dt)
    if D == 3:
        dRm = dT[3:].reshape(3, 3)
        dt = dT[:3]
    else:
        dRm = np.array([[dT[2], -dT[3]], [dT[3], dT[2]]])
        dt = dT[:2]
    okp, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    try:
        E = dg.energy(h, dRm @ R, t + dt)
    finally:
        dg.fr_m(h)
    # the gateway's value comes from the per-point gradient pass, the
    # library's from the forward pass: both within 1e-6 of the FP64 energy
    assert abs(-negE - E) <= 2e-6 * abs(E)
    # FP32 per-pair rounding is ~1e-7 of E; with h = 1e-4 the differences
    # carry ~1e-3 of the largest component
    assert np.abs(g - gfd).max() <= 2e-3 * np.abs(g).max(), (g, gfd)
