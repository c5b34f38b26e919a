"""The MATLAB MEX gateways (dugma_amd/mex) driven through a minimal mxArray
layer: argument and handle checks on the CPU, and the DUGMA.m call
sequence init -> update_sig -> compi -> fr_m on the GPU."""
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
