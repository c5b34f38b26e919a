"""dugma_energy_grad_noalloc (the fused energy + pose gradient with its
partials in the library's file-scope scratch array) gives, through the C ABI,
exactly what dugma_energy_grad gives on the same context and pose: two calls
in a row, the second on a smaller problem; a launch larger than the array is
refused before anything runs."""
import ctypes

import numpy as np
import pytest

import dugma_amd as dg
from dugma_amd import _lib
from dugma_amd import workloads as W

pytestmark = pytest.mark.gpu


def both(h, R, t):
    """(E, gpose) from dugma_energy_grad_noalloc, then from dugma_energy_grad,
    each called through ctypes with its own result buffers."""
    L = _lib.load()
    Rc = np.ascontiguousarray(np.asarray(R, np.float64).T).ravel()
    tc = np.ascontiguousarray(np.ravel(t), np.float64)
    dp = ctypes.POINTER(ctypes.c_double)
    out = []
    for fn in (L.dugma_energy_grad_noalloc, L.dugma_energy_grad):
        E = ctypes.c_double(np.nan)
        g = np.full(6 if h.D == 3 else 3, np.nan)
        rc = fn(h.ptr, Rc.ctypes.data_as(dp), tc.ctypes.data_as(dp), 0, ctypes.byref(E), g.ctypes.data_as(dp))
        assert rc == 0, L.dugma_last_error().decode()
        out.append((E.value, g))
    return out


def test_noalloc_equals_energy_grad_two_calls():
    cases = (("6,000 x 6,000 full covariance", W.config2(n=6000, m=6000)),
             ("1,500 x 1,200 isotropic", W.config1(n=1500, m=1200)))
    for label, d in cases:
        print("no-alloc fused step: %s" % label, flush=True)
        ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
        try:
            rng = np.random.default_rng(7)
            R = W.random_rotation(rng, 2.0) @ d["R_G"]
            t = np.ravel(d["t_G"]) + rng.uniform(-0.003, 0.003, size=3)
            (E0, g0), (E1, g1) = both(h, R, t)
        finally:
            dg.fr_m(h)
        assert np.isfinite(E0) and np.all(np.isfinite(g0))
        assert E0 == E1
        np.testing.assert_array_equal(g0, g1)


def test_noalloc_refuses_a_launch_beyond_the_array(monkeypatch):
    """200,000 x 200,000 at 2^22 target waves needs 4.9e6 partial records
    (the array holds 2^18): the call fails with DUGMA_ERR_SIZE before any
    launch; at the default shape it matches dugma_energy_grad."""
    d = W.config3()
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    try:
        R, t = d["R_G"], np.ravel(d["t_G"])
        (E0, g0), (E1, g1) = both(h, R, t)
        assert E0 == E1
        np.testing.assert_array_equal(g0, g1)
        monkeypatch.setenv("DUGMA_TARGET_WAVES", str(1 << 22))
        monkeypatch.setenv("DUGMA_LANE_POINTS", "1")
        L = _lib.load()
        Rc = np.ascontiguousarray(R.T).ravel()
        dp = ctypes.POINTER(ctypes.c_double)
        E = ctypes.c_double()
        g = np.zeros(6)
        rc = L.dugma_energy_grad_noalloc(h.ptr, Rc.ctypes.data_as(dp), t.ctypes.data_as(dp), 0, ctypes.byref(E),
                                         g.ctypes.data_as(dp))
        assert rc == -4, rc
        assert "scratch" in L.dugma_last_error().decode()
    finally:
        dg.fr_m(h)
