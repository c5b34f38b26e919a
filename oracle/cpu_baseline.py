"""ctypes bindings of the CPU baseline (libdugma_cpu_baseline.so, built from
dugma_cpu_baseline.cpp by oracle/Makefile).

TEST INFRASTRUCTURE ONLY: bench.py's cpu_baseline leg times these loops on
the host cores and tests/test_cpu_baseline.py checks them against the oracle.
The dugma_amd package never imports this module.

Every function returns (result, median seconds of `reps` timed runs after
one untimed run).  Arrays follow the oracle's conventions: points (n, 3),
covariances (n, 3, 3) or, isotropic, one variance per point (n,).
"""
import ctypes
import os
import platform
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_d = ctypes.POINTER(ctypes.c_double)
_f = ctypes.POINTER(ctypes.c_float)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libdugma_cpu_baseline.so")
        if not os.path.exists(path):
            raise RuntimeError("CPU baseline not built: run make -C oracle")
        L = ctypes.CDLL(path)
        ci, cd = ctypes.c_int, ctypes.c_double
        L.dcb_num_threads.restype = ci
        L.dcb_energy.argtypes = [ci, ci, ci, ci, _d, _d, ci, _d, _d, _d, _d, cd, _d, ci, _d]
        L.dcb_energy.restype = cd
        L.dcb_compi.argtypes = [ci, ci, _f, _f, _f, ci, _f, _f, _f, _d, _d, cd, _d, ci, _d]
        L.dcb_update_sig.argtypes = [ci, _f, ci, _f, _d, _d, ci, _d]
        L.dcb_update_sig.restype = ctypes.c_float
        _LIB = L
    return _LIB


def _p(a, ptr):
    return None if a is None else a.ctypes.data_as(ptr)


def _colmajor(R):
    return np.ascontiguousarray(np.asarray(R, np.float64).T).ravel()


def num_threads():
    return lib().dcb_num_threads()


def cpu_model():
    """The host CPU's model name (lscpu's 'Model name', else /proc/cpuinfo)."""
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True,
                             timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


MODES = {"forward": 0, "fused": 1, "fd": 2}


def energy(Y, SY, X, SX, R, t, mode="fused", prec=32, h=1e-4, reps=3):
    """Convolved energy over targets Y and sources X (3-D).  mode: forward,
    fused (E + pose gradient) or fd (central differences, 13 passes).
    Returns (E, gpose or None, median seconds)."""
    Y = np.ascontiguousarray(Y, np.float64)
    X = np.ascontiguousarray(X, np.float64)
    SY = np.asarray(SY, np.float64)
    SX = np.asarray(SX, np.float64)
    iso = SY.ndim == 1
    if iso:
        SYc, SXc = np.ascontiguousarray(SY), np.ascontiguousarray(SX)
    else:  # column-major blocks
        SYc = np.ascontiguousarray(np.transpose(SY, (0, 2, 1)))
        SXc = np.ascontiguousarray(np.transpose(SX, (0, 2, 1)))
    g = np.zeros(6)
    ts = ctypes.c_double()
    E = lib().dcb_energy(int(prec), MODES[mode], int(iso), Y.shape[0], _p(Y, _d), _p(SYc, _d), X.shape[0],
                         _p(X, _d), _p(SXc, _d), _p(_colmajor(R), _d), _p(np.ascontiguousarray(t, np.float64), _d),
                         float(h), _p(g, _d), int(reps), ctypes.byref(ts))
    return E, (None if mode == "forward" else g), ts.value


def compi(Xf, XCinv, gX, Yf, YCinv, gY, R, t, sig, form=0, reps=3):
    """compi's 637 coefficients (3-D): form 0 per-pair expansion, form 1
    per-source moments.  XCinv / YCinv (n, 9) column-major prepped inverses."""
    a = [np.ascontiguousarray(v, np.float32) for v in (Xf, XCinv, gX, Yf, YCinv, gY)]
    coef = np.zeros(637)
    ts = ctypes.c_double()
    lib().dcb_compi(int(form), a[0].shape[0], _p(a[0], _f), _p(a[1], _f), _p(a[2], _f), a[3].shape[0],
                    _p(a[3], _f), _p(a[4], _f), _p(a[5], _f), _p(_colmajor(R), _d),
                    _p(np.ascontiguousarray(t, np.float64), _d), float(sig), _p(coef, _d), int(reps),
                    ctypes.byref(ts))
    return coef, ts.value


def update_sig(Xf, Yf, R, t, reps=3):
    Xf = np.ascontiguousarray(Xf, np.float32)
    Yf = np.ascontiguousarray(Yf, np.float32)
    ts = ctypes.c_double()
    s = lib().dcb_update_sig(Xf.shape[0], _p(Xf, _f), Yf.shape[0], _p(Yf, _f), _p(_colmajor(R), _d),
                             _p(np.ascontiguousarray(t, np.float64), _d), int(reps), ctypes.byref(ts))
    return float(s), ts.value
