"""The DUGMA.m loop (dugma_register) end to end on the GPU."""
import os

import numpy as np
import pytest

import dugma_amd as dg
from dugma_amd import workloads as W

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def test_config1_registration():
    """1,000 x 1,000, isotropic 8-NN sigma, rotation up to 45 degrees."""
    d = W.config1()
    R, t, it, sig = dg.DUGMA(d["X"], d["Y"], d["X_C"], d["Y_C"], np.eye(3), np.zeros(3), 150)
    err = W.rotation_error_deg(R, d["R_G"])
    start = W.rotation_error_deg(np.eye(3), d["R_G"])
    assert it >= 1
    assert err < start  # the loop moves toward the ground truth
    assert np.all(np.isfinite(R)) and np.all(np.isfinite(t))


def test_kinect_fixture_scenes():
    """Decimated Kinect scenes (every 25th point, NaN covariances kept):
    DUGMA's own error measures (main_kinect_real_application.m:33-34)."""
    g = np.load(os.path.join(HERE, "golden", "kinect_fixtures.npz"))
    for k in (2, 11, 20, 22):
        X, XC, Y, YC = g["s%d_X" % k], g["s%d_X_C" % k], g["s%d_Y" % k], g["s%d_Y_C" % k]
        R, t, it, sig = dg.DUGMA(X, Y, XC, YC, np.eye(3), np.zeros(3), 300, 1e-4, 1e-8, 1e-9)
        R_G, t_G = g["s%d_R_G" % k], g["s%d_t_G" % k].ravel()
        err_R = np.linalg.norm(np.eye(3) - R @ np.linalg.pinv(R_G))
        assert np.all(np.isfinite(R)) and abs(np.linalg.det(R) - 1) < 1e-9
        assert err_R < np.linalg.norm(np.eye(3) - np.linalg.pinv(R_G)), (k, err_R)


def test_2d_registration():
    d = W.config5_2d(n=600, m=500)
    R, t, it, sig = dg.DUGMA(d["X"], d["Y"], d["X_C"], d["Y_C"], np.eye(2), np.zeros(2), 200)
    ang = np.degrees(np.arctan2(R[1, 0], R[0, 0]))
    gt = np.degrees(np.arctan2(d["R_G"][1, 0], d["R_G"][0, 0]))
    assert abs(ang - gt) < abs(gt) or abs(gt) < 1.0
