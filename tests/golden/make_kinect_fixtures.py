"""Small fixtures from the reference's Kinect data and published outcomes.

Reads (with scipy.io.loadmat, numeric arrays only):
  dataset_all/Kinect_Application/{k}.mat          X, X_C, Y, Y_C, R_G, t_G
  Our_result_in_Kinect_Application/Ours_{k}_result.mat   R, t, R_G, t_G, time_stop_DUGMA
and writes beside this script:
  kinect_fixtures.npz
  * for scenes 2, 11, 20, 22: every 25th point plus every point whose
    covariance is NaN or has variance >= 1e5 (the init fallback and
    near-zero-weight cases);
  * the published per-scene results of all 30 scenes;
  kinect_full_scenes.npz
  * scenes 2, 8, 20, 22 whole: points, one variance per point (every
    covariance block of these scenes is v I, the NaN ones NaN on the
    diagonal and 0 elsewhere, which v = NaN restores), R_G, t_G.
Run where the reference is present:  python tests/golden/make_kinect_fixtures.py
"""
import os
import sys

import numpy as np
import scipy.io as sio

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def special(C):
    D = C.shape[0]
    n = C.shape[1] // D
    B = np.transpose(C.reshape(D, n, D), (1, 0, 2))
    bad = ~np.isfinite(B).all(axis=(1, 2))
    big = np.nanmax(np.abs(B), axis=(1, 2)) >= 1e5
    return np.nonzero(bad | big)[0]


def subset(P, C, step):
    D = P.shape[0]
    n = P.shape[1]
    keep = sorted(set(range(0, n, step)) | set(special(C).tolist()))
    keep = np.array(keep)
    cols = (keep[:, None] * D + np.arange(D)[None, :]).ravel()
    return P[:, keep], C[:, cols], keep


def main():
    out = {}
    for k in (2, 11, 20, 22):
        d = sio.loadmat(os.path.join(REF, "dataset_all", "Kinect_Application", "%d.mat" % k))
        X, XC, kx = subset(d["X"], d["X_C"], 25)
        Y, YC, ky = subset(d["Y"], d["Y_C"], 25)
        out["s%d_X" % k], out["s%d_X_C" % k] = X, XC
        out["s%d_Y" % k], out["s%d_Y_C" % k] = Y, YC
        out["s%d_R_G" % k], out["s%d_t_G" % k] = d["R_G"], d["t_G"]
        out["s%d_keepX" % k], out["s%d_keepY" % k] = kx, ky
    res = {"R": [], "t": [], "R_G": [], "t_G": [], "time": [], "N": [], "M": []}
    for k in range(1, 31):
        r = sio.loadmat(os.path.join(REF, "Our_result_in_Kinect_Application", "Ours_%d_result.mat" % k))
        d = sio.loadmat(os.path.join(REF, "dataset_all", "Kinect_Application", "%d.mat" % k),
                        variable_names=["X", "Y"])
        res["R"].append(r["R"]); res["t"].append(r["t"].ravel())
        res["R_G"].append(r["R_G"]); res["t_G"].append(r["t_G"].ravel())
        res["time"].append(float(r["time_stop_DUGMA"].ravel()[0]))
        res["N"].append(d["X"].shape[1]); res["M"].append(d["Y"].shape[1])
    for key, v in res.items():
        out["published_" + key] = np.array(v)
    dst = os.path.join(HERE, "kinect_fixtures.npz")
    np.savez_compressed(dst, **out)
    print("wrote", dst, os.path.getsize(dst), "bytes")
    full = {}
    for k in FULL_SCENES:
        d = sio.loadmat(os.path.join(REF, "dataset_all", "Kinect_Application", "%d.mat" % k))
        for name in ("X", "Y"):
            full["s%d_%s" % (k, name)] = d[name]
            full["s%d_v%s" % (k, name)] = variances(d[name + "_C"])
        full["s%d_R_G" % k], full["s%d_t_G" % k] = d["R_G"], d["t_G"].ravel()
    dst = os.path.join(HERE, "kinect_full_scenes.npz")
    np.savez_compressed(dst, **full)
    print("wrote", dst, os.path.getsize(dst), "bytes")


FULL_SCENES = (2, 8, 20, 22)


def variances(C):
    """One variance per point from D x (D*n) blocks that are all v I (with
    v = NaN for the NaN blocks); raises for any other block."""
    D = C.shape[0]
    n = C.shape[1] // D
    B = np.transpose(C.reshape(D, n, D), (1, 0, 2))
    v = B[:, 0, 0].copy()
    rebuilt = v[:, None, None] * np.eye(D)[None]
    off = ~np.eye(D, dtype=bool)
    rebuilt[:, off] = 0.0
    same = np.all((B == rebuilt) | (np.isnan(B) & np.isnan(rebuilt)), axis=(1, 2))
    if not same.all():
        raise ValueError("covariance blocks that are not v I")
    return v


if __name__ == "__main__":
    main()
