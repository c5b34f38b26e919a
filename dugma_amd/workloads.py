"""Synthetic registration workloads (numpy; deterministic per seed).

Seeds are 0xD06A0000 + 1000*config + trial.  Every generator returns
clouds in the MATLAB layout the entry points take: X (D,N) target,
X_C (D,D*N), Y (D,M) source, Y_C (D,D*M), plus the ground-truth pose
(R_G, t_G) with X ~ R_G Y + t_G.

Surfaces (all 3-D configs): r(theta,phi) = 1 + 0.2 * sum_{l=1..4} sum_m
c_lm Y_lm(theta,phi), real spherical harmonics, c_lm ~ N(0,1)/l^2 (l = 0 is
left out because 1/l^2 is undefined there).  Points are drawn area-uniformly
from a fine triangulation of the surface.
"""
import math

import numpy as np
from scipy.special import lpmv

SEED_BASE = 0xD06A0000


def seed(config, trial=0):
    return SEED_BASE + 1000 * config + trial


# ---------------------------------------------------------------------------
# geometry helpers
# ---------------------------------------------------------------------------
def rot_axis_angle(axis, angle):
    axis = np.asarray(axis, float)
    axis = axis / np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + math.sin(angle) * K + (1 - math.cos(angle)) * K @ K


def random_rotation(rng, max_deg):
    axis = rng.normal(size=3)
    ang = math.radians(rng.uniform(0.0, max_deg))
    return rot_axis_angle(axis, ang)


def rot2(theta):
    c, s = math.cos(theta), math.sin(theta)
    return np.array([[c, -s], [s, c]])


def rotation_error_deg(R, R_G):
    c = (np.trace(R @ R_G.T) - 1.0) / 2.0
    return math.degrees(math.acos(min(1.0, max(-1.0, c))))


def _real_sph_harm(l, m, theta, phi):
    am = abs(m)
    norm = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - am) / math.factorial(l + am))
    P = lpmv(am, l, np.cos(theta))
    if m == 0:
        return norm * P
    if m > 0:
        return math.sqrt(2) * norm * P * np.cos(am * phi)
    return math.sqrt(2) * norm * P * np.sin(am * phi)


# Real spherical harmonics here are 4 pi-normalised (mean square 1 over the
# sphere), so 0.2 c_lm is the RMS radial deformation contributed by each term.
# With orthonormal Y_lm (mean square 1 / 4 pi) the same formula stays within
# ~3% of a sphere, whose rotation no registration can recover to 2 degrees.
SH_NORM = math.sqrt(4.0 * math.pi)
R_MIN = 0.25  # coefficient draws whose radius dips below this are redrawn


class Surface:
    """Closed star-shaped surface r(theta, phi) = 1 + 0.2 sum_{l=1..4,m}
    c_lm Y_lm, c_lm ~ N(0, 1) / l^2 (l = 0 left out: 1/l^2 is undefined
    there), Y_lm 4 pi-normalised real spherical harmonics."""

    def __init__(self, rng, lmax=4, nt=160, nph=320):
        th = np.linspace(0, math.pi, nt + 1)
        ph = np.linspace(0, 2 * math.pi, nph + 1)[:-1]
        TH, PH = np.meshgrid(th, ph, indexing="ij")
        while True:
            self.coef = [(l, m, rng.normal() / l ** 2) for l in range(1, lmax + 1)
                         for m in range(-l, l + 1)]
            if self.radius(TH.ravel(), PH.ravel()).min() >= R_MIN:
                break
        V = self.points(TH.ravel(), PH.ravel()).reshape(nt + 1, nph, 3)
        a = np.arange(nt)[:, None]
        b = np.arange(nph)[None, :]
        b2 = (b + 1) % nph
        t1 = np.stack([V[a, b], V[a + 1, b], V[a + 1, b2]], axis=2).reshape(-1, 3, 3)
        t2 = np.stack([V[a, b], V[a + 1, b2], V[a, b2]], axis=2).reshape(-1, 3, 3)
        T = np.stack([t1, t2], axis=1).reshape(-1, 3, 3)  # (nt*nph*2, 3, 3)
        area = 0.5 * np.linalg.norm(np.cross(T[:, 1] - T[:, 0], T[:, 2] - T[:, 0]), axis=1)
        keep = area > 0
        self.T = T[keep]
        self.cdf = np.cumsum(area[keep])
        self.cdf /= self.cdf[-1]

    def radius(self, theta, phi):
        r = np.ones_like(theta)
        for l, m, c in self.coef:
            r = r + 0.2 * SH_NORM * c * _real_sph_harm(l, m, theta, phi)
        return r

    def points(self, theta, phi):
        r = self.radius(theta, phi)
        return np.stack([r * np.sin(theta) * np.cos(phi), r * np.sin(theta) * np.sin(phi),
                         r * np.cos(theta)], axis=-1)

    def sample(self, rng, n):
        k = np.searchsorted(self.cdf, rng.uniform(size=n))
        k = np.minimum(k, len(self.T) - 1)
        u, v = rng.uniform(size=n), rng.uniform(size=n)
        flip = u + v > 1
        u[flip], v[flip] = 1 - u[flip], 1 - v[flip]
        T = self.T[k]
        return T[:, 0] + u[:, None] * (T[:, 1] - T[:, 0]) + v[:, None] * (T[:, 2] - T[:, 0])


def knn_sigma(P, k=8, workers=-1):
    """Mean distance to the k nearest neighbours within the cloud (self excluded)."""
    from scipy.spatial import cKDTree
    d, _ = cKDTree(P).query(P, k=k + 1, workers=workers)
    return d[:, 1:].mean(axis=1)


# Surface covariances of the batched trials' refinement: half the sample
# covariance of a point's 8 nearest neighbours (flat along the surface: its
# spread is the sampling spacing in the tangent plane and the noise and
# curvature along the normal) plus 0.002 times its 8-NN sigma^2 as an
# isotropic floor.  With these, the energy of two independently sampled
# clouds peaks at the true pose (FP64 trials of 500-2,000 points: 32 of 32
# within 2 degrees from the truth and from 20 degrees off, against 0 of 32
# with sigma^2 I).
KNN_COV_SCALE = 0.5
KNN_COV_FLOOR = 0.002


def knn_sigma_cov(P, k=8, scale=KNN_COV_SCALE, floor=KNN_COV_FLOOR, workers=-1):
    """(sigma, C): the mean distance to the k nearest neighbours (self
    excluded) and the covariance scale * (their sample covariance) +
    floor * sigma^2 I, (n, 3, 3)."""
    from scipy.spatial import cKDTree
    d, idx = cKDTree(P).query(P, k=k + 1, workers=workers)
    sig = d[:, 1:].mean(axis=1)
    nb = P[idx[:, 1:]]
    c = nb - nb.mean(axis=1, keepdims=True)
    C = np.einsum("nka,nkb->nab", c, c) / k
    return sig, scale * C + (floor * sig ** 2)[:, None, None] * np.eye(3)


def iso_blocks(sig2):
    """D x (D*n) covariance blocks s_k I for a vector of variances (3-D)."""
    n = len(sig2)
    C = np.zeros((n, 3, 3))
    for a in range(3):
        C[:, a, a] = sig2
    return blocks(C)


def blocks(C):
    """(n,D,D) -> D x (D*n) MATLAB block layout."""
    n, D, _ = C.shape
    return np.ascontiguousarray(np.transpose(C, (1, 0, 2)).reshape(D, D * n))


def unblocks(XC):
    D = XC.shape[0]
    n = XC.shape[1] // D
    return np.transpose(XC.reshape(D, n, D), (1, 0, 2))


# ---------------------------------------------------------------------------
# configurations
# ---------------------------------------------------------------------------
def config1(trial=0, n=1000, m=1000):
    """Plain method, 3-D, FP64: 1,000 x 1,000 points, isotropic 8-NN sigma."""
    rng = np.random.default_rng(seed(1, trial))
    surf = Surface(rng)
    src = surf.sample(rng, m)
    tgt0 = surf.sample(rng, n)
    R = random_rotation(rng, 45.0)
    t = rng.uniform(-0.1, 0.1, size=3)
    tgt = tgt0 @ R.T + t + rng.normal(scale=0.005, size=(n, 3))
    sx, sy = knn_sigma(tgt), knn_sigma(src)
    return dict(X=tgt.T.copy(), X_C=iso_blocks(sx ** 2), Y=src.T.copy(), Y_C=iso_blocks(sy ** 2),
                R_G=R, t_G=t)


def _camera_cov(z):
    sxy = 0.0015 * z
    sz = 0.0012 + 0.0019 * (z - 0.4) ** 2
    C = np.zeros((len(z), 3, 3))
    C[:, 0, 0] = sxy ** 2
    C[:, 1, 1] = sxy ** 2
    C[:, 2, 2] = sz ** 2
    return C


def _to_depth(P, rng):
    """Scale and offset a unit-size cloud so z spans [0.5, 3.0] m."""
    lo, hi = P[:, 2].min(), P[:, 2].max()
    s = 2.5 / (hi - lo)
    Q = P * s
    Q[:, 2] += 0.5 - lo * s
    return Q


def structured_light(rng, P_cam, outlier_frac):
    """Noise from diag(s_xy^2, s_xy^2, s_z^2) in the camera frame, plus
    uniform outliers in the bounding box."""
    n = len(P_cam)
    z = np.clip(P_cam[:, 2], 0.5, 3.0)
    C = _camera_cov(z)
    noise = rng.normal(size=(n, 3)) * np.sqrt(np.diagonal(C, axis1=1, axis2=2))
    P = P_cam + noise
    k = int(round(outlier_frac * n))
    if k:
        lo, hi = P.min(axis=0), P.max(axis=0)
        idx = rng.choice(n, size=k, replace=False)
        P[idx] = rng.uniform(lo, hi, size=(k, 3))
        C[idx] = _camera_cov(np.clip(P[idx, 2], 0.5, 3.0))
    return P, C


def config2(trial=0, n=6000, m=6000):
    """3-D 6,000 x 6,000, structured-light anisotropic covariances, 10% outliers.
    Each cloud is expressed in a world frame rotated away from its camera, so
    every covariance is a full 3x3."""
    rng = np.random.default_rng(seed(2, trial))
    surf = Surface(rng)
    src_cam = _to_depth(surf.sample(rng, m), rng)
    tgt_obj = _to_depth(surf.sample(rng, n), rng)
    R = random_rotation(rng, 30.0)
    c = tgt_obj.mean(axis=0)
    t = c - R @ c + rng.uniform(-0.05, 0.05, size=3)
    tgt_cam = tgt_obj @ R.T + t
    Ys, YC = structured_light(rng, src_cam, 0.10)
    Xs, XC = structured_light(rng, tgt_cam, 0.10)
    # camera -> world rotations make the covariances full
    Wx, Wy = random_rotation(rng, 20.0), random_rotation(rng, 20.0)
    X = Xs @ Wx.T
    XC = np.einsum("ab,nbc,dc->nad", Wx, XC, Wx)
    Y = Ys @ Wy.T
    YC = np.einsum("ab,nbc,dc->nad", Wy, YC, Wy)
    # ground truth in world frames: X = Wx (R Wy^T Y + t)
    R_G = Wx @ R @ Wy.T
    t_G = Wx @ t
    return dict(X=X.T.copy(), X_C=blocks(XC), Y=Y.T.copy(), Y_C=blocks(YC), R_G=R_G, t_G=t_G)


def config3(trial=0, n=200000, m=200000):
    """Large single pair: 200,000 x 200,000, isotropic sigma from 8-NN."""
    rng = np.random.default_rng(seed(3, trial))
    surf = Surface(rng)
    src = surf.sample(rng, m)
    tgt0 = surf.sample(rng, n)
    R = random_rotation(rng, 45.0)
    t = rng.uniform(-0.1, 0.1, size=3)
    tgt = tgt0 @ R.T + t + rng.normal(scale=0.002, size=(n, 3))
    sx, sy = knn_sigma(tgt), knn_sigma(src)
    return dict(X=tgt.T.copy(), X_C=iso_blocks(sx ** 2), Y=src.T.copy(), Y_C=iso_blocks(sy ** 2),
                R_G=R, t_G=t)


def _c4_chunk(args):
    ks, cov = args
    return [config4_trial(k, workers=1, covariances=cov) for k in ks]


def config4(ntrials=4096, procs=None, covariances=False):
    """All batched trials, one seed (and one surface) per trial; large sets
    are generated in worker processes (at most 16).  covariances: also the
    surface covariances CX, CY of the refinement (knn_sigma_cov)."""
    import os
    procs = procs if procs is not None else min(16, os.cpu_count() or 1)
    if ntrials < 256 or procs <= 1:
        return [config4_trial(k, covariances=covariances) for k in range(ntrials)]
    import multiprocessing as mp
    chunks = [(list(range(a, min(a + 64, ntrials))), covariances) for a in range(0, ntrials, 64)]
    with mp.get_context("fork").Pool(procs) as pool:
        out = []
        for part in pool.imap(_c4_chunk, chunks):
            out.extend(part)
    return out


def config4_trial(trial, workers=-1, covariances=False):
    """One batched trial from its own seed: its own surface, N, M
    log-uniform in [500, 8000], angle U(0, 90) deg, noise sigma
    U(0.001, 0.02), outlier fraction U(0, 0.3)."""
    rng = np.random.default_rng(seed(4, trial))
    n = int(round(math.exp(rng.uniform(math.log(500), math.log(8000)))))
    m = int(round(math.exp(rng.uniform(math.log(500), math.log(8000)))))
    surf = Surface(rng, nt=60, nph=120)
    src = surf.sample(rng, m)
    tgt0 = surf.sample(rng, n)
    R = random_rotation(rng, 90.0)
    t = rng.uniform(-0.1, 0.1, size=3)
    noise = rng.uniform(0.001, 0.02)
    outl = rng.uniform(0.0, 0.3)
    tgt = tgt0 @ R.T + t + rng.normal(scale=noise, size=(n, 3))
    k = int(round(outl * n))
    if k:
        lo, hi = tgt.min(axis=0), tgt.max(axis=0)
        idx = rng.choice(n, size=k, replace=False)
        tgt[idx] = rng.uniform(lo, hi, size=(k, 3))
    # isotropic variances per point (Batch takes them directly; iso_blocks
    # turns them into the D x (D*n) layout of init)
    if covariances:
        sx, CX = knn_sigma_cov(tgt, workers=workers)
        sy, CY = knn_sigma_cov(src, workers=workers)
        return dict(X=tgt.T.copy(), sX=sx ** 2, Y=src.T.copy(), sY=sy ** 2, CX=CX, CY=CY,
                    R_G=R, t_G=t, noise=noise, outliers=outl, n=n, m=m)
    sx, sy = knn_sigma(tgt, workers=workers), knn_sigma(src, workers=workers)
    return dict(X=tgt.T.copy(), sX=sx ** 2, Y=src.T.copy(), sY=sy ** 2,
                R_G=R, t_G=t, noise=noise, outliers=outl, n=n, m=m)


def kinect_like(trial=0, n=4600, m=4000, deg=30.0, var=2.25):
    """A registration the size of the reference's Kinect scenes (3,600-5,300
    points per cloud, 8-43 degrees apart, median 30): a rotation of `deg`
    degrees about a random axis of an anisotropic closed surface,
    noise 0.003, isotropic covariances of the Kinect data's size relative to
    the normalised cloud (the regime DUGMA's Lambda / sig weights are made
    for)."""
    rng = np.random.default_rng(seed(6, trial))
    surf = Surface(rng, nt=60, nph=120)
    A = np.diag([1.6, 1.0, 0.6])
    src = surf.sample(rng, m) @ A
    tgt0 = surf.sample(rng, n) @ A
    R = rot_axis_angle(rng.normal(size=3), math.radians(deg))
    t = rng.uniform(-0.1, 0.1, 3)
    tgt = tgt0 @ R.T + t + rng.normal(scale=0.003, size=(n, 3))
    return dict(X=tgt.T.copy(), X_C=iso_blocks(np.full(n, var)), Y=src.T.copy(),
                Y_C=iso_blocks(np.full(m, var)), R_G=R, t_G=t)


def contour2d(rng, n, harmonics=20):
    """Closed 2-D contour from a random Fourier series (coefficients N(0,1)/k)."""
    s = np.sort(rng.uniform(0, 2 * math.pi, size=n))
    r = np.ones(n)
    for k in range(1, harmonics + 1):
        a, b = rng.normal() / k, rng.normal() / k
        r += 0.1 * (a * np.cos(k * s) + b * np.sin(k * s))
    return np.stack([r * np.cos(s), r * np.sin(s)], axis=1)


def config5_2d(trial=0, n=None, m=None):
    """2-D contours, 300-3,000 points, isotropic sigma."""
    rng = np.random.default_rng(seed(5, trial))
    n = n or int(rng.integers(300, 3001))
    m = m or int(rng.integers(300, 3001))
    src = contour2d(np.random.default_rng(seed(5, trial) + 7), m)
    base = contour2d(np.random.default_rng(seed(5, trial) + 7), n)
    th = math.radians(rng.uniform(-45, 45))
    R = rot2(th)
    t = rng.uniform(-0.1, 0.1, size=2)
    tgt = base @ R.T + t + rng.normal(scale=0.003, size=(n, 2))
    from scipy.spatial import cKDTree
    sx = cKDTree(tgt).query(tgt, k=9)[0][:, 1:].mean(axis=1)
    sy = cKDTree(src).query(src, k=9)[0][:, 1:].mean(axis=1)
    CX = np.zeros((n, 2, 2)); CX[:, 0, 0] = CX[:, 1, 1] = sx ** 2
    CY = np.zeros((m, 2, 2)); CY[:, 0, 0] = CY[:, 1, 1] = sy ** 2
    return dict(X=tgt.T.copy(), X_C=blocks(CX), Y=src.T.copy(), Y_C=blocks(CY), R_G=R, t_G=t)


def depth_view(rng, W=320, H=240, f=285.0):
    """3-D view back-projected from a synthetic depth map (smooth random
    height field plus 3 random boxes), pinhole f px, principal point at the
    centre, z in [0.5, 3] m; per-point covariance from the structured-light
    model rotated by the viewing ray (full 3x3)."""
    yy, xx = np.mgrid[0:H, 0:W].astype(float)
    z = np.zeros((H, W))
    for _ in range(6):
        kx, ky = rng.uniform(0.5, 3.0, size=2)
        ph = rng.uniform(0, 2 * math.pi, size=2)
        z += rng.normal() * np.sin(kx * xx / W * 2 * math.pi + ph[0]) * np.cos(ky * yy / H * 2 * math.pi + ph[1])
    z = (z - z.min()) / (z.max() - z.min() + 1e-12)
    z = 1.2 + 1.3 * z
    for _ in range(3):
        x0, y0 = rng.integers(0, W - 40), rng.integers(0, H - 40)
        w, h = rng.integers(20, 80), rng.integers(20, 80)
        z[y0:y0 + h, x0:x0 + w] = rng.uniform(0.5, 1.5)
    z = np.clip(z, 0.5, 3.0)
    X = (xx - W / 2) * z / f
    Y = (yy - H / 2) * z / f
    P = np.stack([X.ravel(), Y.ravel(), z.ravel()], axis=1)
    C = _camera_cov(P[:, 2])
    # rotate the camera-axis covariance onto each viewing ray
    ray = P / np.linalg.norm(P, axis=1, keepdims=True)
    ez = np.array([0.0, 0.0, 1.0])
    out = np.empty_like(C)
    for k0 in range(0, len(P), 20000):
        r = ray[k0:k0 + 20000]
        v = np.cross(np.broadcast_to(ez, r.shape), r)
        s = np.linalg.norm(v, axis=1)
        c = r[:, 2]
        Kx = np.zeros((len(r), 3, 3))
        Kx[:, 0, 1], Kx[:, 0, 2], Kx[:, 1, 2] = -v[:, 2], v[:, 1], -v[:, 0]
        Kx[:, 1, 0], Kx[:, 2, 0], Kx[:, 2, 1] = v[:, 2], -v[:, 1], v[:, 0]
        fac = np.where(s > 1e-12, (1 - c) / np.maximum(s, 1e-12) ** 2, 0.5)
        Rr = np.eye(3)[None] + Kx + fac[:, None, None] * Kx @ Kx
        out[k0:k0 + 20000] = Rr @ C[k0:k0 + 20000] @ np.transpose(Rr, (0, 2, 1))
    return P, out


def config5_3d(trial=0):
    """Two 320 x 240 depth views of the same scene (~76,800 points each)."""
    rng = np.random.default_rng(seed(5, 100 + trial))
    P, C = depth_view(rng)
    R = random_rotation(rng, 10.0)
    t = rng.uniform(-0.05, 0.05, size=3)
    Q = P @ R.T + t
    CQ = np.einsum("ab,nbc,dc->nad", R, C, R)
    return dict(X=Q.T.copy(), X_C=blocks(CQ), Y=P.T.copy(), Y_C=blocks(C), R_G=R, t_G=t)
