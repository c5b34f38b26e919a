#!/usr/bin/env python3
"""Benchmark of the DUGMA registration hot path on one MI355X.

A step is one fused energy + pose-adjoint evaluation (dugma_energy_grad: the
optimiser's gradient callback) over 200,000 x 200,000 3-D points with
isotropic sigma from the 8 nearest neighbours, inputs resident in HBM.  The
value is pair evaluations per second of that whole call.  Beside it the line
reports:
  large      on the same clouds: the same gradient by central finite
             differences (13 forward launches), the k = 8 nearest-neighbour
             search, one full iteration (per-point sigma from the 8
             nearest neighbours of both clouds on the device + the fused
             step), and a registration on the convolved energy;
  full_6k    the fused step on 6,000 x 6,000 full-covariance
             structured-light clouds (10% outliers) and its finite-difference
             gradient;
  points     energy with gradients wrt every point's position and
             covariance on two 320 x 240 depth views;
  contours_2d  the 2-D path: fused step and finite differences on
             3,000 x 3,000 contours;
  batched    4,096 independent registrations in one persistent launch per
             iteration, against one trial per launch;
  reference_path  update_sig, compi and a whole DUGMA.m loop;
  roofline   achieved FP32 FLOP/s of the all-pairs kernel (HIP events on the
             context's stream) against MI355X's 157.3 TFLOP/s;
  cpu_baseline  OpenMP loops of the same computations on the host cores
             (oracle/dugma_cpu_baseline.cpp).
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)



def roofline_constants():
    """The #define values of include/dugma_roofline.h: canonical algorithmic
    FLOP per pair (FMA = 2, exp/rsq/rcp = 1) and the FP32 vector peak."""
    import re
    text = open(os.path.join(ROOT, "include", "dugma_roofline.h")).read()
    return {k: float(v) for k, v in re.findall(r"^#define DUGMA_(\w+) ([0-9.]+)\s*$", text, re.M)}


_RC = roofline_constants()
PEAK_FP32_TFLOPS = _RC["PEAK_FP32_TFLOPS"]  # MI355X dense FP32 vector, AMD public spec
FLOP_3D_FULL_FUSED = int(_RC["FLOP_3D_FULL_FUSED"])
FLOP_3D_FULL_FWD = int(_RC["FLOP_3D_FULL_FWD"])
FLOP_3D_ISO_FUSED = int(_RC["FLOP_3D_ISO_FUSED"])
FLOP_3D_ISO_FWD = int(_RC["FLOP_3D_ISO_FWD"])
# HBM bytes per launch of the headline kernel (fused, isotropic, 200,000 x
# 200,000, the autotuned launch): rocprofv3 FETCH_SIZE and WRITE_SIZE runs,
# each scaled by what the counter reads for a copy kernel of known size in the
# same process (profiles/k_pairs_traffic_200000x200000_iso.txt).
PROFILED_TRAFFIC_200K_ISO = 117246982
TRAFFIC_PLAN = {"lane_points": 1, "target_waves": 524288}  # the launch those counters measured
TRAFFIC_PROFILE = "profiles/k_pairs_traffic_200000x200000_iso.txt"


class Hip:
    """HIP events through the runtime's C API (timing on a given stream)."""

    def __init__(self):
        self.lib = ctypes.CDLL("libamdhip64.so")
        self.lib.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.lib.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.lib.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.lib.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        self.lib.hipEventDestroy.argtypes = [ctypes.c_void_p]

    def copy_calibration(self, mbytes):
        """One device-to-device copy of mbytes MiB (a blit kernel of known
        traffic, for scaling memory counters)."""
        n = int(mbytes) << 20
        a, b = ctypes.c_void_p(), ctypes.c_void_p()
        self.lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.lib.hipFree.argtypes = [ctypes.c_void_p]
        assert self.lib.hipMalloc(ctypes.byref(a), n) == 0 and self.lib.hipMalloc(ctypes.byref(b), n) == 0
        self.lib.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        self.lib.hipMemset(a, 1, n)
        self.sync()
        assert self.lib.hipMemcpy(b, a, n, 3) == 0  # hipMemcpyDeviceToDevice
        self.sync()
        self.lib.hipFree(a)
        self.lib.hipFree(b)

    def event(self):
        e = ctypes.c_void_p()
        assert self.lib.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def sync(self):
        assert self.lib.hipDeviceSynchronize() == 0

    def timed(self, stream, fn, reps):
        """Milliseconds of `reps` calls of fn, bracketed by device syncs and
        timed with events on `stream`."""
        e0, e1 = self.event(), self.event()
        self.sync()
        self.lib.hipEventRecord(e0, ctypes.c_void_p(stream))
        for _ in range(reps):
            fn()
        self.lib.hipEventRecord(e1, ctypes.c_void_p(stream))
        self.lib.hipEventSynchronize(e1)
        self.sync()
        ms = ctypes.c_float()
        self.lib.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
        self.lib.hipEventDestroy(e0)
        self.lib.hipEventDestroy(e1)
        return float(ms.value)


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def start_pose(d, deg=2.0, dt=0.003, seed=0):
    from dugma_amd import workloads as W
    rng = np.random.default_rng(seed)
    return W.random_rotation(rng, deg) @ d["R_G"], np.ravel(d["t_G"]) + rng.uniform(-dt, dt, size=3)


TUNE = True


class CStep:
    """One step as a C caller makes it: dugma_energy_grad (or, with fd_h,
    dugma_energy_grad_fd) through ctypes with the arguments bound once, so
    the timed loop holds the library call and not the Python mirror's
    argument handling.  E and the gradient land in the context's result
    buffers, preset to NaN; every status the calls return is kept."""

    def __init__(self, dg, h, R, t, fd_h=None):
        from dugma_amd import _lib
        self.h = h
        b = h._pb
        b.pose(R, t)
        if fd_h is None:
            self.fn, self.args = _lib.load().dugma_energy_grad, (h.ptr, b.Rp, b.tp, 0, b.Ep, b.gp)
        else:
            self.fn, self.args = (_lib.load().dugma_energy_grad_fd,
                                  (h.ptr, b.Rp, b.tp, 0, ctypes.c_double(fd_h), b.Ep, b.gp))
        self.bad = 0

    def __call__(self):
        if self.fn(*self.args):
            self.bad += 1

    def arm(self):
        b = self.h._pb
        b.E.value = float("nan")
        b.g[:] = np.nan
        self.bad = 0
        return self

    def check(self):
        """Raise unless every timed call returned 0 and the last one left a
        finite energy and gradient."""
        b = self.h._pb
        if self.bad:
            from dugma_amd import _lib
            raise RuntimeError("%d timed calls failed: %s" % (self.bad, _lib.load().dugma_last_error().decode()))
        if not (np.isfinite(b.E.value) and np.all(np.isfinite(b.g))):
            raise RuntimeError("the timed steps did not produce a finite energy and gradient")


def bench_fused(hip, dg, d, steps, warmup, label):
    # init (host arrays -> device state, the one PCIe transfer): the second of
    # two contexts, so that code objects are already loaded
    ok, h0 = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    dg.fr_m(h0)
    t0 = time.perf_counter()
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    init_ms = (time.perf_counter() - t0) * 1e3
    try:
        R, t = start_pose(d)
        plan = h.autotune() if TUNE else (int(os.environ.get("DUGMA_LANE_POINTS", 0)),
                                          int(os.environ.get("DUGMA_TARGET_WAVES", 0)))
        for _ in range(warmup):
            dg.energy_grad(h, R, t)
        step = CStep(dg, h, R, t).arm()
        ms = hip.timed(h.stream(), step, steps)
        step.check()
        # kernel-only duration: HIP events around each pair-kernel launch
        h.set_timing(True)
        for _ in range(max(3, min(steps, 20))):
            dg.energy_grad(h, R, t)
        kms, launches = h.pair_time()
        h.set_timing(False)
        nT, nS, iso = h.energy_info()
        log("%s: %.3f ms/step, pair kernel %.3f ms/launch (%d launches), iso=%s, plan T=%d waves=%d"
            % (label, ms / steps, kms / max(launches, 1), launches, iso, plan[0], plan[1]))
        return dict(ms_per_step=ms / steps, kernel_ms=kms / max(launches, 1), pairs=nT * nS, iso=iso,
                    plan={"lane_points": plan[0], "target_waves": plan[1]}, init_ms=init_ms)
    finally:
        dg.fr_m(h)


def bench_plain(hip, dg, d, steps, step_h=1e-4):
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    try:
        R, t = start_pose(d)
        if TUNE:
            h.autotune()
        dg.energy_grad_fd(h, R, t, step_h)
        step = CStep(dg, h, R, t, fd_h=step_h).arm()
        ms = hip.timed(h.stream(), step, steps)
        step.check()
        nT, nS, _ = h.energy_info()
        return dict(ms_per_gradient=ms / steps, pairs=nT * nS)
    finally:
        dg.fr_m(h)


def bench_full_6k(hip, dg, d2, steps, warmup, plain=True):
    """The fused step on the 6,000 x 6,000 full-covariance structured-light
    clouds, and (plain) its 13-launch central-difference gradient."""
    fused = bench_fused(hip, dg, d2, steps, warmup, "fused 6000x6000 full")
    value = fused["pairs"] / (fused["ms_per_step"] / 1e3) / 1e9
    ach = fused["pairs"] * FLOP_3D_FULL_FUSED / (fused["kernel_ms"] / 1e3) / 1e12
    res = {"workload": "3-D 6,000 x 6,000 full-covariance structured-light clouds, 10% outliers; fused energy + "
                       "6-DoF pose adjoint per step",
           "value": round(value, 3), "unit": "Gpair-evals/s", "ms_per_step": round(fused["ms_per_step"], 4),
           "kernel_ms": round(fused["kernel_ms"], 4), "achieved_tflops": round(ach, 3),
           "frac": round(ach / PEAK_FP32_TFLOPS, 4), "flop_per_pair": FLOP_3D_FULL_FUSED, "launch": fused["plan"],
           "init_ms": round(fused["init_ms"], 3)}
    if plain:
        pl = bench_plain(hip, dg, d2, max(3, steps // 10))
        pv = pl["pairs"] / (pl["ms_per_gradient"] / 1e3) / 1e9
        res["plain_fd"] = {"value": round(pv, 3), "unit": "Gpair-evals/s",
                           "ms_per_gradient": round(pl["ms_per_gradient"], 4),
                           "scheme": "6-DoF central differences, h=1e-4, 13 forward launches",
                           "fused_speedup": round(value / pv, 2)}
        log("plain FD 6K: %.3f ms/gradient, fused speed-up %.2fx" % (pl["ms_per_gradient"], value / pv))
    return res


def bench_register_energy(dg, d):
    """Registration on the convolved energy (Newton steps from each pass's
    pose Hessian, BFGS where that is not positive definite; one fused pass
    per trial pose) on the 200K x 200K clouds, from 5 degrees about a fixed
    axis and 1 cm per axis off the ground truth; beside it the same with
    BFGS alone (DUGMA_REG_NEWTON=0)."""
    from dugma_amd import workloads as W
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    try:
        if TUNE:
            h.autotune()
        R0 = W.rot_axis_angle([1.0, -2.0, 0.5], np.radians(5.0)) @ d["R_G"]
        t0 = np.ravel(d["t_G"]) + np.array([0.01, -0.01, 0.01])
        t1 = time.perf_counter()
        R, t, it, E = dg.register_energy(h, R0, t0)
        sec = time.perf_counter() - t1
        os.environ["DUGMA_REG_NEWTON"] = "0"
        try:
            t2 = time.perf_counter()
            Rb, tb, itb, Eb = dg.register_energy(h, R0, t0)
            secb = time.perf_counter() - t2
        finally:
            del os.environ["DUGMA_REG_NEWTON"]
    finally:
        dg.fr_m(h)
    log("register_energy 200Kx200K: %d steps in %.3f s (BFGS alone: %d steps in %.3f s)" % (it, sec, itb, secb))
    return {"seconds": round(sec, 4), "steps": it, "start_rotation_error_deg": round(W.rotation_error_deg(R0, d["R_G"]), 4),
            "rotation_error_deg": round(W.rotation_error_deg(R, d["R_G"]), 4),
            "translation_error": float(np.linalg.norm(t - np.ravel(d["t_G"]))),
            "bfgs_alone": {"seconds": round(secb, 4), "steps": itb,
                           "rotation_error_deg": round(W.rotation_error_deg(Rb, d["R_G"]), 4)}}


def bench_plain_large(hip, dg, d, big, gradients=2):
    """The finite-difference gradient on the 200K x 200K clouds, beside the fused step measured there."""
    plain = bench_plain(hip, dg, d, gradients)
    speedup = plain["ms_per_gradient"] / big["ms_per_step"]
    log("plain FD 200Kx200K: %.2f ms/gradient, fused speed-up %.2fx" % (plain["ms_per_gradient"], speedup))
    return {"value": round(plain["pairs"] / (plain["ms_per_gradient"] / 1e3) / 1e9, 3), "unit": "Gpair-evals/s",
            "ms_per_gradient": round(plain["ms_per_gradient"], 3),
            "scheme": "6-DoF central differences, h=1e-4, 13 forward launches",
            "fused_speedup": round(speedup, 2)}


def bench_points(hip, dg, W, calls=3):
    """Energy with the gradient wrt every point's position and covariance, on
    two 320 x 240 depth views: two pair passes per call (targets stationary,
    then sources), with the per-point sums copied to the host."""
    d = W.config5_3d()
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    try:
        R, t = start_pose(d)
        dg.energy_grad_points(h, R, t)
        ms = hip.timed(h.stream(), lambda: dg.energy_grad_points(h, R, t), calls) / calls
        h.set_timing(True)
        dg.energy_grad_points(h, R, t)
        kms, launches = h.pair_time()
        h.set_timing(False)
        nT, nS, _ = h.energy_info()
    finally:
        dg.fr_m(h)
    evals = 2 * nT * nS
    ach = evals * FLOP_3D_FULL_FUSED / (kms / 1e3) / 1e12
    log("per-point gradients %dx%d: %.2f ms/call, pair passes %.2f ms (%d launches), %.1f%% of peak"
        % (nT, nS, ms, kms, launches, 100 * ach / PEAK_FP32_TFLOPS))
    return {"workload": "3-D two 320 x 240 depth views, %d x %d points, full covariance; energy + gradient "
                        "wrt every point's position and covariance" % (nT, nS),
            "ms_per_call": round(ms, 3), "kernel_ms": round(kms, 3), "launches": int(launches),
            "value": round(evals / (kms / 1e3) / 1e9, 3), "unit": "Gpair-evals/s (two passes per call)",
            "achieved_tflops": round(ach, 3), "frac": round(ach / PEAK_FP32_TFLOPS, 4),
            "flop_per_pair": FLOP_3D_FULL_FUSED}


def bench_contours_2d(hip, dg, W, steps=200):
    """The 2-D path: fused energy + pose adjoint (theta, t) on the largest
    contours of the mixed 2-D / 3-D configuration (3,000 x 3,000 points,
    isotropic sigma from the 8-NN spacing), and the same gradient by central
    differences."""
    d = W.config5_2d(n=3000, m=3000)
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    try:
        if TUNE:
            h.autotune()
        R, t = d["R_G"], np.ravel(d["t_G"])
        for _ in range(5):
            dg.energy_grad(h, R, t)
        step = CStep(dg, h, R, t).arm()
        ms = hip.timed(h.stream(), step, steps) / steps
        step.check()
        step = CStep(dg, h, R, t, fd_h=1e-4).arm()
        fd = hip.timed(h.stream(), step, steps // 10) / (steps // 10)
        step.check()
        nT, nS, iso = h.energy_info()
    finally:
        dg.fr_m(h)
    log("2-D contours %dx%d: %.4f ms/step, FD %.4f ms/gradient" % (nT, nS, ms, fd))
    return {"workload": "2-D contours, %d x %d points, isotropic; fused energy + (theta, t) adjoint" % (nT, nS),
            "ms_per_step": round(ms, 4), "value": round(nT * nS / (ms / 1e3) / 1e9, 3), "unit": "Gpair-evals/s",
            "plain_fd_ms_per_gradient": round(fd, 4), "fused_speedup": round(fd / ms, 2),
            "fd_scheme": "central differences in (theta, t), h=1e-4, 7 forward launches"}


# Batched schedule.  The isotropic phase: DUGMA's sigma update (the kernel
# variance tracks the mean nearest distance) on top of an annealed extra
# variance, pose steps over-relaxed by 1.9 once annealing has ended, 120
# iterations on every 4th target and source point (1/16 of the pairs).  Each
# trial's points are in spatial order, so every 4th is an evenly spread
# subsample.  Then
# the refinement on the energy with surface covariances (half the 8-NN
# sample covariance plus 0.002 sigma^2 I), one fused energy + pose-gradient
# pass per trial pose, Newton steps from the pass's approximate Hessian (BFGS
# where it is not positive definite) per trial: at most 60 steps on the same coarse
# points (covariances times 4), at most 20 on every 2nd point (covariances
# times 2), then at most 100 on the full clouds, each level starting from the
# previous one's pose and scaled inverse Hessian, until the rotation step is below
# 1e-4 rad and the translation step below 1e-4 (as the reference's own
# driver stops, main_kinect_real_application.m:14-17).  The refinement
# converges from tens of degrees, so the isotropic phase only has to bring
# the trials into its basin.
# Measured on the 4,096 trials (success < 2 deg, registrations/s; coarse
# isotropic / full isotropic iterations, coarse / full refinement steps, tol):
#   160 / 300 / 0 / 100, 1e-4: 89.2%, 1,375/s;  160 / 40 / 0 / 100: 86.7%,
#   2,221-2,355/s;  160 / 0 / 0 / 100: 85.2%, 2,745-2,919/s;
#   160 / 0 / 40 / 100: 88.1%, 3,663-3,735/s (fine level started afresh);
#   with the coarse inverse Hessian carried over: 160 / 0 / 60 / 100, 1e-4:
#   88.4%, 5,432/s;  the same at 2e-4: 88.3%, 5,562/s;  120 / 0 / 60 / 100,
#   2e-4: 88.1%, 5,797/s.  With a launch shape per level and the middle level
#   (coarse isotropic / coarse / middle refinement steps, tol 1e-4): 160 / 60 /
#   0: 88.4%, 6,376/s;  160 / 60 / 20: 88.8%, 6,220/s;  160 / 60 / 40: 88.9%,
#   6,165/s;  100 / 60 / 30: 88.7%, 6,318/s;  80 / 60 / 30: 88.6%, 6,310/s
#   (run-to-run spread of the rate about 4%).  With the points in spatial
#   order and the refinement's far-field skip: 160 / 60 / 20: 91.9%, 7,983/s;
#   120 / 60 / 20: 91.8%, 8,822/s;  100 / 60 / 20: 91.5%, 8,992/s;  80 / 60 /
#   20: 91.4%, 9,708/s;  120 / 60 / 0: 91.5%, 9,180/s;  120 / 40 / 10: 91.6%,
#   8,301/s.
BATCH_COARSE_ITERS = 120
BATCH_REFINE_ITERS = 100
BATCH_REFINE_COARSE_ITERS = 60
BATCH_REFINE_MID_ITERS = 20
BATCH_MODE = dict(anneal=True, anneal_decay=0.95, sigma_update=True, tol_rot=1e-4, tol_trans=1e-4, relax=1.9,
                  coarse_iters=BATCH_COARSE_ITERS, refine_iters=BATCH_REFINE_ITERS,
                  refine_coarse_iters=BATCH_REFINE_COARSE_ITERS, refine_mid_iters=BATCH_REFINE_MID_ITERS)
BATCH_MAX_ITER = BATCH_COARSE_ITERS  # no full-cloud isotropic iterations
# the isotropic phase alone, for comparison: up to 300 full-cloud iterations
BATCH_ISO_MAX_ITER = 300 + BATCH_COARSE_ITERS


def bench_batched(dg, W, ntrials, baseline_trials, max_iter=BATCH_MAX_ITER):
    """4,096 independent registrations: one persistent pair launch per
    iteration over a (trial, tile) queue, iterations replayed from a hipGraph;
    against the same kernels one trial per launch on a subset."""
    t0 = time.time()
    trials = W.config4(ntrials=ntrials, covariances=True)
    log("generated %d trials in %.1f s" % (ntrials, time.time() - t0))
    b = dg.Batch(trials)
    try:
        b.register(max_iter=2, **BATCH_MODE)  # warm-up: code objects, graph path
        ta = time.perf_counter()
        shape = b.autotune(sigma_update=BATCH_MODE["sigma_update"])
        tune_s = time.perf_counter() - ta
        info = b.info()
        b.register(max_iter=2, **BATCH_MODE)
        t1 = time.perf_counter()
        out = b.register(max_iter=max_iter, use_graph=True, check_every=10, **BATCH_MODE)
        wall = time.perf_counter() - t1
        # diagnostic: the same registrations started at the ground truth (how
        # far the energy maximum itself sits from it): one full-cloud
        # isotropic step, then the full-cloud refinement
        R_G = np.array([tr["R_G"] for tr in trials])
        t_G = np.array([tr["t_G"] for tr in trials])
        outg = b.register(R_G, t_G, max_iter=max(1, max_iter - BATCH_COARSE_ITERS), use_graph=True,
                          **{**BATCH_MODE, "anneal": False, "coarse_iters": 0, "refine_coarse_iters": 0,
                             "refine_mid_iters": 0})
        # the isotropic iterations alone (no refinement), for comparison
        t3 = time.perf_counter()
        outi = b.register(max_iter=BATCH_ISO_MAX_ITER, use_graph=True, check_every=10,
                          **{**BATCH_MODE, "refine_iters": 0})
        wall_iso = time.perf_counter() - t3
    finally:
        b.free()
    errs = np.array([W.rotation_error_deg(out["R"][k], tr["R_G"]) for k, tr in enumerate(trials)])
    errs_g = np.array([W.rotation_error_deg(outg["R"][k], tr["R_G"]) for k, tr in enumerate(trials)])
    errs_i = np.array([W.rotation_error_deg(outi["R"][k], tr["R_G"]) for k, tr in enumerate(trials)])
    mn = np.array([min(tr["n"], tr["m"]) for tr in trials])
    bands = {}
    for lo, hi in ((500, 1000), (1000, 2000), (2000, 4000), (4000, 8001)):
        sel = (mn >= lo) & (mn < hi)
        if sel.any():
            bands["%d-%d" % (lo, min(hi, 8000))] = {"trials": int(sel.sum()),
                                                    "success": round(float((errs[sel] < 2.0).mean()), 4),
                                                    "success_from_ground_truth": round(float((errs_g[sel] < 2.0).mean()), 4)}
    sub = trials[:baseline_trials]
    b1 = dg.Batch(sub)
    try:
        for lvl, sh in enumerate(shape["levels"]):  # same sums as the batched run
            b1.set_level_shape(lvl, sh["waves"], sh["chunks"], sh["grid_waves"])
        b1.register(max_iter=2, one_trial_per_launch=True, **BATCH_MODE)
        t2 = time.perf_counter()
        out1 = b1.register(max_iter=max_iter, one_trial_per_launch=True, **BATCH_MODE)
        wall1 = time.perf_counter() - t2
    finally:
        b1.free()
    same = bool(np.array_equal(out1["R"], out["R"][:baseline_trials]))
    rps = ntrials / wall
    rps1 = baseline_trials / wall1
    log("batched: %d trials in %.3f s (%.1f reg/s), one-trial-per-launch %.1f reg/s, success %.3f, shape %s"
        % (ntrials, wall, rps, rps1, float((errs < 2.0).mean()), shape))
    return {"workload": "4,096 trials, N and M log-uniform in [500, 8000], rotation U(0,90) deg, "
                        "noise U(0.001,0.02), outliers U(0,0.3); isotropic 8-NN sigma, then surface covariances",
            "value": round(rps, 2), "unit": "registrations/s", "wall_s": round(wall, 4),
            "iteration": "isotropic phase: majorise-minimise pose step; DUGMA sigma update%s; steps "
                         "over-relaxed by %g after annealing; %d iterations on every 4th point, %d on the "
                         "full clouds" %
                         (" + annealed extra variance (decay %g)" % BATCH_MODE["anneal_decay"]
                          if BATCH_MODE["anneal"] else "", BATCH_MODE["relax"],
                          min(BATCH_MODE["coarse_iters"], max_iter), max(0, max_iter - BATCH_MODE["coarse_iters"])),
            "converged_fraction": round(float((out["status"] == 1).mean()), 4),
            "iterations_mean": float(out["iters"].mean()), "max_iter": max_iter,
            "pairs_per_iteration": info["pairs_per_iteration"], "work_items": info["items"],
            "autotune": {**shape, "seconds": round(tune_s, 3)},
            "success_rate": round(float((errs < 2.0).mean()), 4),
            "rotation_error_below_5deg": round(float((errs < 5.0).mean()), 4),
            "rotation_error_median_deg": round(float(np.median(errs)), 4),
            "success_rate_from_ground_truth": round(float((errs_g < 2.0).mean()), 4),
            "success_by_min_points": bands,
            "refine": "Newton / BFGS per trial on the energy with surface covariances (half the 8-NN sample covariance "
                      "+ 0.002 sigma^2 I), one fused energy + pose-gradient pass per trial pose: at most %d steps "
                      "on every 4th point (covariances x 4), %d on every 2nd (x 2), then at most %d on the full "
                      "clouds; stop at rotation step < %g rad and translation step < %g"
                      % (BATCH_MODE["refine_coarse_iters"], BATCH_MODE["refine_mid_iters"],
                         BATCH_MODE["refine_iters"], BATCH_MODE["tol_rot"], BATCH_MODE["tol_trans"]),
            "without_refinement": {"value": round(ntrials / wall_iso, 2), "unit": "registrations/s",
                                   "max_iter": BATCH_ISO_MAX_ITER,
                                   "success_rate": round(float((errs_i < 2.0).mean()), 4),
                                   "rotation_error_below_5deg": round(float((errs_i < 5.0).mean()), 4)},
            "one_trial_per_launch": {"value": round(rps1, 2), "trials": baseline_trials,
                                     "wall_s": round(wall1, 4), "same_poses": same},
            "speedup": round(rps / rps1, 2)}


def bench_knn(dg, d, k=8, reps=10):
    """k nearest transformed sources of every target (the sigma update of the
    200K x 200K configuration), wall time per call including the copy of the
    N x k indices and distances to the host.  Above 2^24 pairs the library
    searches a uniform grid (exact: the all-pairs results, bit for bit), so
    the rate is effective: N x M pairs per second of the call."""
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    try:
        R, t = d["R_G"], d["t_G"]
        for _ in range(2):
            dg.knn(h, dg.KNN_TARGET_TO_SOURCE, k, R, t)  # warm-up
        t0 = time.perf_counter()
        for _ in range(reps):
            dg.knn(h, dg.KNN_TARGET_TO_SOURCE, k, R, t)
        ms = (time.perf_counter() - t0) / reps * 1e3
        pairs = float(h.N) * h.M
    finally:
        dg.fr_m(h)
    log("knn k=%d: %.3f ms per call" % (k, ms))
    return {"k": k, "ms_per_call": round(ms, 3), "value": round(pairs / (ms / 1e3) / 1e9, 3),
            "unit": "effective Gpair-distance-evals/s", "search": "uniform grid, exact (all-pairs results)"}


def bench_reference_path(dg, W, d2):
    """The reference's own GPU calls and loop: update_sig and compi per call
    on the 6,000 x 6,000 workload, and one whole DUGMA.m registration
    (dugma_register) of a Kinect-sized problem, timed end to end."""
    ok, h = dg.init(d2["X"], d2["X_C"], d2["Y"], d2["Y_C"])
    try:
        R, t = start_pose(d2)
        sig = dg.update_sig(h, R, t, h.N, h.M, 3)
        dg.compi(h, R, t, h.N, h.M, 3, sig)
        reps = 10
        t0 = time.perf_counter()
        for _ in range(reps):
            sig = dg.update_sig(h, R, t, h.N, h.M, 3)
        us_ms = (time.perf_counter() - t0) / reps * 1e3
        t0 = time.perf_counter()
        for _ in range(reps):
            dg.compi(h, R, t, h.N, h.M, 3, sig)
        co_ms = (time.perf_counter() - t0) / reps * 1e3
        pairs = float(h.N) * h.M
    finally:
        dg.fr_m(h)
    k = W.kinect_like()
    dg.DUGMA(k["X"], k["Y"], k["X_C"], k["Y_C"], np.eye(3), np.zeros(3), 2)  # warm-up
    t0 = time.perf_counter()
    Rk, tk, it, sg = dg.DUGMA(k["X"], k["Y"], k["X_C"], k["Y_C"], np.eye(3), np.zeros(3), 300, 1e-4, 1e-8, 1e-9)
    wall = time.perf_counter() - t0
    log("update_sig %.3f ms, compi %.3f ms per call (6K x 6K); DUGMA loop %d iterations in %.3f s"
        % (us_ms, co_ms, it, wall))
    scenes = bench_kinect_scenes(dg, W)
    return {"update_sig": {"ms_per_call": round(us_ms, 4), "value": round(pairs / (us_ms / 1e3) / 1e9, 3),
                           "unit": "Gpair-evals/s", "workload": "6,000 x 6,000"},
            "compi": {"ms_per_call": round(co_ms, 4), "value": round(pairs / (co_ms / 1e3) / 1e9, 3),
                      "unit": "Gpair-evals/s", "workload": "6,000 x 6,000, 637 coefficients"},
            "dugma_loop": {"workload": "4,600 x 4,000 Kinect-sized synthetic registration, Kinect-scale "
                                       "covariances, 30 degrees about a random axis; Max_Iteration 300, accuracies "
                                       "1e-4 / 1e-8 / 1e-9 (main_kinect_real_application.m:14-17)",
                           "seconds": round(wall, 4), "iterations": int(it),
                           "ms_per_iteration": round(wall / max(it, 1) * 1e3, 3),
                           "rotation_error_deg": round(float(W.rotation_error_deg(Rk, k["R_G"])), 4),
                           "start_error_deg": round(float(W.rotation_error_deg(np.eye(3), k["R_G"])), 4),
                           "reference_published_seconds": "24.5-631.7 per Kinect scene (median 145), "
                                                          "GTX 1080 Ti + MATLAB fmincon, printing included"},
            "kinect_scenes": scenes}


KINECT_SCENES = (2, 8, 20, 22)  # the whole scenes committed under tests/golden


def bench_kinect_scenes(dg, W):
    """The reference's own Kinect scenes (dataset_all/Kinect_Application,
    committed whole for scenes 2, 8, 20 and 22) registered with its driver's
    settings (main_kinect_real_application.m:14-21: R0 = I, t0 = 0, 300
    iterations, accuracies 1e-4 / 1e-8 / 1e-9) by dugma_register, timed end
    to end, beside the outcome and wall time the reference published for
    each scene (Ours_k_result.mat; GTX 1080 Ti with MATLAB's fmincon)."""
    here = os.path.dirname(os.path.abspath(__file__))
    g = np.load(os.path.join(here, "tests", "golden", "kinect_full_scenes.npz"))
    pub = np.load(os.path.join(here, "tests", "golden", "kinect_fixtures.npz"))
    rows = []
    for k in KINECT_SCENES:
        d = {"X": g["s%d_X" % k], "Y": g["s%d_Y" % k]}
        for name in ("X", "Y"):
            v = g["s%d_v%s" % (k, name)]
            d[name + "_C"] = W.blocks(v[:, None, None] * np.eye(3))
        R_G, t_G = g["s%d_R_G" % k], g["s%d_t_G" % k]
        args = (d["X"], d["Y"], d["X_C"], d["Y_C"], np.eye(3), np.zeros(3))
        dg.DUGMA(*args, 2)  # warm-up at this size
        t0 = time.perf_counter()
        R, t, it, sig = dg.DUGMA(*args, 300, 1e-4, 1e-8, 1e-9)
        wall = time.perf_counter() - t0
        Rp, tp = pub["published_R"][k - 1], pub["published_t"][k - 1]
        rows.append({"scene": k, "N": int(d["X"].shape[1]), "M": int(d["Y"].shape[1]), "seconds": round(wall, 4),
                     "iterations": int(it), "rotation_error_deg": round(float(W.rotation_error_deg(R, R_G)), 4),
                     "translation_error": round(float(np.linalg.norm(t - t_G)), 5),
                     "published_rotation_error_deg": round(float(W.rotation_error_deg(Rp, R_G)), 4),
                     "published_translation_error": round(float(np.linalg.norm(tp - t_G)), 5),
                     "published_seconds": round(float(pub["published_time"][k - 1]), 2),
                     "gap_to_published_pose_deg": round(float(W.rotation_error_deg(R, Rp)), 4)})
    log("Kinect scenes %s: %s s (published %s s)" % (list(KINECT_SCENES), [r["seconds"] for r in rows],
                                                      [r["published_seconds"] for r in rows]))
    return {"settings": "R0 = I, t0 = 0, Max_Iteration 300, accuracies 1e-4 / 1e-8 / 1e-9 "
                        "(main_kinect_real_application.m:14-21)",
            "scenes": rows,
            "note": "published_seconds: the reference's own run (GTX 1080 Ti, MATLAB fmincon, coefficients "
                    "printed every iteration), not a like-for-like comparison of hardware"}


def bench_iteration(hip, dg, d, steps=5):
    """One iteration of the large configuration as the device runs it: the
    per-point sigma of both clouds from their 8 nearest neighbours
    (dugma_update_sigma_knn: two within-cloud k-NN searches, mean
    distances, covariances rewritten in place), then the fused step at the
    new covariances.  Both calls through ctypes with bound arguments."""
    from dugma_amd import _lib
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    try:
        R, t = start_pose(d)
        if TUNE:
            h.autotune()
        un = ctypes.c_int(0)
        sfn = _lib.load().dugma_update_sigma_knn
        sargs = (h.ptr, dg.SIGMA_TARGET | dg.SIGMA_SOURCE, 8, None, None, ctypes.byref(un))
        step = CStep(dg, h, R, t)
        bad = [0]

        def sigma():
            if sfn(*sargs):
                bad[0] += 1

        def iteration():
            sigma()
            step()

        for _ in range(2):
            iteration()
        step.arm()
        sig_ms = hip.timed(h.stream(), sigma, steps) / steps
        it_ms = hip.timed(h.stream(), iteration, steps) / steps
        step.check()
        if bad[0]:
            raise RuntimeError("%d sigma updates failed: %s" % (bad[0], _lib.load().dugma_last_error().decode()))
        nT, nS, iso = h.energy_info()
    finally:
        dg.fr_m(h)
    log("iteration 200Kx200K: sigma update %.3f ms, sigma + fused step %.3f ms" % (sig_ms, it_ms))
    return {"what": "per-point sigma of both clouds from their 8 nearest neighbours (device k-NN, mean distance, "
                    "covariances rewritten in place), then the fused energy + pose adjoint",
            "ms_per_iteration": round(it_ms, 3), "sigma_update_ms": round(sig_ms, 3), "isotropic": bool(iso),
            "unchanged_points": int(un.value), "value": round(nT * nS / (it_ms / 1e3) / 1e9, 3),
            "unit": "Gpair-evals/s (fused pairs per iteration time)"}


def bench_iteration_surface(hip, dg, d, steps=3):
    """The same clouds with surface covariances: dugma_update_cov_knn on both
    clouds (8-NN sample covariance x 0.5 + 0.002 sigma^2 I, on the device),
    then the fused step with full covariances (107 FLOP per pair)."""
    from dugma_amd import _lib
    ok, h = dg.init(d["X"], d["X_C"], d["Y"], d["Y_C"])
    try:
        R, t = start_pose(d)
        un = ctypes.c_int(0)
        cfn = _lib.load().dugma_update_cov_knn
        cargs = (h.ptr, dg.SIGMA_TARGET | dg.SIGMA_SOURCE, 8, ctypes.c_double(0.5), ctypes.c_double(0.002),
                 None, None, ctypes.byref(un))
        bad = [0]

        def cov():
            if cfn(*cargs):
                bad[0] += 1

        cov()
        if TUNE:
            h.autotune()
        step = CStep(dg, h, R, t)
        step()
        cov_ms = hip.timed(h.stream(), cov, steps) / steps
        step.arm()
        ms = hip.timed(h.stream(), step, steps) / steps
        step.check()
        if bad[0]:
            raise RuntimeError("%d covariance updates failed: %s" % (bad[0], _lib.load().dugma_last_error().decode()))
        h.set_timing(True)
        dg.energy_grad(h, R, t)
        kms, launches = h.pair_time()
        h.set_timing(False)
        nT, nS, iso = h.energy_info()
    finally:
        dg.fr_m(h)
    kms /= max(launches, 1)
    ach = nT * nS * FLOP_3D_FULL_FUSED / (kms / 1e3) / 1e12
    log("surface iteration 200Kx200K: covariance update %.3f ms, full-covariance fused step %.3f ms (kernel %.3f, "
        "%.1f%% of peak)" % (cov_ms, ms, kms, 100 * ach / PEAK_FP32_TFLOPS))
    return {"what": "surface covariances of both clouds from their 8 nearest neighbours on the device "
                    "(0.5 x sample covariance + 0.002 sigma^2 I), then the fused energy + pose adjoint with full "
                    "covariances", "covariance_update_ms": round(cov_ms, 3), "ms_per_step": round(ms, 3),
            "kernel_ms": round(kms, 3), "isotropic": bool(iso), "value": round(nT * nS / (ms / 1e3) / 1e9, 3),
            "unit": "Gpair-evals/s", "achieved_tflops": round(ach, 3), "frac": round(ach / PEAK_FP32_TFLOPS, 4),
            "flop_per_pair": FLOP_3D_FULL_FUSED, "unchanged_points": int(un.value)}


def cpu_baseline(d3, d2, n200=1024, n_compi=300, reps=5):
    """OpenMP loops of the same computations on the host cores
    (oracle/dugma_cpu_baseline.cpp; steady_clock, median of `reps` runs after
    one untimed run).  The headline leg is the fused FP32 energy + pose
    gradient on n200 of the 200,000 targets against all 200,000 sources."""
    from dugma_amd import workloads as W
    from oracle import cpu_baseline as CB
    from oracle import oracle as O

    def leg(pairs, seconds, **extra):
        return {"value": round(pairs / seconds / 1e9, 4), "unit": "Gpair-evals/s", "seconds": round(seconds, 5),
                **extra}

    t_all = time.perf_counter()
    Y, X = d3["X"].T, d3["Y"].T
    vy, vx = W.unblocks(d3["X_C"])[:, 0, 0], W.unblocks(d3["Y_C"])[:, 0, 0]
    R, t = start_pose(d3)
    n200 = min(n200, Y.shape[0])
    pairs = n200 * X.shape[0]
    scale = Y.shape[0] / n200
    large = {}
    for name, mode, prec in (("forward_fp32", "forward", 32), ("fused_fp32", "fused", 32),
                             ("fused_fp64", "fused", 64)):
        _, _, ts = CB.energy(Y[:n200], vy[:n200], X, vx, R, t, mode=mode, prec=prec, reps=reps)
        large[name] = leg(pairs, ts, full_size_seconds=round(ts * scale, 3))
    large["fd_fp32_full_size_seconds"] = round(13 * large["forward_fp32"]["full_size_seconds"], 2)
    large["sample"] = ("%d of the %d targets x all %d sources (%.3g pairs); full-size seconds scale the sample "
                       "by %d/%d, finite differences extrapolated as 13 forward passes"
                       % (n200, Y.shape[0], X.shape[0], pairs, Y.shape[0], n200))
    # 6,000 x 6,000 full covariance
    Y6, X6 = d2["X"].T, d2["Y"].T
    SY6, SX6 = W.unblocks(d2["X_C"]), W.unblocks(d2["Y_C"])
    R6, t6 = start_pose(d2)
    p6 = Y6.shape[0] * X6.shape[0]
    full = {}
    for prec in (32, 64):
        for mode in ("forward", "fused", "fd"):
            _, _, ts = CB.energy(Y6, SY6, X6, SX6, R6, t6, mode=mode, prec=prec, reps=reps)
            full["%s_fp%d" % (mode, prec)] = leg(p6, ts)
    # the reference's own loops on the 6K clouds
    Xf, Yf = d2["X"].T.astype(np.float32), d2["Y"].T.astype(np.float32)
    sig, ts = CB.update_sig(Xf, Yf, R6, t6, reps=reps)
    ref = {"update_sig": leg(p6, ts, what="quantised nearest distance, FP32 (update_sig.cu:238-260)")}
    XCi, gX, _ = O.prep_inverse(SY6, 3)
    YCi, gY, _ = O.prep_inverse(SX6, 3)
    nc = min(n_compi, Xf.shape[0])
    _, ts = CB.compi(Xf[:nc], XCi[:nc], gX[:nc], Yf, YCi, gY, R6, t6, sig, form=0, reps=max(1, reps // 2))
    ref["compi_per_pair"] = leg(nc * Yf.shape[0], ts, full_size_seconds=round(ts * Xf.shape[0] / nc, 3),
                                what="P (FP64 exp) and the 637-coefficient expansion of every pair "
                                     "(compi.cu:747-3350), on %d of the 6,000 targets" % nc)
    _, ts = CB.compi(Xf, XCi, gX, Yf, YCi, gY, R6, t6, sig, form=1, reps=reps)
    ref["compi_moments"] = leg(p6, ts, what="P and 20 target features per pair, one expansion per source")
    head = large["fused_fp32"]
    return {"value": head["value"], "unit": "Gpair-evals/s", "cores": CB.num_threads(), "kind": "port",
            "cpu_model": CB.cpu_model(), "nproc": os.cpu_count(),
            "sample": "fused FP32 energy + pose gradient (FP64 per-target sums) on %d of the 200,000 targets x all "
                      "200,000 sources of the headline workload, median of %d runs after one untimed run"
                      % (n200, reps),
            "build": "g++ -O3 -march=x86-64-v3 -ffast-math -fopenmp (vectorised expf/exp)",
            "legs": {"large_200k_iso": large, "full_6k": full, "reference_loops_6k": ref},
            "wall_s": round(time.perf_counter() - t_all, 2)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-large", action="store_true", help="skip the 200K extras (FD, k-NN, iteration, registration)")
    ap.add_argument("--no-6k", action="store_true", help="skip the 6,000 x 6,000 full-covariance measurement")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched registrations")
    ap.add_argument("--no-points", action="store_true", help="skip the per-point gradients and the 2-D contours")
    ap.add_argument("--no-loop", action="store_true", help="skip update_sig / compi / DUGMA-loop timings")
    ap.add_argument("--trials", type=int, default=4096, help="batched trials")
    ap.add_argument("--only", choices=["fused", "fused6k", "plain", "iteration", "large", "points", "batch", "loop",
                                       "cpu"], default=None, help="run one measurement (for profiler runs)")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the launch shape from DUGMA_LANE_POINTS / DUGMA_TARGET_WAVES")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="first do one device-to-device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path (targets or whole trials "
              "would be sharded across GPUs with one all-reduce per iteration); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import dugma_amd as dg
    from dugma_amd import workloads as W
    hip = Hip()
    global TUNE
    TUNE = not args.no_tune
    if args.copy_mb:
        hip.copy_calibration(args.copy_mb)
    only = args.only
    if only == "cpu":
        print(json.dumps({"only": only, **cpu_baseline(W.config3(), W.config2(n=6000, m=6000))}), flush=True)
        return
    if only in ("points", "batch", "loop", "fused6k", "plain"):
        if only == "points":
            res = {**bench_points(hip, dg, W), "contours_2d": bench_contours_2d(hip, dg, W)}
        elif only == "loop":
            res = bench_reference_path(dg, W, W.config2(n=6000, m=6000))
        elif only == "batch":
            res = bench_batched(dg, W, args.trials, min(128, args.trials))
        else:
            d2 = W.config2(n=6000, m=6000)
            res = bench_full_6k(hip, dg, d2, args.steps, args.warmup, plain=(only == "plain"))
        print(json.dumps({"only": only, **res}), flush=True)
        return

    log("generating 200,000 x 200,000 workload")
    d3 = W.config3()
    if only == "iteration":
        print(json.dumps({"only": only, **bench_iteration(hip, dg, d3),
                          "surface": bench_iteration_surface(hip, dg, d3)}), flush=True)
        return
    fused = bench_fused(hip, dg, d3, args.steps, args.warmup, "fused 200Kx200K iso")
    value = fused["pairs"] / (fused["ms_per_step"] / 1e3) / 1e9
    achieved = fused["pairs"] * FLOP_3D_ISO_FUSED / (fused["kernel_ms"] / 1e3) / 1e12
    out = {
        "metric": "Gpair-evals/s energy+adjoint (fused vs plain FD), batched registrations/s",
        "value": round(value, 3),
        "unit": "Gpair-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(fused["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "3-D 200,000 x 200,000 points (spherical-harmonic surface), isotropic sigma from the "
                               "8 nearest neighbours; fused energy + 6-DoF pose adjoint per step (per-pair FP32, "
                               "FP64 sums)",
                   "target_points": int(d3["X"].shape[1]), "source_points": int(d3["Y"].shape[1]),
                   "parallelism": "single GPU"},
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                     "traffic": PROFILED_TRAFFIC_200K_ISO if fused["plan"] == TRAFFIC_PLAN else None,
                     "traffic_source": TRAFFIC_PROFILE,
                     "kernel_ms": round(fused["kernel_ms"], 4), "flop_per_pair": FLOP_3D_ISO_FUSED,
                     "launch": fused["plan"]},
        "init_ms": round(fused["init_ms"], 3),
    }
    if only == "fused":
        print(json.dumps(out), flush=True)
        return
    if not args.no_large:
        out["large"] = {"plain_fd": bench_plain_large(hip, dg, d3, fused), "knn": bench_knn(dg, d3),
                        "iteration": bench_iteration(hip, dg, d3),
                        "iteration_surface": bench_iteration_surface(hip, dg, d3),
                        "register_energy": bench_register_energy(dg, d3)}
        if only == "large":
            print(json.dumps({"only": only, **out["large"]}), flush=True)
            return
    d2 = None
    if not args.no_6k:
        log("generating 6000 x 6000 workload")
        d2 = W.config2(n=6000, m=6000)
        out["full_6k"] = bench_full_6k(hip, dg, d2, args.steps, args.warmup)
    if not args.no_points:
        out["points"] = bench_points(hip, dg, W)
        out["contours_2d"] = bench_contours_2d(hip, dg, W)
    if not args.no_batch:
        out["batched"] = bench_batched(dg, W, args.trials, min(128, args.trials))
    if not args.no_loop:
        out["reference_path"] = bench_reference_path(dg, W, d2 if d2 is not None else W.config2(n=6000, m=6000))
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(d3, d2 if d2 is not None else W.config2(n=6000, m=6000))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
