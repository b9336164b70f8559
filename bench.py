#!/usr/bin/env python3
"""Benchmark of the fused collide-stream path on one MI355X.

Workload: the D3Q19 precision-study vessel (Y-bifurcation, parent radius 24,
daughters radius 18 at 35 degrees, 3-mode wall perturbation, ~3M nodes),
Womersley inlet (Wo = 4, period 4000 steps), zero-pressure outlets, half-way
bounce-back walls. The headline `value` is the whole-job MLUPS of the
FP32-shifted path; FP64 and FP16-storage runs of the same workload are
reported beside it with their velocity / WSS error against FP64.

A "step" is one fused kernel launch = one reference iteration over all nodes.
Inputs are resident in HBM when the timed region starts.

    python bench.py [--gpus 1] [--steps 8000] [--warmup 100]

The FP32/FP16-vs-FP64 errors are taken after the workload's own 8000 steps
(--precision-steps), whatever --steps is; that FP32 run is also the
sustained measurement behind `roofline` (its mean step-kernel launch time),
and the --steps batch is reported beside it as `roofline.burst`.
"""
import argparse
import json
import math
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "reza-seti_amd"))

import lbm  # noqa: E402
import workloads  # noqa: E402

METRIC = "MLUPS & %HBM roofline, FP32 vs FP64 at ≤1e-4 WSS err; adjoint/forward time"
PEAK_HBM_GBS = 8000.0
WO, PERIOD, U_MEAN, U_AMP = 4.0, 4000, 0.05, 0.025
WSS_EVERY = 1  # wall shear stress and its time sums every step, as Simulation::Iterate does


def womersley_basis(r, wo=None):
    """Three inlet modes: mean Poiseuille (mean 1) and the real / imaginary
    parts of the Womersley profile (Womersley number wo) normalised to unit
    mean over the inlet nodes, so u = U_mean b0 + U_amp (b1 cos wt + b2 sin wt)."""
    from scipy.special import jv

    r = np.clip(r, 0.0, 1.0)
    b0 = 2.0 * (1.0 - r * r)
    b0 /= b0.mean()
    a = (1j ** 1.5) * (WO if wo is None else wo)
    P = 1.0 - jv(0, a * r) / jv(0, a)
    P /= P.mean()
    return np.stack([b0, P.real, -P.imag])


def coefficients(t0, n):
    t = np.arange(t0, t0 + n, dtype=np.float64)
    w = 2.0 * math.pi / PERIOD
    return np.stack([np.full(n, U_MEAN), U_AMP * np.cos(w * t), U_AMP * np.sin(w * t)], 1)


def build_geometry(workdir):
    t0 = time.time()
    cfg, info = workloads.y_bifurcation(workdir)
    t1 = time.time()
    geom = lbm.Geometry.from_config(cfg, normals=2)
    t2 = time.time()
    ijk, kind, _ = geom.nodes()
    b = geom.info.bbox
    inl = np.where(kind == lbm.KIND_INLET)[0]
    x = b[0] + ijk[inl, 0] * geom.info.spacing
    y = b[2] + ijk[inl, 1] * geom.info.spacing
    r = np.hypot(x - info["parent_axis"][0], y - info["parent_axis"][1]) / info["parent_radius"]
    print(f"# geometry: {info['n_triangles']} triangles (mesh {t1 - t0:.1f}s), {geom.n} nodes, "
          f"grid {tuple(geom.info.grid)}, tables {t2 - t1:.1f}s", file=sys.stderr, flush=True)
    return geom, womersley_basis(r), info


def tau_for():
    R = 24.0
    nu = R * R * (2 * math.pi / PERIOD) / (WO * WO)
    return 3.0 * nu + 0.5


PRECISION_STEPS = 8000  # the precision study's own horizon: two Womersley periods


def bifurcation_solver(geom, basis, prec):
    s = lbm.Solver(geom, precision=prec, wall_mode=lbm.WALL_HALFWAY_BB, inlet_mode=lbm.INLET_NEE_VELOCITY,
                   outlet_mode=lbm.OUTLET_NEE_PRESSURE, tau=tau_for(), dt=1e-5, parity_checks=0,
                   wss_every=WSS_EVERY)
    s.set_inlet_profile(basis)
    s.set_kernel_timing(True)
    return s


def run_precision(geom, basis, prec, steps, warmup):
    """warmup untimed steps, then `steps` timed in one batch (HIP events on
    the solver's stream; the stress-flush launches timed apart)."""
    s = bifurcation_solver(geom, basis, prec)
    if warmup:
        s.step_coefficients(coefficients(0, warmup))
    s.synchronize()
    coef = coefficients(warmup, steps)
    t0 = time.perf_counter()
    s.step_coefficients(coef)
    s.synchronize()
    wall = time.perf_counter() - t0
    ev_ms, n = s.last_batch_ms()
    step_ms, flush_ms = s.last_batch_kernel_ms()
    assert n == steps
    return s, wall, ev_ms, step_ms, flush_ms


REF_FAMILY_STEPS = 500


def run_reference_family(geom, steps=REF_FAMILY_STEPS, warmup=20):
    """The reference-parity kernel family on the same vessel: FP64, the
    reference's own rules (Zou-He inlet at the constant speed U_mean, outlet
    copy, sequential wall copy, f* == 0 push check, WSS every step), i.e.
    lbm_params_default as a drop-in for Simulation::Iterate runs it."""
    s = lbm.Solver(geom, tau=tau_for(), dt=1e-5, cte_vel=U_MEAN)
    s.step(warmup)
    s.synchronize()
    t0 = time.perf_counter()
    s.step(steps)
    s.synchronize()
    wall = time.perf_counter() - t0
    ev_ms, _ = s.last_batch_ms()
    info = s.info()
    launch_s = ev_ms / 1e3 / steps
    achieved = geom.n * info.bytes_per_update / launch_s / 1e9
    s.close()
    return {"workload": "same vessel, FP64, reference rules (Zou-He inlet, outlet copy, wall copy rule, "
                        "f* == 0 push check, WSS every step): the lbm_kernel family",
            "steps": steps, "mlups": round(geom.n * steps / wall / 1e6, 1), "launch_us": round(launch_s * 1e6, 2),
            "special_nodes": int(info.n_special), "achieved_gbs": round(achieved, 1),
            "roofline_frac": round(achieved / PEAK_HBM_GBS, 4)}


ADJ_STEPS, ADJ_EVERY, ADJ_MODES = 4000, 200, 16
ADJ_SEGMENT_S0, ADJ_SEGMENT_LEN = 300.0, 20.0


def adjoint_case(geom, info):
    """Adjoint workload on the same vessel: 16 Zernike inlet coefficients
    (seed 3) times a pulsatile factor, Brinkman porosity 1e-3 |N(0,1)|
    (seed 4), objective = time-averaged sum of |WSS|^2 over a 20-lu segment of
    one daughter branch plus the inlet-outlet pressure drop."""
    ijk, kind, _ = geom.nodes()
    b = geom.info.bbox
    h = geom.info.spacing
    pts = np.array([b[0], b[2], b[4]])[None, :] + ijk * h
    inl = np.where(kind == lbm.KIND_INLET)[0]
    x = pts[inl, 0] - info["parent_axis"][0]
    y = pts[inl, 1] - info["parent_axis"][1]
    rho = np.clip(np.hypot(x, y) / info["parent_radius"], 0.0, 1.0)
    basis = workloads.zernike_basis(rho, np.arctan2(y, x), ADJ_MODES)
    rng = np.random.default_rng(3)
    theta = np.concatenate([[U_MEAN], 0.1 * U_MEAN * rng.standard_normal(ADJ_MODES - 1)])
    alpha = (1e-3 * np.abs(np.random.default_rng(4).standard_normal(geom.n))).astype(np.float32)
    nw, _ = geom.nearwall()
    mask = workloads.branch_segment_mask(pts[nw], info["branch_point"], info["daughter_dirs"][0], ADJ_SEGMENT_S0,
                                         ADJ_SEGMENT_LEN, info["daughter_radius"] + 6.0).astype(np.uint8)
    wave = 1.0 + 0.5 * np.sin(2.0 * math.pi * np.arange(ADJ_STEPS) / PERIOD)
    return dict(basis=basis, theta=theta, alpha=alpha, mask=mask, wave=wave)


def adjoint_solver(geom, case, prec):
    s = lbm.Solver(geom, precision=prec, wall_mode=lbm.WALL_HALFWAY_BB, inlet_mode=lbm.INLET_NEE_VELOCITY,
                   outlet_mode=lbm.OUTLET_NEE_PRESSURE, tau=tau_for(), dt=1e-5, parity_checks=0, wss_every=1)
    s.set_inlet_profile(case["basis"])
    s.set_porosity(case["alpha"])
    return s


def run_adjoint(geom, info, steps=ADJ_STEPS):
    """Times forward, adjoint forward (checkpointed) and reverse passes in FP32
    and checks the gradient against the FP64 adjoint and against central
    differences of the FP64 forward along a random inlet direction."""
    case = adjoint_case(geom, info)
    wave = case["wave"][:steps]
    coef = lambda th: wave[:, None] * th[None, :]  # noqa: E731

    def timed(fn):
        t0 = time.perf_counter()
        r = fn()
        return r, time.perf_counter() - t0

    # Every FP32 pass below runs on one solver, returned to rest between them
    # (lbm_set_populations(NULL)), so that the forward pass and the gradient
    # are timed on the same state buffers: how much of a solver's state the
    # Infinity Cache holds has differed between solvers of one process.
    a32 = adjoint_solver(geom, case, lbm.PREC_F32)
    c = coef(case["theta"])
    # Objective weights: the stress term scaled to the size of the pressure
    # term over the whole run (the segment lies ~900 lu downstream of the
    # inlet, which the flow reaches only after ~1600 steps).
    jw = a32.adjoint_forward(steps, c, wss_weight=1.0, dp_weight=0.0, wss_mask=case["mask"],
                             checkpoint_every=ADJ_EVERY)
    a32.set_populations(None)
    jp = a32.adjoint_forward(steps, c, wss_weight=0.0, dp_weight=1.0, checkpoint_every=ADJ_EVERY)
    a32.set_populations(None)
    w_wss = abs(jp / jw) if jw > 0 else 1.0
    obj = dict(wss_weight=w_wss, dp_weight=1.0, wss_mask=case["mask"], checkpoint_every=ADJ_EVERY)
    # The forward pass alone.
    _, t_fwd = timed(lambda: (a32.step_coefficients(c), a32.synchronize()))
    a32.set_populations(None)
    # The gradient as an optimisation loop sees it: the second forward/backward
    # pair (the first allocates the reverse pass's workspace, which later pairs
    # reuse).
    a32.adjoint_forward(steps, c, **obj)
    a32.adjoint_backward(0)
    a32.set_populations(None)
    J32, t_af = timed(lambda: a32.adjoint_forward(steps, c, **obj))
    (gc32, ga32), t_bw = timed(lambda: a32.adjoint_backward(0))
    a32.close()
    a64 = adjoint_solver(geom, case, lbm.PREC_F64)
    J64 = a64.adjoint_forward(steps, c, **obj)
    gc64, ga64 = a64.adjoint_backward(0)
    a64.close()
    gt32 = (wave[:, None] * gc32).sum(0)
    gt64 = (wave[:, None] * gc64).sum(0)
    v = np.random.default_rng(7).standard_normal(ADJ_MODES)
    v /= np.linalg.norm(v)
    hstep = 1e-4 * np.linalg.norm(case["theta"])
    Jpm = []
    for sgn in (1.0, -1.0):
        sol = adjoint_solver(geom, case, lbm.PREC_F64)
        Jpm.append(sol.adjoint_forward(steps, coef(case["theta"] + sgn * hstep * v), **obj))
        sol.close()
    fd = (Jpm[0] - Jpm[1]) / (2.0 * hstep)
    ad = float(gt64 @ v)
    return {
        "workload": f"adjoint on the same vessel: {steps} FP32 steps, checkpoint every {ADJ_EVERY}, "
                    f"{ADJ_MODES} Zernike inlet coefficients + per-node Brinkman porosity, "
                    f"J = mean sum |WSS|^2 over a {ADJ_SEGMENT_LEN:g}-lu daughter segment "
                    f"({int(case['mask'].sum())} near-wall nodes) + pressure drop",
        "timing": "one FP32 solver returned to rest between passes: its forward pass of 4000 steps, and its "
                  "second checkpointed forward + reverse pair (the first allocates the reverse workspace)",
        "forward_s": round(t_fwd, 4), "adjoint_forward_s": round(t_af, 4), "backward_s": round(t_bw, 4),
        "adjoint_over_forward": round(t_bw / t_fwd, 3), "gradient_over_forward": round((t_af + t_bw) / t_fwd, 3),
        "J_f32": J32, "J_f64": J64, "J_wss_unit_weight": jw, "J_pressure_drop": jp, "wss_weight": w_wss,
        "fd_directional_rel": abs(ad - fd) / abs(fd) if fd != 0 else None,
        "f32_vs_f64_inlet_grad_rel_l2": float(np.linalg.norm(gt32 - gt64) / np.linalg.norm(gt64)),
        "f32_vs_f64_porosity_grad_rel_l2": float(np.linalg.norm(ga32 - ga64) / np.linalg.norm(ga64)),
    }


STENOSIS_TAU = 0.55


def run_stenosis(steps, warmup=20):
    """D3Q27 stenosed tube (radius 64, length 1024, 50% area cosine
    stenosis at z = 384 over 128 lu, ~12.7M nodes), parabolic NEE inlet at
    Re 200 on the diameter, NEE zero-pressure outlet, bounce-back walls;
    FP32-shifted and FP64 step rates with the stress computed every step."""
    with tempfile.TemporaryDirectory() as wd:
        cfg, info = workloads.stenosed_tube(wd, tau=STENOSIS_TAU)
        geom = lbm.Geometry.from_config(cfg, q=27, normals=2)
    ijk, kind, _ = geom.nodes()
    b = geom.info.bbox
    h = geom.info.spacing
    inl = np.where(kind == lbm.KIND_INLET)[0]
    x = b[0] + ijk[inl, 0] * h - info["axis"][0]
    y = b[2] + ijk[inl, 1] * h - info["axis"][1]
    prof = np.clip(1.0 - (x * x + y * y) / info["radius"] ** 2, 0.0, None)
    prof /= prof.mean()
    nu = (STENOSIS_TAU - 0.5) / 3.0
    u_mean = 200.0 * nu / (2.0 * info["radius"])
    out = {"workload": f"D3Q27 stenosed tube, {geom.n} nodes, parabolic inlet Re 200 (U_mean {u_mean:.4f}), "
                       f"tau {STENOSIS_TAU}, bounce-back walls, NEE inlet/outlet", "nodes": geom.n, "steps": steps}
    for name, prec, nbytes in (("f32", lbm.PREC_F32, 2 * 27 * 4), ("f64", lbm.PREC_F64, 2 * 27 * 8)):
        s = lbm.Solver(geom, precision=prec, wall_mode=lbm.WALL_HALFWAY_BB, inlet_mode=lbm.INLET_NEE_VELOCITY,
                       outlet_mode=lbm.OUTLET_NEE_PRESSURE, tau=STENOSIS_TAU, dt=1e-5, parity_checks=0, wss_every=1)
        s.set_inlet_profile(prof[None, :])
        s.step_coefficients(np.full((warmup, 1), u_mean))
        s.synchronize()
        t0 = time.perf_counter()
        s.step_coefficients(np.full((steps, 1), u_mean))
        s.synchronize()
        wall = time.perf_counter() - t0
        ev_ms, _ = s.last_batch_ms()
        mlups = geom.n * steps / wall / 1e6
        out[f"{name}_mlups"] = round(mlups, 1)
        out[f"{name}_roofline_frac"] = round(geom.n * nbytes / (ev_ms / 1e3 / steps) / 1e9 / PEAK_HBM_GBS, 4)
        s.close()
    return out


LARGE_SCALE, LARGE_WO, LARGE_PERIOD, LARGE_WSS_EVERY = 4.3, 5.0, 8000, 100


def large_share_geometry(host_chain=False):
    """The large share's geometry, built the reference's way from a surface:
    the tortuous vessel's STL (workloads.tortuous_vessel_stl, marching
    tetrahedra at 8 lu) through the set-up chain with ray labels,
    classification, order, push table and near-wall normals on the GPU
    (host_chain=True: labels and tables on 16 host threads instead). Returns
    (geometry, vessel info, set-up seconds)."""
    import tempfile

    t0 = time.perf_counter()
    with tempfile.TemporaryDirectory() as d:
        cfg, info = workloads.tortuous_vessel_stl(d, scale=LARGE_SCALE, mesh_h=8.0)
        t1 = time.perf_counter()
        print(f"# large share: STL of {info['n_triangles']} triangles in {t1 - t0:.1f}s", file=sys.stderr, flush=True)
        on_gpu = not host_chain
        geom = lbm.Geometry.from_config(cfg, normals=2, label_device=on_gpu, classify_device=on_gpu,
                                        nthreads=int(os.environ.get("OMP_NUM_THREADS", "16")))
    t2 = time.perf_counter()
    print(f"# large share: {geom.n} nodes, geometry {t2 - t1:.1f}s", file=sys.stderr, flush=True)
    return geom, info, {"stl": round(t1 - t0, 1), "geometry": round(t2 - t1, 1)}


def run_large_share(steps, warmup=20):
    """Large single-GPU share: the tortuous vessel (Catmull-Rom centreline
    through a seeded 12-point random walk, radius 20-40 and a saccular bulge
    of radius 60, all lengths x4.3: ~300M nodes) built from its STL through
    the GPU set-up chain, D3Q19 FP32-shifted, Womersley inlet (Wo 5, period
    8000 steps, U_mean 0.05), NEE zero-pressure outlet, bounce-back walls,
    wall shear stress every 100 steps. Returns the step rate and its roofline
    fraction, with the set-up times."""
    geom, info, setup = large_share_geometry()
    t2 = time.perf_counter()
    grid = tuple(int(v) for v in geom.info.grid)
    ijk, kind, _ = geom.nodes()
    ii = np.where(kind == lbm.KIND_INLET)[0]
    b, h = geom.info.bbox, geom.info.spacing
    c0, r0 = info["centreline"][0], float(info["radius"][0])
    r = np.hypot(b[0] + ijk[ii, 0] * h - c0[0], b[2] + ijk[ii, 1] * h - c0[1]) / r0
    del ijk, kind
    basis = womersley_basis(r, LARGE_WO)
    w = 2.0 * math.pi / LARGE_PERIOD
    tau = 3.0 * r0 * r0 * w / (LARGE_WO * LARGE_WO) + 0.5

    def coef(t0_, n):
        t = np.arange(t0_, t0_ + n, dtype=np.float64)
        return np.stack([np.full(n, U_MEAN), U_AMP * np.cos(w * t), U_AMP * np.sin(w * t)], 1)

    s = lbm.Solver(geom, precision=lbm.PREC_F32, wall_mode=lbm.WALL_HALFWAY_BB, inlet_mode=lbm.INLET_NEE_VELOCITY,
                   outlet_mode=lbm.OUTLET_NEE_PRESSURE, tau=tau, dt=1e-5, parity_checks=0, wss_every=LARGE_WSS_EVERY)
    s.set_inlet_profile(basis)
    t3 = time.perf_counter()
    print(f"# large share: solver {t3 - t2:.1f}s", file=sys.stderr, flush=True)
    s.step_coefficients(coef(0, warmup))
    s.synchronize()
    ta = time.perf_counter()
    s.step_coefficients(coef(warmup, steps))
    s.synchronize()
    wall = time.perf_counter() - ta
    ev_ms, _ = s.last_batch_ms()
    n = geom.n
    achieved = n * 152.0 / (ev_ms / 1e3 / steps) / 1e9
    res = {"workload": f"tortuous vessel x{LARGE_SCALE} (12-point walk, radius {20 * LARGE_SCALE:.0f}-{40 * LARGE_SCALE:.0f}, "
                       f"bulge {60 * LARGE_SCALE:.0f}) from its STL through GPU labels, classification and normals, "
                       f"D3Q19 FP32-shifted, Womersley Wo {LARGE_WO} period {LARGE_PERIOD}, "
                       f"WSS every {LARGE_WSS_EVERY} steps",
           "nodes": n, "grid": grid, "triangles": info["n_triangles"], "steps": steps, "tau": round(tau, 4),
           "f32_mlups": round(n * steps / wall / 1e6, 1), "roofline_frac": round(achieved / PEAK_HBM_GBS, 4),
           "achieved_gbs": round(achieved, 1), "launch_us": round(ev_ms * 1e3 / steps, 2),
           "setup_s": dict(setup, solver=round(t3 - t2, 1))}
    s.close()
    return res


def large_share_child(steps, timeout_s=900):
    """The large share in a child process (tens of GB of host tables and a
    45 GB state): a failure there is reported, not fatal to the bench line."""
    import subprocess

    cmd = [sys.executable, os.path.abspath(__file__), "--large-share-only", "--large-share-steps", str(steps)]
    try:
        r = subprocess.run(cmd, stdout=subprocess.PIPE, timeout=timeout_s, text=True)
    except subprocess.TimeoutExpired:
        return {"error": f"large share did not finish within {timeout_s} s"}
    if r.returncode != 0 or not r.stdout.strip():
        return {"error": f"large share exited with status {r.returncode}"}
    res = json.loads(r.stdout.strip().splitlines()[-1])
    print(f"# large share: {res['nodes']} nodes, {res['f32_mlups']:.0f} MLUPS, roofline {res['roofline_frac']:.3f}",
          file=sys.stderr, flush=True)
    return res


def host_description():
    """Core counts and CPU model of the host this runs on."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count()
    return {"cpu_model": model, "host_cpus": os.cpu_count(), "cpus_available": affinity,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def cpu_threads():
    """The host's CPU share for this process: the cores it may run on, at most
    OMP_NUM_THREADS where that is set (the share of one GPU on a shared
    node)."""
    host = host_description()
    n = host["cpus_available"] or 1
    if host["omp_num_threads"]:
        n = min(n, int(host["omp_num_threads"]))
    return max(1, n), host


def cpu_baseline(geom, budget_s=12.0):
    """The oracle (FP64 restatement of the reference step with the
    reference's own boundary rules: Zou-He inlet, outlet copy, wall copy,
    wall shear stress every step, Simulation.cpp:553-568) on the host's
    cores, on a bounded sample of the same workload, plus a one-thread row
    (one reference MPI rank runs single-threaded)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    ncores, host = cpu_threads()
    ref = oracle.Reference.from_geometry(geom, tau=tau_for(), dt=1e-5, wall_mode=0, inlet_mode=0, outlet_mode=0)
    t0 = time.perf_counter()
    ref.step(np.full(2, U_MEAN), nthreads=ncores)
    per = (time.perf_counter() - t0) / 2
    n = int(max(3, min(400, budget_s / max(per, 1e-6))))
    t0 = time.perf_counter()
    ref.step(np.full(n, U_MEAN), nthreads=ncores)
    dt = time.perf_counter() - t0
    t0 = time.perf_counter()
    ref.step(np.full(1, U_MEAN), nthreads=1)
    per1 = time.perf_counter() - t0
    n1 = int(max(1, min(40, 0.25 * budget_s / max(per1, 1e-6))))
    t0 = time.perf_counter()
    ref.step(np.full(n1, U_MEAN), nthreads=1)
    dt1 = time.perf_counter() - t0
    return {"value": round(geom.n * n / dt / 1e6, 3), "unit": "MLUPS", "cores": ncores, "kind": "port",
            "sample": f"{n} steps ({dt:.1f} s) of the same {geom.n}-node bifurcation: the oracle's FP64 restatement "
                      f"of the reference step with the reference rules (Zou-He inlet at U {U_MEAN}, outlet copy, "
                      f"wall copy, WSS every step), OpenMP on {ncores} host threads",
            "single_thread": {"value": round(geom.n * n1 / dt1 / 1e6, 3), "steps": n1}, "host": host}


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "columns_step_bifurcation_f32_traffic.md")
LARGE_PROFILE = os.path.join(ROOT, "profiles", "columns_step_tortuous_vessel_large_share.md")


def profiled_traffic(path=TRAFFIC_PROFILE):
    """HBM bytes per step launch measured by the two rocprofv3 counter passes
    (FETCH_SIZE, WRITE_SIZE, copy-kernel calibrated) recorded in a profile
    summary; None when the summary is absent."""
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def rel_l2(a, b):
    d = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / d) if d > 0 else float("nan")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precisions", default="f32,f64,f16")
    ap.add_argument("--precision-steps", type=int, default=PRECISION_STEPS,
                    help="horizon of the FP32/FP16-vs-FP64 comparison (the workload's 8000 steps)")
    ap.add_argument("--no-reference-family", action="store_true")
    ap.add_argument("--reference-family-steps", type=int, default=REF_FAMILY_STEPS)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-stenosis", action="store_true")
    ap.add_argument("--stenosis-steps", type=int, default=500)
    ap.add_argument("--adjoint-steps", type=int, default=ADJ_STEPS)
    ap.add_argument("--no-large-share", action="store_true", help="skip the ~300M-node large share")
    ap.add_argument("--large-share-steps", type=int, default=300)
    ap.add_argument("--large-share-only", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.large_share_only:  # child process of the default run
        print(json.dumps(run_large_share(args.large_share_steps)))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the whole lattice runs on a single MI355X "
              "(no sharding or collectives); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    with tempfile.TemporaryDirectory() as wd:
        geom, basis, info = build_geometry(wd)
    N = geom.n
    precs = {"f32": lbm.PREC_F32, "f64": lbm.PREC_F64, "f16": lbm.PREC_F16}
    names = args.precisions.split(",")
    if "f32" not in names:
        names = ["f32"] + names

    # Precision study over the workload's own horizon (8000 steps from rest),
    # whatever --steps says. Each run is one batch with HIP-event timing, so
    # the FP32 run is also the sustained measurement of the step kernel. The
    # fields compared against FP64 are read right after each run.
    study = {}
    for name in names:
        s, wall, ev_ms, step_ms, flush_ms = run_precision(geom, basis, precs[name], args.precision_steps, 0)
        w, _, m = s.wss()
        _, u = s.macroscopic()
        study[name] = dict(solver=s, wall=wall, step_ms=step_ms, flush_ms=flush_ms, wss=w, tawss=m, vel=u,
                           mlups=N * args.precision_steps / wall / 1e6)
        print(f"# {name}: {args.precision_steps} steps, {study[name]['mlups']:.0f} MLUPS, step kernel "
              f"{step_ms * 1e3 / args.precision_steps:.2f} us", file=sys.stderr, flush=True)
        if name == "f32":
            # The timed region of the bench contract (W warm-up steps, K timed
            # steps) on the same solver, returned to rest: a second FP32
            # solver made later in the process has run its cached-input steps
            # at 78-87 us instead of 62-71 us (the Infinity Cache holding less
            # of its state), the first one has not.
            s.set_populations(None)
            if args.warmup:
                s.step_coefficients(coefficients(0, args.warmup))
            s.synchronize()
            t0 = time.perf_counter()
            s.step_coefficients(coefficients(args.warmup, args.steps))
            s.synchronize()
            timed_wall = time.perf_counter() - t0
            timed_step_ms, _ = s.last_batch_kernel_ms()
            sinfo = s.info()
            value = N * args.steps / timed_wall / 1e6
            print(f"# timed region: {args.steps} FP32 steps, {value:.0f} MLUPS", file=sys.stderr, flush=True)

    copy_gbps = lbm.copy_probe(1 << 30, iters=20)
    print(f"# copy probe: {copy_gbps:.0f} GB/s (1 GiB read + write)", file=sys.stderr, flush=True)
    bytes_per_launch = N * sinfo.bytes_per_update
    # The dominant kernel's own average launch time: the batch's HIP-event
    # time less the event-timed stress-flush launches (one per 8 steps).
    sus = study["f32"]
    per_launch_s = sus["step_ms"] / 1e3 / args.precision_steps
    burst_launch_s = timed_step_ms / 1e3 / args.steps
    achieved = bytes_per_launch / per_launch_s / 1e9
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "MLUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(timed_wall * 1e3 / args.steps, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "precision-study Y-bifurcation (parent radius 24, daughters 18 at 35 degrees), "
                                "2.85M nodes, D3Q19, FP32-shifted populations",
                   "nodes": N, "explicit_row_nodes": int(sinfo.n_special), "triangles": info["n_triangles"],
                   "inlet": f"Womersley Wo={WO} period {PERIOD} steps, U_mean {U_MEAN}",
                   "outlet": "NEE pressure rho=1", "walls": "half-way bounce-back", "tau": round(tau_for(), 6)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": profiled_traffic(),
                     "bytes_per_launch_algorithmic": bytes_per_launch,
                     "launch_us": round(per_launch_s * 1e6, 3), "sustained_steps": args.precision_steps,
                     "kernel": "columns_kernel_occ (FP32 step; the two cache-hint variants alternate)",
                     "flush_us_per_step": round(sus["flush_ms"] * 1e3 / args.precision_steps, 3),
                     "burst": {"steps": args.steps, "launch_us": round(burst_launch_s * 1e6, 3),
                               "frac": round(bytes_per_launch / burst_launch_s / 1e9 / PEAK_HBM_GBS, 4)},
                     "copy_probe_gbps": round(copy_gbps, 1)},
        "precision_steps": args.precision_steps,
        "f32_sustained_mlups": round(sus["mlups"], 1),
    }
    if "f64" in study:
        out["fp64_mlups"] = round(study["f64"]["mlups"], 1)
        out["fp32_over_fp64"] = round(sus["mlups"] / study["f64"]["mlups"], 3)
        w64, m64, u64 = study["f64"]["wss"], study["f64"]["tawss"], study["f64"]["vel"]
        for name in ("f32", "f16"):
            if name not in study:
                continue
            w, m, u = study[name]["wss"], study[name]["tawss"], study[name]["vel"]
            out[f"{name}_wss_rel_l2"] = rel_l2(w, w64)
            out[f"{name}_tawss_rel_l2"] = rel_l2(m, m64)
            out[f"{name}_velocity_rel_l2"] = rel_l2(u, u64)
    if "f16" in study:
        out["fp16_mlups"] = round(study["f16"]["mlups"], 1)
        f16_launch = study["f16"]["step_ms"] / 1e3 / args.precision_steps
        out["fp16_roofline_frac"] = round(N * 76.0 / f16_launch / 1e9 / PEAK_HBM_GBS, 4)
    if "f64" in study:
        f64_launch = study["f64"]["step_ms"] / 1e3 / args.precision_steps
        out["fp64_roofline_frac"] = round(N * 304.0 / f64_launch / 1e9 / PEAK_HBM_GBS, 4)
    for v in study.values():
        v["solver"].close()
    if not args.no_reference_family:
        out["reference_family"] = run_reference_family(geom, args.reference_family_steps)
        print(f"# reference family (FP64, reference rules): {out['reference_family']['mlups']:.0f} MLUPS, "
              f"roofline {out['reference_family']['roofline_frac']:.3f}", file=sys.stderr, flush=True)
    out["adjoint_over_forward"] = None
    if not args.no_adjoint:
        adj = run_adjoint(geom, info, args.adjoint_steps)
        print(f"# adjoint: forward {adj['forward_s']:.3f}s, checkpointed forward {adj['adjoint_forward_s']:.3f}s, "
              f"reverse {adj['backward_s']:.3f}s", file=sys.stderr, flush=True)
        out["adjoint_over_forward"] = adj["adjoint_over_forward"]
        out["adjoint"] = adj
    if not args.no_stenosis:
        out["d3q27_stenosis"] = run_stenosis(args.stenosis_steps)
        print(f"# D3Q27 stenosis: f32 {out['d3q27_stenosis']['f32_mlups']:.0f} MLUPS, "
              f"f64 {out['d3q27_stenosis']['f64_mlups']:.0f} MLUPS", file=sys.stderr, flush=True)
    if not args.no_large_share:
        ls = large_share_child(args.large_share_steps)
        if "nodes" in ls:
            # The state (45 GB) is far beyond the Infinity Cache: this is the
            # DRAM-bound figure of the same FP32 step, with its counted traffic.
            ls["traffic"] = profiled_traffic(LARGE_PROFILE)
            ls["bytes_per_launch_algorithmic"] = ls["nodes"] * 152
        out["large_share"] = ls
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(geom)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
