"""lbm_run: the reference driver (LBM.cpp main) over the C ABI, through the
C++ drop-in classes (Simulation / SubDomain / Geometry / LatticeModel)."""
import os
import subprocess

import numpy as np
import pytest

import oracle
import workloads
from conftest import ROOT

EXE = os.path.join(ROOT, "reza-seti_amd", "lbm_run")


def test_driver_reports_a_missing_config():
    if not os.path.exists(EXE):
        pytest.skip("lbm_run not built")
    r = subprocess.run([EXE, "/nonexistent/config.txt"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Unable to open input file" in r.stderr


@pytest.mark.gpu
def test_driver_runs_the_reference_loop_bit_exact(tmp_path):
    import lbm

    cfg = workloads.straight_tube(str(tmp_path), radius=6.0, length=30.0, cte_vel=0.03, max_iter=150)
    dump = tmp_path / "f.bin"
    out = tmp_path / "out"
    out.mkdir()
    text = open(cfg).read().replace("IterToWrite: 100000", "IterToWrite: 100")
    open(cfg, "w").write(text)
    r = subprocess.run([EXE, cfg, "--dump", str(dump), "--out", str(out)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "Iterations: 150" in r.stdout
    geom = lbm.Geometry.from_config(cfg)
    f = np.fromfile(dump, dtype=np.float64).reshape(geom.n, 19)
    ref = oracle.Reference.from_geometry(geom, tau=0.8, dt=1e-5)
    ref.step(np.full(150, 0.03))
    assert np.array_equal(f, ref.f())
    # Simulation::WriteOutput's three files (Simulation.cpp:835-856).
    raw = (out / "100_FlowHemodynamics.csv").read_text().splitlines()
    assert raw[0] == "X,Y,Z,WSS,VelMag" and len(raw) == geom.info.n_nearwall + 1
    smooth = (out / "FlowHemodynamics" / "Iteration_100_100_FlowHemodynamics.csv").read_text().splitlines()
    assert smooth == raw  # pitch 1 > the 0.25 smoothing radius: no neighbours, values kept
    # Iteration_100 follows the 101st iteration: its fields are those that
    # iteration's collision set, as the oracle keeps them after 101 steps.
    from test_output import read_vtu

    ref101 = oracle.Reference.from_geometry(geom, tau=0.8, dt=1e-5)
    ref101.step(np.full(101, 0.03))
    rho, u = ref101.macro()
    ijk, _, _ = geom.nodes()
    order = np.lexsort((ijk[:, 0], ijk[:, 1], ijk[:, 2]))
    npts, _, a = read_vtu(out / "Fluid_Variables" / "100.vtu")
    assert npts == geom.n
    assert np.allclose(a["Density"], rho[order], rtol=1e-14, atol=0)
    assert np.allclose(a["Velocity"], u[order], rtol=0, atol=1e-14)
    vm = np.sqrt((u ** 2).sum(1))
    nodes, _ = geom.nearwall()
    col = np.loadtxt(out / "100_FlowHemodynamics.csv", delimiter=",", skiprows=1, ndmin=2)[:, 4]
    assert np.allclose(col, vm[nodes], rtol=1e-5, atol=1e-15)


def _pulsatile_schedule(dt, systole, diastole):
    """Simulation::Iterate's Pulsatile branch (Simulation.cpp:549-602) restated
    over the loop variables alone: the iterations, the inlet speed of each
    and the (phase, iteration) of every output."""
    t, it, speeds, outs, ns, nd = 0.0, 0, [], [], 0, 0
    while t <= 2.7:
        speeds.append(oracle.inlet_velocity(t, 0.9))
        tic = np.fmod(t, 0.9)
        if abs(tic - systole) < dt / 2:
            ns += 1
            outs.append((f"Systole_{ns}", it))
        elif abs(tic - diastole) < dt / 2:
            nd += 1
            outs.append((f"Diastole_{nd}", it))
        t += dt
        it += 1
    return it, np.array(speeds), outs


@pytest.mark.gpu
def test_driver_runs_the_pulsatile_loop(tmp_path):
    import lbm

    dt = 0.01
    cfg = workloads.axis_tube(str(tmp_path), axis="z", radius=5.0, length=16.0, tau=1.0, dt=dt,
                              flow_type="Pulsatile")
    n, speeds, outs = _pulsatile_schedule(dt, 0.15, 0.9)
    assert any(p.startswith("Systole") for p, _ in outs) and any(p.startswith("Diastole") for p, _ in outs)
    dump, out = tmp_path / "f.bin", tmp_path / "out"
    out.mkdir()
    r = subprocess.run([EXE, cfg, "--dump", str(dump), "--out", str(out)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    assert f"Iterations: {n}," in r.stdout  # the t <= 2.7 stop
    got = sorted(os.listdir(out / "FlowHemodynamics"))
    assert got == sorted(f"{p}_{it}_FlowHemodynamics.csv" for p, it in outs)
    assert sorted(os.listdir(out / "Fluid_Variables")) == sorted(f"{it}.vtu" for _, it in outs)
    geom = lbm.Geometry.from_config(cfg)
    ref = oracle.Reference.from_geometry(geom, tau=1.0, dt=dt)
    ref.step(speeds, nthreads=8)
    f = np.fromfile(dump, dtype=np.float64).reshape(geom.n, 19)
    assert np.array_equal(f, ref.f())
    # the last output's OSI and TAWSS columns from the oracle's accumulators
    # at that iteration (Simulation.cpp:813-824)
    phase, it = outs[-1]
    ref2 = oracle.Reference.from_geometry(geom, tau=1.0, dt=dt)
    ref2.step(speeds[: it + 1], nthreads=8)
    _, acc, mag = ref2.wss()
    osi = 0.5 * (1.0 - np.sqrt((acc ** 2).sum(1)) / mag)
    tawss = mag / 0.9 * dt
    raw = np.loadtxt(out / f"{it}_FlowHemodynamics.csv", delimiter=",", skiprows=1, ndmin=2)
    head = (out / f"{it}_FlowHemodynamics.csv").read_text().splitlines()[0]
    assert head == "X,Y,Z,WSS,OSI,TAWSS,VelMag"
    np.testing.assert_allclose(raw[:, 4], osi, rtol=1e-5, atol=1e-12)
    np.testing.assert_allclose(raw[:, 5], tawss, rtol=1e-5, atol=0)


@pytest.mark.gpu
def test_driver_runs_the_reference_case(tmp_path):
    # The reference's own case: its config.txt and JT.stl at grid_size 0.025
    # (6.6M nodes, one inlet and four outlet blocks, Constant CteVel 0.1,
    # tau 0.68), through lbm_run, bit for bit against the oracle.
    import lbm
    import sys

    from conftest import GOLDEN

    sys.path.insert(0, GOLDEN)
    import cases

    steps = 300
    cfg = cases.jt_case(str(tmp_path), 0.025)
    dump = tmp_path / "f.bin"
    print("lbm_run on the reference case", flush=True)
    r = subprocess.run([EXE, cfg, "--steps", str(steps), "--dump", str(dump)], capture_output=True, text=True,
                       timeout=200)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    assert f"Iterations: {steps}," in r.stdout
    geom = lbm.Geometry.from_config(cfg)
    assert geom.n > 6_000_000 and geom.info.n_triangles == 9712
    print(f"oracle: {geom.n} nodes, {steps} steps", flush=True)
    ref = oracle.Reference.from_geometry(geom, tau=0.68, dt=1e-5)
    ref.step(np.full(steps, 0.1), nthreads=min(16, os.cpu_count() or 1))
    f = np.fromfile(dump, dtype=np.float64).reshape(geom.n, 19)
    assert np.isfinite(f).all()
    assert np.array_equal(f, ref.f())
