"""D3Q27 on the GPU: the stenosed-tube workload at a small size.

The reference builds D3Q27 tables (LatticeModel.cpp:249-375) and collides
with them (Simulation.cpp:149-207), but its Zou-He inlet is D3Q19 only, so
the D3Q27 checks use the reference wall/outlet copy rules without an inlet
rule (bit-exact against the oracle), the half-way bounce-back family with
NEE inlet/outlet (FP64 against the oracle to rounding), FP32 against FP64,
and the periodic Poiseuille closed form.
"""
import numpy as np
import pytest

import lbm
import oracle
import workloads

pytestmark = pytest.mark.gpu

TAU, DT, U = 0.7, 1e-5, 0.03


@pytest.fixture(scope="module")
def sten(tmp_path_factory):
    wd = str(tmp_path_factory.mktemp("sten"))
    cfg, _ = workloads.stenosed_tube(wd, radius=8.0, length=40.0, z_center=20.0, sten_length=16.0, segments=64,
                                   tau=TAU)
    return lbm.Geometry.from_config(cfg, q=27)


def test_reference_rules_d3q27_bit_exact(sten):
    sol = lbm.Solver(sten, tau=TAU, dt=DT, cte_vel=U, inlet_mode=lbm.INLET_NONE)
    ref = oracle.Reference.from_geometry(sten, tau=TAU, dt=DT, inlet_mode=2)
    rng = np.random.default_rng(27)
    _, _, w = lbm.lattice(27)
    f0 = w[None, :] * (1.0 + 0.05 * rng.uniform(-1, 1, size=(sten.n, 27)))
    sol.set_populations(f0)
    ref.set_f(f0)
    done = 0
    for target in (1, 20, 100):
        sol.step(target - done)
        ref.step(np.full(target - done, U), nthreads=8)
        done = target
        d = np.max(np.abs(sol.populations() - ref.f()))
        print(f"step {done}: max |f - f_oracle| = {d:.3e}", flush=True)
        assert d == 0.0
    w, acc, mag = sol.wss()
    rw, racc, rmag = ref.wss()
    assert np.array_equal(w, rw) and np.array_equal(acc, racc) and np.array_equal(mag, rmag)


def test_bounce_back_nee_d3q27_matches_oracle(sten):
    sol = lbm.Solver(sten, tau=TAU, dt=DT, cte_vel=U, wall_mode=lbm.WALL_HALFWAY_BB,
                     inlet_mode=lbm.INLET_NEE_VELOCITY, outlet_mode=lbm.OUTLET_NEE_PRESSURE, parity_checks=0)
    ref = oracle.Reference.from_geometry(sten, tau=TAU, dt=DT, wall_mode=1, inlet_mode=1, outlet_mode=1)
    done = 0
    for target in (1, 40, 200):
        sol.step(target - done)
        ref.step(np.full(target - done, U), nthreads=8)
        done = target
        d = np.max(np.abs(sol.populations() - ref.f()))
        print(f"step {done}: max |f - f_oracle| = {d:.3e}", flush=True)
        assert d < 1e-12
    w, _, mag = sol.wss()
    rw, _, rmag = ref.wss()
    assert np.max(np.abs(w - rw)) <= 1e-9 * np.max(np.abs(rw))
    assert np.max(np.abs(mag - rmag)) <= 1e-9 * np.max(np.abs(rmag))


def test_fp32_d3q27_tracks_fp64(sten):
    out = {}
    for prec in (lbm.PREC_F64, lbm.PREC_F32):
        sol = lbm.Solver(sten, precision=prec, tau=TAU, dt=DT, cte_vel=U, wall_mode=lbm.WALL_HALFWAY_BB,
                         inlet_mode=lbm.INLET_NEE_VELOCITY, outlet_mode=lbm.OUTLET_NEE_PRESSURE, parity_checks=0)
        sol.step(600)
        _, u = sol.macroscopic()
        w, _, mag = sol.wss()
        out[prec] = (u, w, mag)
    u64, w64, m64 = out[lbm.PREC_F64]
    u32, w32, m32 = out[lbm.PREC_F32]
    ru = np.linalg.norm(u32 - u64) / np.linalg.norm(u64)
    rw = np.linalg.norm(w32 - w64) / np.linalg.norm(w64)
    rm = np.linalg.norm(m32 - m64) / np.linalg.norm(m64)
    print(f"FP32 vs FP64 (D3Q27): velocity {ru:.2e}, WSS {rw:.2e}, TAWSS {rm:.2e}")
    assert ru < 1e-4 and rw < 1e-4 and rm < 1e-4


def test_periodic_poiseuille_d3q27():
    R, L, g, tau = 8.0, 12, 2e-6, 0.8
    lab, origin, (ax, ay) = workloads.tube_mask(R, L)
    geom = lbm.Geometry.from_labels(lab, origin, 1.0, periodic=(0, 0, 1), normals=False, q=27)
    sol = lbm.Solver(geom, tau=tau, dt=DT, wall_mode=lbm.WALL_HALFWAY_BB, inlet_mode=lbm.INLET_NONE,
                     outlet_mode=lbm.OUTLET_NONE, parity_checks=0, force=(0.0, 0.0, g), wss_every=0)
    s0 = sol.reduce()
    sol.step(6000)
    ijk, _, _ = geom.nodes()
    rho, u = sol.macroscopic()
    uz = u[:, 2] + 0.5 * g / rho
    r2 = (ijk[:, 0] - ax) ** 2 + (ijk[:, 1] - ay) ** 2
    nu = (tau - 0.5) / 3.0
    core = r2 < (R - 2.5) ** 2
    A = np.stack([np.ones_like(r2), r2], 1)[core]
    coef, *_ = np.linalg.lstsq(A, uz[core], rcond=None)
    assert coef[1] == pytest.approx(-g / (4 * nu), rel=0.02)
    assert abs(sol.reduce().mass - s0.mass) < 1e-10 * s0.mass


def test_d3q27_output_files(sten, tmp_path):
    # The VTU carries the last iteration's collision as the oracle keeps it;
    # the CSV's VelMag is 0, since only the D3Q19 collision sets it in the
    # reference (Simulation.cpp:136-139).
    from test_output import read_vtu

    sol = lbm.Solver(sten, tau=TAU, dt=DT, cte_vel=U, inlet_mode=lbm.INLET_NONE)
    ref = oracle.Reference.from_geometry(sten, tau=TAU, dt=DT, inlet_mode=2)
    sol.step(25)
    ref.step(np.full(25, U), nthreads=8)
    sol.write_output(sten, 25, "Iteration_25", tmp_path, pulsatile=False)
    rho, u = ref.macro()
    ijk, _, _ = sten.nodes()
    order = np.lexsort((ijk[:, 0], ijk[:, 1], ijk[:, 2]))
    npts, _, a = read_vtu(tmp_path / "Fluid_Variables" / "25.vtu")
    assert npts == sten.n
    assert np.allclose(a["Density"], rho[order], rtol=1e-14, atol=0)
    assert np.allclose(a["Velocity"], u[order], rtol=0, atol=1e-14)
    raw = np.loadtxt(tmp_path / "25_FlowHemodynamics.csv", delimiter=",", skiprows=1, ndmin=2)
    assert raw.shape == (sten.info.n_nearwall, 5) and np.all(raw[:, 4] == 0.0)
