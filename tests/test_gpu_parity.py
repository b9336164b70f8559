"""FP64 parity of the fused HIP kernel against the oracle (reference rules).

The oracle restates Simulation::Iterate (Simulation.cpp:553-568) in C; both
run the same tables from the same initial state. The bar is bit equality of
populations, post-collision populations and WSS accumulators (the
reference's own tolerance target is 1e-12 relative).
"""
import numpy as np
import pytest

import lbm
import oracle
import workloads

pytestmark = pytest.mark.gpu

TAU, DT, U = 0.8, 1e-5, 0.05


def _pair(geom, **kw):
    sol = lbm.Solver(geom, tau=TAU, dt=DT, cte_vel=U, **kw)
    ref = oracle.Reference.from_geometry(geom, tau=TAU, dt=DT)
    return sol, ref


def _assert_same(a, b, what):
    d = np.max(np.abs(a - b)) if a.size else 0.0
    assert d == 0.0, f"{what}: max |diff| = {d:.3e}"


def test_tube_2000_steps_bit_exact(tube_geometry):
    sol, ref = _pair(tube_geometry)
    done = 0
    for target in (1, 10, 100, 2000):
        sol.step(target - done)
        ref.step(np.full(target - done, U), nthreads=8)
        done = target
        print(f"step {done}", flush=True)
        _assert_same(sol.populations(), ref.f(), f"f after {done} steps")
        _assert_same(sol.post_collision(), _collided(ref), f"f* after {done} steps")
        w, acc, mag = sol.wss()
        rw, racc, rmag = ref.wss()
        _assert_same(w, rw, "wss")
        _assert_same(acc, racc, "accumulated tau")
        _assert_same(mag, rmag, "accumulated |tau|")


def _collided(ref):
    """f* of the collision that opens the next iteration. The oracle's next
    step recomputes its collision from f, so running it here is harmless."""
    ref.collide()
    return ref.fs()


def test_random_populations_bit_exact(tube_geometry):
    rng = np.random.default_rng(12345)
    _, _, w = lbm.lattice(19)
    f0 = w[None, :] * (1.0 + 0.05 * rng.uniform(-1, 1, size=(tube_geometry.n, 19)))
    sol, ref = _pair(tube_geometry)
    sol.set_populations(f0)
    ref.set_f(f0)
    _assert_same(sol.populations(), f0, "set/get round trip")
    sol.step(5)
    ref.step(np.full(5, U))
    _assert_same(sol.populations(), ref.f(), "f after 5 steps from random f")


@pytest.mark.parametrize("case", ["periodic-tube", "periodic-box"])
def test_periodic_wraps_bit_exact(case):
    # Periodic wraps give plain fluid slots whose neighbours are not at their
    # column segment's offsets (the step reads their pull-table rows) among
    # slots whose sources come from the segments: a tube with wall nodes and
    # a wrap along z, and a box that wraps along every axis (no wall at all).
    if case == "periodic-tube":
        lab, origin, _ = workloads.tube_mask(5.0, 8)
        periodic = (0, 0, 1)
    else:
        lab, origin, periodic = np.ones((7, 6, 5), np.uint8), (0.0, 0.0, 0.0), (1, 1, 1)
    geom = lbm.Geometry.from_labels(lab, origin, 1.0, periodic=periodic, normals=False)
    sol = lbm.Solver(geom, tau=TAU, dt=DT, inlet_mode=lbm.INLET_NONE, outlet_mode=lbm.OUTLET_NONE)
    ref = oracle.Reference.from_geometry(geom, tau=TAU, dt=DT, inlet_mode=2, outlet_mode=2, periodic=periodic)
    rng = np.random.default_rng(77)
    _, _, w = lbm.lattice(19)
    f0 = w[None, :] * (1.0 + 0.05 * rng.uniform(-1, 1, size=(geom.n, 19)))
    sol.set_populations(f0)
    ref.set_f(f0)
    print(f"{case}: {geom.n} nodes", flush=True)
    for target in (1, 20):
        sol.step(target if target == 1 else target - 1)
        ref.step(np.zeros(target if target == 1 else target - 1))
        _assert_same(sol.populations(), ref.f(), f"{case}: f after {target} steps")
        _assert_same(sol.post_collision(), _collided(ref), f"{case}: f* after {target} steps")


def test_pulsatile_inlet_bit_exact(tmp_path):
    cfg = workloads.straight_tube(str(tmp_path), radius=6.0, length=30.0)
    geom = lbm.Geometry.from_config(cfg)
    sol = lbm.Solver(geom, tau=TAU, dt=DT, flow_type=lbm.FLOW_PULSATILE)
    ref = oracle.Reference.from_geometry(geom, tau=TAU, dt=DT)
    t, u = 0.0, []
    for _ in range(50):
        u.append(oracle.inlet_velocity(t, 0.9))
        t += DT
    sol.step(50)
    ref.step(np.array(u))
    _assert_same(sol.populations(), ref.f(), "f after 50 pulsatile steps")


def test_reduction_kernel_matches_host_sums(tube_geometry):
    sol = lbm.Solver(tube_geometry, tau=TAU, dt=DT, cte_vel=U)
    sol.step(50)
    s = sol.reduce()
    fs = sol.post_collision()
    c, _, _ = lbm.lattice(19)
    mass = fs.sum()
    mom = fs @ c.astype(np.float64)
    assert abs(s.mass - mass) <= 1e-12 * mass
    np.testing.assert_allclose(np.array(s.momentum), mom.sum(axis=0), rtol=1e-10, atol=1e-12)
    w, _, _ = sol.wss()
    assert abs(s.wss_l2 - np.sqrt((w * w).sum())) <= 1e-12 * max(1.0, s.wss_l2)
    assert s.wss_max == w.max()
    assert s.iteration == 50


@pytest.mark.parametrize("prec,tol", [(lbm.PREC_F32, 2e-5), (lbm.PREC_F16, 2e-2)])
def test_reduced_precision_tracks_fp64(tube_geometry, prec, tol):
    a = lbm.Solver(tube_geometry, tau=TAU, dt=DT, cte_vel=U)
    b = lbm.Solver(tube_geometry, tau=TAU, dt=DT, cte_vel=U, precision=prec, parity_checks=0)
    a.step(200)
    b.step(200)
    fa, fb = a.populations(), b.populations()
    _, _, w = lbm.lattice(19)
    err = np.max(np.abs(fa - fb) / w[None, :])
    print(f"precision {prec}: max |f - f64| / w = {err:.3e}")
    assert err < tol
