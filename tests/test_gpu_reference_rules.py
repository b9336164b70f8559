"""Reference boundary rules on the GPU beyond the -z inlet / +z outlet tube.

Each case runs the fused FP64 kernel (reference-rule family) and the oracle
on the same tables and compares populations, post-collision populations and
the wall-shear-stress accumulators bit for bit:

* the six branches of Boundary::ApplyBc (Boundary.cpp:111-289) and the six
  outlet normals of Simulation::ApplyOutletBc (Simulation.cpp:650-741), on a
  straight tube along each axis with the inlet at either end (the geometry
  tables of every one of these tubes equal the compiled reference's,
  tests/test_reference_build.py). The -y branch divides by 1 - 1.0
  (Boundary.cpp:206) and the +y branch imposes a speed of 1.0 whatever the
  configured speed: both quirks are kept, so the -y run carries inf/NaN
  exactly where the oracle does;
* the pulsatile inlet of Simulation::Iterate (Simulation.cpp:558-562,
  519-534) across all three polynomial pieces, the zero tail and the wrap of
  the 0.9 s cycle;
* the f* == 0.0 push skip (Simulation.cpp:245-251) on a wall node, whose
  receivers keep their previous values, and on a plain fluid node, which the
  product refuses with LBM_ERR_PARITY because it keeps no previous values
  for plain fluid nodes.
"""
import numpy as np
import pytest

import lbm
import oracle
import workloads

pytestmark = pytest.mark.gpu


def _same(a, b, what):
    assert a.shape == b.shape, what
    assert np.array_equal(np.isnan(a), np.isnan(b)), f"{what}: NaN positions differ"
    fin = ~np.isnan(a)
    eq = np.array_equal(a[fin], b[fin])
    if not eq:
        # the reference rules can drive values to +-inf; the difference is
        # reported over the entries that are finite in both
        both = np.isfinite(a) & np.isfinite(b)
        d = np.max(np.abs(a[both] - b[both])) if both.any() else float("nan")
        pytest.fail(f"{what}: max |diff| over finite entries = {d:.3e}")


def _collided(ref):
    ref.collide()
    return ref.fs()


def _compare(sol, ref, what):
    _same(sol.populations(), ref.f(), f"f {what}")
    _same(sol.post_collision(), _collided(ref), f"f* {what}")
    w, acc, mag = sol.wss()
    rw, racc, rmag = ref.wss()
    _same(w, rw, f"wss {what}")
    _same(acc, racc, f"accumulated tau {what}")
    _same(mag, rmag, f"accumulated |tau| {what}")


@pytest.mark.parametrize("axis,end", [(a, e) for a in "xyz" for e in ("min", "max")])
def test_zou_he_branch_and_outlet_normal_bit_exact(tmp_path, axis, end):
    U, TAU, DT = 0.02, 0.8, 1e-5
    cfg = workloads.axis_tube(str(tmp_path), axis=axis, inlet_end=end, cte_vel=U)
    geom = lbm.Geometry.from_config(cfg)
    inl = np.array(geom.info.inlet_normal)
    onrm = geom.outlet_normals()[0]
    print(f"tube along {axis}, inlet normal {inl}, outlet normal {onrm}, {geom.n} nodes", flush=True)
    sol = lbm.Solver(geom, tau=TAU, dt=DT, cte_vel=U)
    ref = oracle.Reference.from_geometry(geom, tau=TAU, dt=DT)
    done = 0
    for target in (1, 5, 50, 200):
        sol.step(target - done)
        ref.step(np.full(target - done, U), nthreads=8)
        done = target
        _compare(sol, ref, f"after {done} steps")
    f = ref.f()
    if axis == "y" and end == "min":
        assert np.isnan(f).any()  # the reference's division by 1 - 1.0
    else:
        assert np.isfinite(f).all()


def _pulsatile_speeds(n, dt):
    t, u = 0.0, []
    for _ in range(n):
        u.append(oracle.inlet_velocity(t, 0.9))
        t += dt
    return np.array(u)


@pytest.mark.parametrize("dt,checkpoints", [(2e-3, (45, 46, 135, 136, 405, 406, 450, 451, 500)),
                                            (0.05, (2, 3, 6, 7, 17, 18, 19, 40, 56))])
def test_pulsatile_inlet_all_pieces_bit_exact(tmp_path, dt, checkpoints):
    # dt = 2e-3: t crosses 0.09 (piece 1 -> 2) at step 45, 0.27 (-> 3) at
    # 135, 0.81 (-> zero tail) at 405 and the 0.9 wrap at 450. The reference
    # applies these speeds (up to 0.5) directly as lattice velocities, and
    # the populations grow without bound over such a run (to ~1e207 after
    # 500 steps here): the values stay finite and must agree bit for bit.
    # dt = 0.05 crosses the same bounds within 20 steps and runs past the
    # 2.7 s at which Iterate stops (55 iterations) with bounded populations.
    TAU = 1.0
    cfg = workloads.axis_tube(str(tmp_path), axis="z", inlet_end="min", radius=6.0, length=30.0)
    geom = lbm.Geometry.from_config(cfg)
    n = max(checkpoints)
    u = _pulsatile_speeds(n, dt)
    pieces = set()
    for t in np.arange(n) * dt:
        x = np.fmod(t, 0.9)
        pieces.add(0 if x < 0.09 else 1 if x < 0.27 else 2 if x < 0.81 else 3)
    assert pieces == {0, 1, 2, 3}
    sol = lbm.Solver(geom, tau=TAU, dt=dt, flow_type=lbm.FLOW_PULSATILE)
    ref = oracle.Reference.from_geometry(geom, tau=TAU, dt=dt)
    done = 0
    for target in checkpoints:
        sol.step(target - done)
        ref.step(u[done:target], nthreads=8)
        done = target
        print(f"dt {dt}: step {done}", flush=True)
        _compare(sol, ref, f"after {done} pulsatile steps (dt {dt})")
    assert np.isfinite(ref.f()).all()
    assert sol.reduce().time == pytest.approx(n * dt, rel=1e-12)


def _tube(tmp_path):
    cfg = workloads.axis_tube(str(tmp_path), axis="z", inlet_end="min", radius=5.0, length=24.0)
    return lbm.Geometry.from_config(cfg)


def test_zero_wall_node_is_not_streamed_bit_exact(tmp_path):
    # A wall node whose populations are all 0.0 collides to f* = 0 exactly
    # (u = 0 on walls, rho = 0): none of its values is pushed and its
    # receivers, all near-wall or boundary nodes, keep their previous values.
    geom = _tube(tmp_path)
    _, kind, _ = geom.nodes()
    nb = geom.neighbours()
    nw, _ = geom.nearwall()
    nwset = set(nw.tolist())
    walls = np.where(kind == lbm.KIND_WALL)[0]
    pick = next(int(x) for x in walls if any(int(nb[q, x]) in nwset for q in range(19) if nb[q, x] >= 0))
    _, _, w = lbm.lattice(19)
    rng = np.random.default_rng(7)
    f0 = w[None, :] * (1.0 + 0.05 * rng.uniform(-1, 1, size=(geom.n, 19)))
    f0[pick] = 0.0
    sol = lbm.Solver(geom, tau=0.8, dt=1e-5, cte_vel=0.02)
    ref = oracle.Reference.from_geometry(geom, tau=0.8, dt=1e-5)
    sol.set_populations(f0)
    ref.set_f(f0)
    _same(sol.post_collision()[pick], np.zeros(19), "f* of the zero wall node")
    sol.step(1)
    ref.step([0.02])
    _compare(sol, ref, "after the zero wall node's step")
    sol.step(4)
    ref.step(np.full(4, 0.02))
    _compare(sol, ref, "4 steps later")


def test_zero_plain_fluid_node_is_refused(tmp_path):
    # With tau = 2/3 (1/tau = 1.5 exactly) a plain fluid node holding only
    # f_rest = 3 has rho = 3, u = 0, f_eq,rest = 1 and f*_rest = 3 + 1.5 (1 - 3)
    # = 0.0 exactly. The reference keeps the old f_rest there (no push); the
    # product keeps previous values only for boundary and near-wall nodes, so
    # it stops with LBM_ERR_PARITY instead of diverging silently.
    geom = _tube(tmp_path)
    _, kind, _ = geom.nodes()
    nb = geom.neighbours()
    nw, _ = geom.nearwall()
    bulk = np.where((kind == lbm.KIND_FLUID) & ~np.isin(np.arange(geom.n), nw) & (nb >= 0).all(axis=0))[0]
    pick = int(bulk[len(bulk) // 2])
    _, _, w = lbm.lattice(19)
    f0 = np.tile(w, (geom.n, 1))
    f0[pick] = 0.0
    f0[pick, 18] = 3.0
    tau = 2.0 / 3.0
    assert 1.0 / tau == 1.5
    ref = oracle.Reference.from_geometry(geom, tau=tau, dt=1e-5)
    ref.set_f(f0)
    ref.collide()
    assert ref.fs()[pick, 18] == 0.0
    ref.step([0.02])
    assert ref.f()[pick, 18] == 3.0  # the reference keeps the old value
    sol = lbm.Solver(geom, tau=tau, dt=1e-5, cte_vel=0.02)
    sol.set_populations(f0)
    assert sol.post_collision()[pick, 18] == 0.0
    with pytest.raises(lbm.LbmError) as e:
        sol.step(1)
    assert e.value.code == lbm.ERR_PARITY
    # without the check the step runs (and differs from the reference there)
    quiet = lbm.Solver(geom, tau=tau, dt=1e-5, cte_vel=0.02, parity_checks=0)
    quiet.set_populations(f0)
    quiet.step(1)
