"""OSI and TAWSS (Simulation::ComputeOSIAndTAWSS, Simulation.cpp:804-826) on
the solver path, pinned by the values the formulas take for stress
histories whose shape is known:

  TAWSS = (sum_t |tau_t|) / 0.9 * dt,  OSI = (1 - |sum_t tau_t| / sum_t |tau_t|) / 2.

The state f(t) is set before every step (lbm_set_populations, which leaves
the stress accumulators alone), so each step's stress is a function of the
set state only:
* the same state every step gives the same stress vector every step: OSI is 0
  (up to the rounding of the sums) and TAWSS is T |tau| dt / 0.9;
* states w (1 + d xi) and w (1 - d xi) alternately give, to first order in
  d, opposite stress vectors (the stress vanishes at rest, f = w, and is
  smooth in f): OSI is 1/2 up to O(d).
Both stress families run: the reference loops (reference wall rule) and the
column kernels' form (half-way bounce-back).
"""
import numpy as np
import pytest

import lbm
import workloads

pytestmark = pytest.mark.gpu

DT = 1e-3


def _solver(geom, family):
    if family == "reference":
        return lbm.Solver(geom, tau=0.8, dt=DT, cte_vel=0.0)
    return lbm.Solver(geom, tau=0.8, dt=DT, cte_vel=0.0, wall_mode=lbm.WALL_HALFWAY_BB,
                      inlet_mode=lbm.INLET_NONE, outlet_mode=lbm.OUTLET_NONE, parity_checks=0)


def _states(geom, d, seed):
    _, _, w = lbm.lattice(19)
    xi = np.random.default_rng(seed).uniform(-1, 1, size=(geom.n, 19))
    return w[None, :] * (1.0 + d * xi), w[None, :] * (1.0 - d * xi)


def _run(sol, seq):
    mags = []
    for f in seq:
        _, acc0, mag0 = sol.wss()
        sol.set_populations(f)
        _, acc1, mag1 = sol.wss()
        assert np.array_equal(acc0, acc1) and np.array_equal(mag0, mag1)  # set leaves the sums alone
        sol.step(1)
        mags.append(sol.wss()[0])
    return np.array(mags)


@pytest.fixture(scope="module")
def tube(tmp_path_factory):
    d = tmp_path_factory.mktemp("osi")
    return lbm.Geometry.from_config(workloads.axis_tube(str(d), axis="z", radius=5.0, length=16.0))


@pytest.mark.parametrize("family", ["reference", "columns"])
def test_constant_stress_gives_zero_osi(tube, family):
    T = 12
    P, _ = _states(tube, 1e-3, 1)
    sol = _solver(tube, family)
    mags = _run(sol, [P] * T)
    assert np.all(mags == mags[0][None, :])  # the same stress every step
    osi, tawss = sol.osi_tawss()
    _, acc, mag = sol.wss()
    live = mag > 0
    assert live.sum() > 0.9 * len(mag)
    assert np.max(np.abs(osi[live])) < 1e-14
    # TAWSS: the sum of the per-step magnitudes in step order, / 0.9 * dt
    s = np.zeros_like(mags[0])
    for m in mags:
        s = s + m
    assert np.array_equal(mag, s)
    np.testing.assert_allclose(tawss, s / 0.9 * DT, rtol=1e-15, atol=0)
    np.testing.assert_allclose(tawss[live], T * mags[0][live] * DT / 0.9, rtol=1e-14)


@pytest.mark.parametrize("family", ["reference", "columns"])
def test_reversing_stress_gives_half_osi(tube, family):
    d = 1e-7
    P, M = _states(tube, d, 2)
    sol = _solver(tube, family)
    mags = _run(sol, [P, M] * 6)
    osi, tawss = sol.osi_tawss()
    _, _, mag = sol.wss()
    live = mag > 1e-6 * mag.max()
    assert live.sum() > 0.9 * len(mag)
    dev = np.max(np.abs(osi[live] - 0.5))
    print(f"{family}: max |OSI - 1/2| = {dev:.2e} at d = {d}")
    assert dev < 1e-5
    np.testing.assert_allclose(tawss, mags.sum(axis=0) / 0.9 * DT, rtol=1e-14)
