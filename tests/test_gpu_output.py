"""Simulation::WriteOutput through the C ABI (lbm_write_output): the VTU
carries the last collision's density and velocity, the near-wall CSV the
solver's WSS / OSI / TAWSS, the smoothed copy the smoothing of that CSV."""
import numpy as np
import pytest

import lbm
import oracle
from test_output import read_vtu

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("wall_mode", [lbm.WALL_REFERENCE, lbm.WALL_HALFWAY_BB])
def test_write_output_files(tmp_path, tube_geometry, wall_mode):
    g = tube_geometry
    kw = {} if wall_mode == lbm.WALL_REFERENCE else dict(inlet_mode=lbm.INLET_NEE_VELOCITY,
                                                         outlet_mode=lbm.OUTLET_NEE_PRESSURE)
    sol = lbm.Solver(g, tau=0.8, dt=1e-5, cte_vel=0.04, wall_mode=wall_mode, parity_checks=0, **kw)
    with pytest.raises(lbm.LbmError):  # no iteration yet: nothing to write
        sol.write_output(g, 0, "Iteration_0", tmp_path, pulsatile=True)
    sol.step(30)
    sol.write_output(g, 30, "Iteration_30", tmp_path, pulsatile=True)
    # VTU: the density and velocity that the 30th iteration's collision set
    # (of f after 29 steps), as the oracle's reference step keeps them.
    ref = oracle.Reference.from_geometry(g, tau=0.8, dt=1e-5, wall_mode=0 if wall_mode == lbm.WALL_REFERENCE else 1,
                                         inlet_mode=0 if wall_mode == lbm.WALL_REFERENCE else 1,
                                         outlet_mode=0 if wall_mode == lbm.WALL_REFERENCE else 1)
    ref.step(np.full(30, 0.04))
    rho, u = ref.macro()
    ijk, kind, _ = g.nodes()
    order = np.lexsort((ijk[:, 0], ijk[:, 1], ijk[:, 2]))
    npts, ncells, a = read_vtu(tmp_path / "Fluid_Variables" / "30.vtu")
    assert npts == g.n and ncells > 0
    # Reference family: populations bit-exact with the oracle, so only the
    # rounding between the moments of f and of f* remains. Bounce-back
    # family: populations within 1e-12 of the oracle (FMA contraction).
    tol = 1e-14 if wall_mode == lbm.WALL_REFERENCE else 1e-11
    assert np.allclose(a["Density"], rho[order], rtol=tol, atol=0)
    assert np.allclose(a["Velocity"], u[order], rtol=0, atol=tol)
    # the step after it would differ: the fields are those of iteration 30
    ref.step(np.full(1, 0.04))
    assert not np.allclose(a["Velocity"], ref.macro()[1][order], rtol=0, atol=1e-12)
    # Near-wall CSV: the solver's fields as printed (6 significant digits).
    wss = sol.wss()[0]
    osi, tawss = sol.osi_tawss()
    nodes, _ = g.nearwall()
    vm = np.sqrt((u[nodes] ** 2).sum(1))
    raw = np.loadtxt(tmp_path / "30_FlowHemodynamics.csv", delimiter=",", skiprows=1, ndmin=2)
    assert open(tmp_path / "30_FlowHemodynamics.csv").readline().strip() == "X,Y,Z,WSS,OSI,TAWSS,VelMag"
    assert raw.shape == (len(nodes), 7)
    for col, want in ((3, wss), (4, osi), (5, tawss), (6, vm)):
        # VelMag: |u| from the moments of f* (the reference takes f's), so a
        # node still at rest prints ~1e-17 where the oracle has exactly 0.
        atol = 1e-15 if col == 6 else 1e-300
        assert np.allclose(raw[:, col], want, rtol=1e-5, atol=atol, equal_nan=True), col
    sm = tmp_path / "FlowHemodynamics" / "Iteration_30_30_FlowHemodynamics.csv"
    lbm.smooth_surface_csv(tmp_path / "30_FlowHemodynamics.csv", tmp_path / "again.csv", 0.25, True)
    assert open(sm).read() == open(tmp_path / "again.csv").read()
