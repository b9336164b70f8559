"""Simulation::WriteOutput through the C ABI (lbm_write_output): the VTU
carries the last collision's density and velocity, the near-wall CSV the
solver's WSS / OSI / TAWSS, the smoothed copy the smoothing of that CSV."""
import numpy as np
import pytest

import lbm
from test_output import read_vtu

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("wall_mode", [lbm.WALL_REFERENCE, lbm.WALL_HALFWAY_BB])
def test_write_output_files(tmp_path, tube_geometry, wall_mode):
    g = tube_geometry
    kw = {} if wall_mode == lbm.WALL_REFERENCE else dict(inlet_mode=lbm.INLET_NEE_VELOCITY,
                                                         outlet_mode=lbm.OUTLET_NEE_PRESSURE)
    sol = lbm.Solver(g, tau=0.8, dt=1e-5, cte_vel=0.04, wall_mode=wall_mode, parity_checks=0, **kw)
    sol.step(30)
    sol.write_output(g, 30, "Iteration_30", tmp_path, pulsatile=True)
    # VTU: moments of the post-collision state, in z-y-x order.
    fs = sol.post_collision()
    c, _, _ = lbm.lattice(19)
    rho = fs.sum(1)
    u = (fs @ c) / rho[:, None]
    ijk, kind, _ = g.nodes()
    if wall_mode == lbm.WALL_REFERENCE:
        u[kind == lbm.KIND_WALL] = 0.0
    order = np.lexsort((ijk[:, 0], ijk[:, 1], ijk[:, 2]))
    npts, ncells, a = read_vtu(tmp_path / "Fluid_Variables" / "30.vtu")
    assert npts == g.n and ncells > 0
    assert np.allclose(a["Density"], rho[order], rtol=1e-14, atol=0)
    assert np.allclose(a["Velocity"], u[order], rtol=1e-12, atol=1e-16)
    # Near-wall CSV: the solver's fields as printed (6 significant digits).
    wss = sol.wss()[0]
    osi, tawss = sol.osi_tawss()
    nodes, _ = g.nearwall()
    vm = np.sqrt((u[nodes] ** 2).sum(1))
    raw = np.loadtxt(tmp_path / "30_FlowHemodynamics.csv", delimiter=",", skiprows=1, ndmin=2)
    assert open(tmp_path / "30_FlowHemodynamics.csv").readline().strip() == "X,Y,Z,WSS,OSI,TAWSS,VelMag"
    assert raw.shape == (len(nodes), 7)
    for col, want in ((3, wss), (4, osi), (5, tawss), (6, vm)):
        assert np.allclose(raw[:, col], want, rtol=1e-5, atol=1e-300, equal_nan=True), col
    sm = tmp_path / "FlowHemodynamics" / "Iteration_30_30_FlowHemodynamics.csv"
    lbm.smooth_surface_csv(tmp_path / "30_FlowHemodynamics.csv", tmp_path / "again.csv", 0.25, True)
    assert open(sm).read() == open(tmp_path / "again.csv").read()
