"""Synthetic workload generators (CPU): the tortuous vessel of the large
single-GPU share at a small scale."""
import math

import numpy as np
import pytest

import lbm
import workloads


@pytest.fixture(scope="module")
def vessel():
    return workloads.tortuous_vessel(scale=0.5, device="cpu")


def test_vessel_volume_matches_the_swept_tube(vessel):
    fluid, origin, info = vessel
    # Swept tube volume (pi r^2 ds) plus the bulge ball, less their overlap
    # bounded above by the smaller of the two: the node count lies between.
    s = np.linspace(0.0, info["length"], len(info["radius"]))
    tube = float(np.trapezoid(math.pi * info["radius"] ** 2, s))
    ball = 4.0 / 3.0 * math.pi * info["bulge_radius"] ** 3
    n = int(fluid.sum())
    assert max(tube, ball) * 0.95 <= n <= (tube + ball) * 1.05, (n, tube, ball)


def test_vessel_ends_on_the_z_faces_only(vessel):
    fluid, _, _ = vessel
    assert fluid[:, :, 0].sum() > 0 and fluid[:, :, -1].sum() > 0
    assert fluid[0].sum() == fluid[-1].sum() == fluid[:, 0].sum() == fluid[:, -1].sum() == 0


def test_vessel_geometry_builds_with_outward_normals(vessel):
    fluid, origin, info = vessel
    nz = fluid.shape[2]
    g = lbm.Geometry.from_labels(fluid, origin, 1.0,
                                 inlets=[lbm.plane_block("z", float(origin[2]), (0, 0, -1))],
                                 outlets=[lbm.plane_block("z", float(origin[2] + nz - 1), (0, 0, 1))],
                                 triangles=info["triangles"], normals=1, nthreads=8)
    assert g.n == int(fluid.sum())
    assert g.info.n_inlet > 0 and g.info.n_outlet > 0 and g.info.n_wall > 0
    nodes, nrm = g.nearwall()
    ijk, _, _ = g.nodes()
    p = origin[None, :] + ijk[nodes]
    # Outward: from the nearest centreline point (or the bulge centre) to the node.
    C = info["centreline"]
    k = np.argmin(((p[:, None, :] - C[None, ::4, :]) ** 2).sum(-1), axis=1)
    radial = p - C[::4][k]
    in_bulge = np.linalg.norm(p - info["bulge_center"], axis=1) < 0.9 * info["bulge_radius"]
    radial[in_bulge] = (p - info["bulge_center"])[in_bulge]
    cosang = (radial * nrm).sum(1) / np.maximum(np.linalg.norm(radial, axis=1) * np.linalg.norm(nrm, axis=1), 1e-12)
    assert np.mean(cosang > 0.5) > 0.95


def test_vessel_stl_is_closed_and_matches_the_labels(vessel, tmp_path):
    """The vessel's STL (the large share's surface, here at scale 0.5): every
    edge is shared by exactly two facets (marching tetrahedra interpolate a
    shared edge once, so the copies are bit-identical), and the set-up chain
    from it labels nearly the same points as the ball union."""
    fluid, _, _ = vessel
    cfg, info = workloads.tortuous_vessel_stl(str(tmp_path), scale=0.5, mesh_h=2.0)
    g = lbm.Geometry.from_config(cfg, normals=0)
    tris = g.triangles()[:, :9].reshape(-1, 3, 3)  # 3 vertices, then the file normal
    assert len(tris) == info["n_triangles"]
    v, inv = np.unique(tris.reshape(-1, 3), axis=0, return_inverse=True)
    t = inv.reshape(-1, 3)
    e = np.sort(np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]]), axis=1)
    _, count = np.unique(e, axis=0, return_counts=True)
    assert (count == 2).all()
    n = int(fluid.sum())
    assert abs(g.n - n) <= 0.01 * n, (g.n, n)
    assert g.info.n_inlet > 0 and g.info.n_outlet > 0
