"""Voxeliser, classifier and node tables against the brute-force oracle.

The product evaluates the reference's ray predicate only for triangles whose
xy box (plus a margin) holds the ray; the oracle evaluates every point
against every triangle (ComplexGeometry.cpp:440-514) and classifies nodes by
a dense scan (Subdomain.cpp:364-471). Labels, node order, kinds, outlet ids
and the push table must agree exactly; near-wall normals bit for bit.
"""
import numpy as np
import pytest

import lbm
import oracle
import workloads


def _geom_from_surface(tmp_path, tris, nrm, grid_size, inlets, outlets, q=19):
    stl = tmp_path / "s.stl"
    workloads.write_stl(str(stl), tris, nrm)
    cfg = tmp_path / "config.txt"
    workloads.write_config(str(cfg), "s.stl", tau=0.8, grid_size=grid_size, inlets=inlets, outlets=outlets)
    return lbm.Geometry.from_config(str(cfg), q=q, verify_stride=3)


def _compare(g, bf, check_normals=True):
    assert (g.labels() == bf["fluid"]).all()
    ijk, kind, oid = g.nodes()
    assert (ijk == bf["ijk"]).all()
    assert (kind == bf["kind"]).all()
    assert (oid == bf["outlet_id"]).all()
    assert (g.neighbours() == bf["nb"]).all()
    if check_normals:
        nw, nrm = g.nearwall()
        for a in range(0, len(nw), max(1, len(nw) // 40)):
            p = ijk[nw[a]]
            x = bf["bbox"][0] + p[0] * bf["spacing"]
            y = bf["bbox"][2] + p[1] * bf["spacing"]
            z = bf["bbox"][4] + p[2] * bf["spacing"]
            assert (oracle.normal_of(bf, x, y, z) == nrm[a]).all()


def test_tube_matches_brute_force(tmp_path):
    tris, nrm = workloads.tube_surface(6.0, 20.0, segments=32)
    ins = [dict(plane="z", coordinate=0.0, normal=(0, 0, -1))]
    outs = [dict(plane="z", coordinate=20.0, normal=(0, 0, 1))]
    g = _geom_from_surface(tmp_path, tris, nrm, 0.5, ins, outs)
    _compare(g, oracle.brute_force_geometry(tris, nrm, 0.5, ins, outs))
    i = g.info
    assert i.n_inlet > 0 and i.n_outlet > 0 and i.n_wall > 0 and i.n_nearwall > 0


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_blob_matches_brute_force(tmp_path, seed):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=4) * 0.15

    def radial(d):
        return 1.0 + a[0] * d[:, 0] * d[:, 1] + a[1] * d[:, 2] ** 2 + a[2] * np.sin(3 * d[:, 0]) + a[3] * d[:, 1]

    tris, nrm = workloads.icosphere((0.3, -0.2, 0.1), 7.3, subdiv=3, rotation_seed=seed, radial=radial)
    g = _geom_from_surface(tmp_path, tris, nrm, 0.45, [], [])
    _compare(g, oracle.brute_force_geometry(tris, nrm, 0.45, [], []))


def test_outlet_precedence_and_ranges(tmp_path):
    tris, nrm = workloads.tube_surface(5.0, 12.0, segments=24)
    ins = [dict(plane="z", coordinate=0.0, normal=(0, 0, -1))]
    outs = [dict(plane="z", coordinate=12.0, normal=(0, 0, 1), range_x=(-100.0, 0.5)),
            dict(plane="z", coordinate=12.0, normal=(0, 0, 1)),
            dict(plane="x", coordinate=-3.0, normal=(-1, 0, 0))]
    g = _geom_from_surface(tmp_path, tris, nrm, 0.5, ins, outs)
    bf = oracle.brute_force_geometry(tris, nrm, 0.5, ins, outs)
    _compare(g, bf)
    _, kind, oid = g.nodes()
    assert set(oid[kind == lbm.KIND_OUTLET]) >= {0, 1}


def test_shared_edge_double_count_is_reproduced(tmp_path):
    # A box whose top face diagonal passes exactly over grid columns: the
    # reference's inclusive inside test counts both triangles there, which
    # flips the parity (ComplexGeometry.cpp:257-259). Both sides must agree.
    lo, hi = np.array([0.0, 0.0, 0.0]), np.array([4.0, 4.0, 3.0])
    corners = np.array([[x, y, z] for x in (lo[0], hi[0]) for y in (lo[1], hi[1]) for z in (lo[2], hi[2])])
    quads = [((0, 1, 3, 2), (-1, 0, 0)), ((4, 6, 7, 5), (1, 0, 0)), ((0, 4, 5, 1), (0, -1, 0)),
             ((2, 3, 7, 6), (0, 1, 0)), ((0, 2, 6, 4), (0, 0, -1)), ((1, 5, 7, 3), (0, 0, 1))]
    tris, nrm = [], []
    for (a, b, c, d), n in quads:
        tris += [corners[[a, b, c]], corners[[a, c, d]]]
        nrm += [n, n]
    tris, nrm = np.array(tris), np.array(nrm, dtype=np.float64)
    g = _geom_from_surface(tmp_path, tris, nrm, 0.5, [], [])
    bf = oracle.brute_force_geometry(tris, nrm, 0.5, [], [])
    _compare(g, bf, check_normals=False)
    lab = g.labels()
    assert lab[2, 2, 1] == 0  # on the diagonal: counted twice -> "outside"
    assert lab[1, 2, 1] == 1


def test_direct_ray_predicate_against_oracle(tmp_path):
    tris, nrm = workloads.tube_surface(6.0, 20.0, segments=32)
    bf = oracle.brute_force_geometry(tris, nrm, 0.5, [], [])
    g = _geom_from_surface(tmp_path, tris, nrm, 0.5, [], [])
    assert (g.labels() == bf["fluid"]).all()


def test_periodic_labels_and_wrap():
    lab, origin, _ = workloads.tube_mask(4.0, 10)
    g = lbm.Geometry.from_labels(lab, origin, 1.0, periodic=(0, 0, 1))
    ijk, kind, _ = g.nodes()
    nb = g.neighbours()
    c, _, _ = lbm.lattice(19)
    nz = lab.shape[2]
    # no wall nodes on the periodic faces: kinds depend on (i, j) only
    per_col = {}
    for n in range(g.n):
        per_col.setdefault((ijk[n, 0], ijk[n, 1]), set()).add(int(kind[n]))
    assert all(len(v) == 1 for v in per_col.values())
    # +z neighbour of the top layer is the bottom layer of the same column
    top = np.where(ijk[:, 2] == nz - 1)[0]
    for n in top[:20]:
        m = nb[4, n]
        assert m >= 0 and ijk[m, 2] == 0 and (ijk[m, :2] == ijk[n, :2]).all()


def test_bc_neighbour_points_behind_the_plane(tube_geometry):
    ijk, kind, _ = tube_geometry.nodes()
    bcn = tube_geometry.bc_neighbour()
    for n in np.where(kind == lbm.KIND_OUTLET)[0]:
        m = bcn[n]
        assert m >= 0 and (ijk[m] == ijk[n] - [0, 0, 1]).all()
    for n in np.where(kind == lbm.KIND_INLET)[0]:
        m = bcn[n]
        assert m >= 0 and (ijk[m] == ijk[n] + [0, 0, 1]).all()
    assert (bcn[kind <= lbm.KIND_WALL] == -1).all()


def test_tube_counts(tube_geometry):
    i = tube_geometry.info
    # ~25k nodes for the radius-10, length-80 tube; every node is classified
    assert 24000 < i.n_nodes < 26000
    assert i.n_fluid + i.n_wall + i.n_inlet + i.n_outlet == i.n_nodes
    assert i.n_inlet == i.n_outlet  # identical end discs
    assert tuple(i.grid) == (21, 21, 81) and i.spacing == 1.0


def test_cylinder_normals_are_radial_mid_tube(tube_geometry):
    ijk, _, _ = tube_geometry.nodes()
    nw, nrm = tube_geometry.nearwall()
    b = tube_geometry.info.bbox
    mid = np.abs(ijk[nw, 2] - 40) <= 2
    x = b[0] + ijk[nw, 0] - 0.37
    y = b[2] + ijk[nw, 1] - 0.21
    radial = np.stack([x, y, 0 * x], 1)
    radial /= np.linalg.norm(radial, axis=1, keepdims=True)
    cos = (nrm[mid] * radial[mid]).sum(1)
    assert np.all(cos > 0.95)
