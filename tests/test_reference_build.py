"""The product's host chain against the reference's own compiled code.

oracle/_ref/libref.so is built from the reference sources (ComplexGeometry,
Subdomain, Boundary, Node, LatticeModel, VelocitySet, GeometryReader, Utility
and the header-defined parser in LBM.h) by oracle/ref/Makefile, unmodified,
and driven on one MPI rank. These tests skip where it is absent (it is only
built where the reference sources exist).
"""
import os

import numpy as np
import pytest

import lbm
import oracle
import workloads
from conftest import GOLDEN

refbind = pytest.importorskip("refbind")
pytestmark = pytest.mark.skipif(not refbind.available(), reason="oracle/_ref/libref.so not built here")
JT_STL = "/root/reference/JT.stl"


def _compare_tables(g, r):
    ijk, kind, oid = g.nodes()
    assert tuple(g.info.grid) == r["grid"]
    assert g.info.spacing == r["spacing"]
    assert np.array_equal(np.array(g.info.bbox), r["bbox"])
    assert np.array_equal(g.labels(), r["labels"])
    assert np.array_equal(ijk, r["ijk"])
    assert np.array_equal(kind, r["kind"])
    assert np.array_equal(oid, r["outlet_id"])
    assert np.array_equal(g.neighbours(), r["nb"])
    assert np.array_equal(np.array(g.info.inlet_normal), r["inlet_normal"])
    nw, nrm = g.nearwall()
    # IdentifyNextToWallLattices (Subdomain.cpp:483) tests the uninitialised
    # BoundaryType_ of plain fluid nodes against Wall; where the heap happens
    # to hold that value the reference drops the node. The product treats
    # the field as unset (never Wall); compare modulo exactly those nodes.
    garbage_wall = (kind == 0) & (r["fluid_bt"] == 2)
    keep = ~garbage_wall[nw]
    nw, nrm = nw[keep], nrm[keep]
    assert len(nw) == len(r["nw_ijk"])
    pos = {tuple(p): i for i, p in enumerate(r["nw_ijk"])}
    idx = np.array([pos[tuple(ijk[n])] for n in nw], dtype=np.int64)
    assert np.array_equal(nrm, r["nw_normal"][idx])   # bit for bit, any order
    # The same uninitialised field also zeroes the velocity of a plain fluid
    # node's collision when it reads Wall (Simulation.cpp:92): record how
    # often this run's heap made it so.
    print(f"plain fluid nodes whose uninitialised BoundaryType_ read Wall: {int(garbage_wall.sum())}"
          f" of {int((kind == 0).sum())}")


def test_lattice_tables_equal_reference():
    c, opp, w = refbind.lattice()
    c2, opp2, w2 = lbm.lattice(19)
    assert np.array_equal(c, c2) and np.array_equal(opp, opp2) and np.array_equal(w, w2)


def test_tube_tables_equal_reference(tmp_path):
    cfg = workloads.straight_tube(str(tmp_path), radius=6.0, length=24.0)
    _compare_tables(lbm.Geometry.from_config(cfg), refbind.build(cfg))


@pytest.mark.parametrize("seed", [4, 5])
def test_blob_with_outlets_equal_reference(tmp_path, seed):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=3) * 0.12

    def radial(d):
        return 1.0 + a[0] * d[:, 0] * d[:, 2] + a[1] * d[:, 1] ** 2 + a[2] * d[:, 0]

    tris, nrm = workloads.icosphere((0.11, 0.07, -0.3), 6.1, subdiv=3, rotation_seed=seed, radial=radial)
    stl = tmp_path / "s.stl"
    workloads.write_stl(str(stl), tris, nrm)
    zmin = tris[..., 2].min()
    cfg = tmp_path / "config.txt"
    workloads.write_config(str(cfg), "s.stl", tau=0.8, grid_size=0.4,
                           inlets=[dict(plane="x", coordinate=float(tris[..., 0].min()), normal=(-1, 0, 0))],
                           outlets=[dict(plane="z", coordinate=float(zmin), normal=(0, 0, -1),
                                         range_x=(-100.0, 0.5)),
                                    dict(plane="z", coordinate=float(zmin), normal=(0, 0, -1))])
    _compare_tables(lbm.Geometry.from_config(str(cfg)), refbind.build(str(cfg)))


@pytest.mark.skipif(not os.path.exists(JT_STL), reason="reference JT.stl not present")
def test_clinical_stl_coarse_grid_equal_reference(tmp_path):
    # The reference's own vessel surface (9,712 facets) and its config.txt
    # inlet/outlet blocks, on a coarse pitch so the reference's O(points x
    # triangles) voxeliser finishes quickly.
    os.symlink(JT_STL, tmp_path / "JT.stl")
    text = open(os.path.join(GOLDEN, "reference_config.txt")).read().replace("grid_size: 0.025", "grid_size: 0.4")
    (tmp_path / "config.txt").write_text(text)
    g = lbm.Geometry.from_config(str(tmp_path / "config.txt"))
    r = refbind.build(str(tmp_path / "config.txt"))
    _compare_tables(g, r)
    assert g.info.n_triangles == 9712
    assert np.array_equal(g.triangles(), refbind.read_stl(JT_STL))


def test_stl_reader_equal_reference(tmp_path):
    tris, nrm = workloads.tube_surface(3.0, 5.0, segments=12)
    p = tmp_path / "a.stl"
    workloads.write_stl(str(p), tris, nrm)
    text = p.read_text().splitlines()
    # damage a few facets the way real files break: a vertex line short of a
    # coordinate, a misspelt keyword, a missing endloop
    text[3] = "      vertex 1.0 2.0"
    text[9] = "    outer lop"
    text[20] = "  endfacetx"
    q = tmp_path / "b.stl"
    q.write_text("\n".join(text) + "\n")
    for path in (p, q):
        cfg = tmp_path / f"{path.stem}.txt"
        workloads.write_config(str(cfg), path.name, tau=0.8)
        g = lbm.Geometry.from_config(str(cfg), normals=False)
        assert np.array_equal(g.triangles(), refbind.read_stl(str(path)))


def _configs(tmp_path):
    yield os.path.join(GOLDEN, "reference_config.txt")
    p = tmp_path / "q.txt"
    p.write_text('parameters:\n  NumCycle: 2.9\n  FlowType: "Pulsatile"\n  stl_file: "a b.stl"\n'
                 "  time_resolution: 2.5e-6\n\ninlets:\n    plane: z\n    coordinate: 1.5\n"
                 "    normal: [0, 0, -1]\n   \noutlets:\n    plane: x\n    coordinate: 2\n"
                 "    range_y: [ 1 ,2,]\n    normal: [1, 0, 0] # c\n\nequations:\n  E1: [1, 2,3]\n")
    yield str(p)


def test_config_parser_equal_reference(tmp_path):
    for path in _configs(tmp_path):
        a, b = lbm.parse_config(path), refbind.parse_config(path)
        for k in ("stl_file", "flow_type", "dt", "rho", "nu_blood", "tau", "dx", "pressure", "reynolds_number",
                  "cte_vel", "systole", "diastole", "num_out", "num_cycle", "max_iter", "iter_to_write",
                  "n_inlets", "n_outlets", "n_equations"):
            assert getattr(a, k) == getattr(b, k), k
        for blk in ("inlets", "outlets"):
            for i in range(getattr(a, "n_" + blk)):
                x, y = getattr(a, blk)[i], getattr(b, blk)[i]
                assert x.plane == y.plane and x.normal_len == y.normal_len
                assert list(x.normal) == list(y.normal)
                for r in ("range_x", "range_y", "range_z"):
                    assert list(getattr(x, r)) == list(getattr(y, r))
                assert x.coordinate == y.coordinate or (np.isnan(x.coordinate) and y.n_inlets is None)


@pytest.mark.parametrize("normal", [(0, 0, -1), (0, 0, 1), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0),
                                    (1, 0, -1), (-0.5, 0.3, 0.8)])
def test_zou_he_equals_reference_bit_for_bit(normal):
    rng = np.random.default_rng(21)
    _, _, w = lbm.lattice(19)
    for _ in range(25):
        f = w * (1 + 0.1 * rng.uniform(-1, 1, 19))
        U = 0.08 * rng.uniform(-1, 1)
        # the -y branch divides by (1 - 1.0) (Boundary.cpp:206): inf/nan on both sides
        assert np.array_equal(oracle.zouhe(f, U, normal), refbind.zouhe(f, U, normal), equal_nan=True)


def _waveform_times():
    p = 0.9
    ts = list(np.linspace(0.0, 2.7, 40001))
    for k in range(4):
        for b in (0.0, 0.1 * p, 0.3 * p, 0.9 * p):
            x = k * p + b
            ts += [x, np.nextafter(x, -np.inf), np.nextafter(x, np.inf)]
    for dt in (1e-5, 2e-3, 0.01, 0.05):  # the reference's accumulated total_runtime (Simulation.cpp:598-600)
        t = 0.0
        while t <= 2.7:
            ts.append(t)
            t += dt
    return np.array(ts)


def test_inlet_waveform_equals_reference_bit_for_bit():
    # Pulsatile::InletVelequation1..3 compiled from the reference header, the
    # product's lbm_inlet_velocity (what lbm_step and reza::Simulation use)
    # and the oracle's restatement, on a dense sweep of [0, 2.7], every
    # branch bound and its neighbouring doubles, and the times the
    # reference's t += dt accumulation produces.
    ts = _waveform_times()
    seen = set()
    for t in ts:
        r = refbind.inlet_velocity(t, 0.9)
        assert lbm.inlet_velocity(t, 0.9) == r, t
        assert oracle.inlet_velocity(t, 0.9) == r, t
        x = np.fmod(t, 0.9)
        seen.add(0 if x < 0.09 else 1 if x < 0.27 else 2 if x < 0.81 else 3)
    assert seen == {0, 1, 2, 3}
    for period in (0.8, 1.0):  # the bounds scale with the period argument
        for t in np.linspace(0.0, 3.0, 3001):
            assert lbm.inlet_velocity(t, period) == refbind.inlet_velocity(t, period)


def test_inlet_waveform_pieces_meet_the_branch_bounds():
    # The selected piece at each bound is the one the reference's half-open
    # intervals pick: [0, 0.09) eq1, [0.09, 0.27) eq2, [0.27, 0.81) eq3,
    # [0.81, 0.9) zero.
    p = 0.9
    assert lbm.inlet_velocity(0.0, p) == refbind.pulsatile_piece(1, 0.0) == 0.1423
    for b, piece in ((0.1 * p, 2), (0.3 * p, 3)):
        assert lbm.inlet_velocity(b, p) == refbind.pulsatile_piece(piece, b)
        below = np.nextafter(b, 0.0)
        assert lbm.inlet_velocity(below, p) == refbind.pulsatile_piece(piece - 1, below)
    assert lbm.inlet_velocity(0.9 * p, p) == 0.0
    below = np.nextafter(0.9 * p, 0.0)
    assert lbm.inlet_velocity(below, p) == refbind.pulsatile_piece(3, below)
