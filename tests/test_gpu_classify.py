"""Node classification and connectivity on the GPU (classify_device): the
boundary points (Geometry::findBoundaryPoints, ComplexGeometry.cpp:517-564),
node kinds with the plane tests of Geometry::getInOutBP
(ComplexGeometry.cpp:567-662) and CreateNodes' precedence
(Subdomain.cpp:364-442), the node order (Subdomain.cpp:598-611) and the push
table (SubDomain::Connectivity, Subdomain.cpp:447-471) against the host
chain, which the compiled reference pins (tests/test_reference_fixtures.py):
every table, count and near-wall normal identical, bit for bit, on meshes
(tubes along each axis, the reference's JT.stl, blobs with two outlets, a
stenosis in D3Q27, a bifurcation) and on seeded random labels with every
combination of periodic axes and blocks that cut the box."""
import time

import numpy as np
import pytest

import lbm
import workloads
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

import sys  # noqa: E402

sys.path.insert(0, GOLDEN)
import cases  # noqa: E402

INFO = ("n_nodes", "n_fluid", "n_wall", "n_inlet", "n_outlet", "n_nearwall", "n_boundary_points",
        "upstream_out_of_box", "q", "n_outlets", "has_inlet_normal", "spacing")


def same(host, dev):
    for k in INFO:
        assert getattr(host.info, k) == getattr(dev.info, k), k
    assert list(host.info.grid) == list(dev.info.grid)
    assert list(host.info.inlet_normal) == list(dev.info.inlet_normal)
    assert np.array_equal(host.labels(), dev.labels())
    for a, b in zip(host.nodes(), dev.nodes()):
        assert np.array_equal(a, b)
    assert np.array_equal(host.neighbours(), dev.neighbours())
    assert np.array_equal(host.bc_neighbour(), dev.bc_neighbour())
    assert np.array_equal(host.outlet_normals(), dev.outlet_normals())
    hn, hr = host.nearwall()
    dn, dr = dev.nearwall()
    assert np.array_equal(hn, dn) and np.array_equal(hr, dr)


def from_config_both(cfg, **kw):
    t0 = time.perf_counter()
    host = lbm.Geometry.from_config(cfg, **kw)
    t1 = time.perf_counter()
    dev = lbm.Geometry.from_config(cfg, classify_device=True, **kw)
    t2 = time.perf_counter()
    print(f"{host.n} nodes: host {t1 - t0:.2f}s, device classification {t2 - t1:.2f}s", flush=True)
    return host, dev


@pytest.mark.parametrize("axis,end", [(a, e) for a in "xyz" for e in ("min", "max")])
def test_axis_tubes(tmp_path, axis, end):
    same(*from_config_both(workloads.axis_tube(str(tmp_path), axis=axis, inlet_end=end)))


@pytest.mark.parametrize("name", sorted(cases.CASES))
def test_reference_cases(tmp_path, name):
    same(*from_config_both(cases.CASES[name](str(tmp_path))))


def test_stenosis_d3q27(tmp_path):
    cfg, _ = workloads.stenosed_tube(str(tmp_path), radius=24.0, length=256.0, z_center=96.0, sten_length=48.0)
    same(*from_config_both(cfg, q=27))


def test_bifurcation(tmp_path):
    cfg, _ = workloads.y_bifurcation(str(tmp_path), mesh_h=4.0, parent_radius=12.0, daughter_radius=9.0,
                                      parent_length=120.0, daughter_length=160.0)
    host, dev = from_config_both(cfg)
    same(host, dev)
    assert host.n > 10000


def labels_both(fluid, origin, spacing, inlets, outlets, **kw):
    host = lbm.Geometry.from_labels(fluid, origin, spacing, inlets=inlets, outlets=outlets, **kw)
    dev = lbm.Geometry.from_labels(fluid, origin, spacing, inlets=inlets, outlets=outlets, classify_device=True, **kw)
    return host, dev


@pytest.mark.parametrize("periodic", [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 1, 1)])
@pytest.mark.parametrize("q", [19, 27])
def test_random_labels(periodic, q):
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(sum(periodic) * 10 + q)
    shape = (37, 29, 300)  # columns longer than one 256-point chunk
    field = gaussian_filter(rng.standard_normal(shape), 2.0, mode="wrap")
    fluid = (field > np.quantile(field, 0.35)).astype(np.uint8)
    origin, h = (-1.25, 0.5, 3.0), 0.5
    # an inlet on the z = min face; outlets on the x faces, the first cut to
    # part of the y range, the second facing inwards (the node behind it lies
    # outside the box unless x is periodic), the third overlapping the first
    inl = [lbm.plane_block("z", origin[2], (0.0, 0.0, -1.0))]
    outs = [lbm.plane_block("x", origin[0], (-1.0, 0.0, 0.0), range_y=(origin[1], origin[1] + 7.0)),
            lbm.plane_block("x", origin[0] + (shape[0] - 1) * h, (-1.0, 0.0, 0.0)),
            lbm.plane_block("x", origin[0], (-1.0, 0.0, 0.0))]
    host, dev = labels_both(fluid, origin, h, inl, outs, q=q, periodic=periodic, normals=False)
    same(host, dev)
    i = host.info
    assert i.n_inlet > 0 and i.n_outlet > 0 and 0 < i.n_boundary_points < i.n_nodes
    assert (i.upstream_out_of_box > 0) == (periodic[0] == 0)


def test_periodic_tube_and_edge_cases():
    fluid, origin, _ = workloads.tube_mask(6.0, 40)
    host, dev = labels_both(fluid, origin, 1.0, [], [], periodic=(0, 0, 1), normals=False)
    same(host, dev)
    assert host.info.n_boundary_points < host.n  # the periodic ends are no boundary points
    one = np.zeros((5, 4, 3), np.uint8)
    one[2, 1, 1] = 1
    same(*labels_both(one, (0.0, 0.0, 0.0), 1.0, [], [], normals=False))
    same(*labels_both(np.zeros((3, 3, 3), np.uint8), (0.0, 0.0, 0.0), 1.0, [], [], normals=False))
    full = np.ones((4, 5, 6), np.uint8)
    same(*labels_both(full, (0.0, 0.0, 0.0), 1.0, [], [], periodic=(1, 1, 1), normals=False))
