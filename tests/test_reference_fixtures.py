"""The product against fixtures written from the reference's own compiled
code (tests/golden/make_reference_tables.py, make_zouhe_vectors.py), so the
comparisons also run where the reference sources are absent:

* the host set-up chain (labels, node kinds and order, outlet blocks, push
  table, near-wall set and normals) on the reference's JT.stl at grid 0.2
  and on two seeded blobs with two outlets;
* the same on the GPU: ray labels cast on the device (label_device) and
  near-wall normals summed on the device (normals = 2), bit for bit;
* lbm_geometry_from_tables fed the reference's tables (its push table, its
  nextToTheWall order and normals): the solver it makes steps bit for bit as
  the one built from the config, and reports the stress in the reference's
  near-wall order;
* Boundary::ApplyBc evaluated by the device code of the fused step (lbm_zouhe)
  against the compiled reference's, bit for bit.

The reference reads the uninitialised BoundaryType_ of plain fluid nodes
when it builds the near-wall set (Subdomain.cpp:483); where that memory
happened to read Wall, the compiled run dropped the node. The product treats
the field as unset, so the comparisons skip exactly the nodes that run's
fixture marks (fluid_bt == 2).
"""
import os
import sys

import numpy as np
import pytest

import lbm
from conftest import GOLDEN

sys.path.insert(0, GOLDEN)
import cases  # noqa: E402

TABLES = np.load(os.path.join(GOLDEN, "reference_tables.npz"))


def ref_tables(name):
    r = {k.split("/", 1)[1]: TABLES[k] for k in TABLES.files if k.startswith(name + "/")}
    shape = tuple(int(v) for v in r.pop("labels_shape"))
    r["labels"] = np.unpackbits(r["labels"])[: int(np.prod(shape))].reshape(shape)
    return r


def compare(g, r):
    ijk, kind, oid = g.nodes()
    assert tuple(g.info.grid) == tuple(r["grid"])
    assert g.info.spacing == float(r["spacing"])
    assert np.array_equal(g.labels(), r["labels"])
    assert np.array_equal(ijk, r["ijk"])
    assert np.array_equal(kind, r["kind"])
    assert np.array_equal(oid, r["outlet_id"])
    assert np.array_equal(g.neighbours(), r["nb"])
    nw, nrm = g.nearwall()
    keep = ~((kind == 0) & (r["fluid_bt"] == 2))[nw]
    nw, nrm = nw[keep], nrm[keep]
    assert len(nw) == len(r["nw_ijk"])
    pos = {tuple(p): i for i, p in enumerate(r["nw_ijk"])}
    idx = np.array([pos[tuple(ijk[n])] for n in nw], dtype=np.int64)
    assert np.array_equal(nrm, r["nw_normal"][idx])  # bit for bit, any order


def from_reference(g, r):
    """lbm_geometry_from_tables over the reference's tables."""
    ijk, _, _ = g.nodes()
    index = {tuple(p): n for n, p in enumerate(ijk)}
    nw = np.array([index[tuple(p)] for p in r["nw_ijk"]], np.int32)
    b = g.info.bbox
    return lbm.Geometry.from_tables(19, r["grid"], (b[0], b[2], b[4]), float(r["spacing"]), r["ijk"], r["kind"],
                                    r["nb"], outlet_id=r["outlet_id"], inlet_normal=r["inlet_normal"],
                                    outlet_normals=g.outlet_normals(), nearwall=nw, normals=r["nw_normal"]), nw


@pytest.mark.parametrize("name", sorted(cases.CASES))
def test_host_chain_equals_reference_fixture(tmp_path, name):
    compare(lbm.Geometry.from_config(cases.CASES[name](str(tmp_path))), ref_tables(name))


@pytest.mark.parametrize("name", sorted(cases.CASES))
def test_tables_entry_takes_the_reference_tables(tmp_path, name):
    g = lbm.Geometry.from_config(cases.CASES[name](str(tmp_path)))
    r = ref_tables(name)
    t, nw = from_reference(g, r)
    assert t.n == g.n and tuple(t.info.grid) == tuple(g.info.grid)
    for a, b in zip(t.nodes(), g.nodes()):
        assert np.array_equal(a, b)
    assert np.array_equal(t.neighbours(), g.neighbours())
    assert np.array_equal(t.bc_neighbour(), g.bc_neighbour())
    assert np.array_equal(t.labels(), g.labels())
    tnw, tnrm = t.nearwall()
    assert np.array_equal(tnw, nw) and np.array_equal(tnrm, r["nw_normal"])
    for k in ("n_wall", "n_inlet", "n_outlet", "upstream_out_of_box", "has_inlet_normal"):
        assert getattr(t.info, k) == getattr(g.info, k), k
    assert list(t.info.inlet_normal) == list(g.info.inlet_normal)


def test_tables_entry_refuses_inconsistent_tables(tmp_path):
    g = lbm.Geometry.from_config(cases.CASES["blob_4"](str(tmp_path)))
    ijk, kind, oid = g.nodes()
    nb = g.neighbours()
    nw, nrm = g.nearwall()
    b = g.info.bbox
    base = dict(outlet_id=oid, inlet_normal=g.info.inlet_normal, outlet_normals=g.outlet_normals(), nearwall=nw,
                normals=nrm)

    def make(**kw):
        args = dict(ijk=ijk, kind=kind, nb=nb)
        args.update({k: v for k, v in kw.items() if k in args})
        rest = dict(base)
        rest.update({k: v for k, v in kw.items() if k in rest})
        return lbm.Geometry.from_tables(19, g.info.grid, (b[0], b[2], b[4]), g.info.spacing, args["ijk"],
                                        args["kind"], args["nb"], **rest)

    make()  # the consistent tables pass
    bad_nb = nb.copy()
    i = int(np.flatnonzero(nb[0] >= 0)[0])
    bad_nb[0, i] = i  # the node itself, not the node at x + c_0
    dup = ijk.copy()
    dup[1] = dup[0]
    bad_kind = kind.copy()
    bad_kind[np.flatnonzero(kind == lbm.KIND_WALL)[0]] = lbm.KIND_OUTLET  # an outlet node without a block
    wall_nw = nw.copy()
    wall_nw[0] = np.flatnonzero(kind == lbm.KIND_WALL)[0]
    for kw in (dict(nb=bad_nb), dict(ijk=dup), dict(kind=bad_kind), dict(nearwall=wall_nw)):
        with pytest.raises(lbm.LbmError) as e:
            make(**kw)
        assert e.value.code == lbm.ERR_ARG


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(cases.CASES))
def test_device_labels_and_normals_equal_reference_fixture(tmp_path, name):
    cfg = cases.CASES[name](str(tmp_path))
    print(f"{name}: device labels and normals", flush=True)
    g = lbm.Geometry.from_config(cfg, label_device=True, normals=2)
    compare(g, ref_tables(name))


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["jt_0.2", "blob_4"])
def test_solver_on_the_reference_tables_equals_the_config_path(tmp_path, name):
    g = lbm.Geometry.from_config(cases.CASES[name](str(tmp_path)))
    r = ref_tables(name)
    t, nw = from_reference(g, r)
    print(f"{name}: {g.n} nodes, {len(nw)} reference near-wall nodes", flush=True)
    a = lbm.Solver(g, tau=0.8, dt=1e-5, cte_vel=0.02)
    b = lbm.Solver(t, tau=0.8, dt=1e-5, cte_vel=0.02)
    for k in (1, 30):
        a.step(k)
        b.step(k)
        assert np.array_equal(a.populations(), b.populations())
        assert np.array_equal(a.post_collision(), b.post_collision())
    gnw, _ = g.nearwall()
    row = {int(n): i for i, n in enumerate(gnw)}
    common = np.array([i for i, n in enumerate(nw) if int(n) in row])
    wa, acc_a, mag_a = a.wss()
    wb, acc_b, mag_b = b.wss()
    sel = np.array([row[int(nw[i])] for i in common])
    # the stress of a node does not depend on which other nodes are near-wall
    assert np.array_equal(wb[common], wa[sel])
    assert np.array_equal(acc_b[common], acc_a[sel])
    assert np.array_equal(mag_b[common], mag_a[sel])


@pytest.mark.gpu
def test_device_zou_he_equals_reference_vectors():
    z = np.load(os.path.join(GOLDEN, "zouhe_reference.npz"))
    for k, n in enumerate(z["normals"]):
        got = lbm.zouhe(z["f"][k], z["vel"][k], n)
        # the -y branch divides by 1 - 1.0 (Boundary.cpp:206): inf/nan on both sides
        assert np.array_equal(got, z["out"][k], equal_nan=True), f"normal {n}"
