"""The C-ABI library loads and exports every function include/lbm_c_api.h declares
(no compute calls: runs without a GPU)."""
import os
import re

import lbm
from conftest import ROOT


def _declared():
    text = open(os.path.join(ROOT, "include", "lbm_c_api.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lbm_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    L = lbm.lib()
    names = _declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_the_header():
    bound = {n for n, _, _ in lbm.SIGNATURES}
    assert set(_declared()) == bound


def test_errors_are_reported_through_the_abi():
    import pytest

    with pytest.raises(lbm.LbmError) as e:
        lbm.Geometry.from_config("/nonexistent/config.txt")
    assert e.value.code == lbm.ERR_IO
    assert "Unable to open" in str(e.value)


def test_invalid_solver_arguments_are_rejected_before_any_device_call():
    # These checks precede every HIP call, so they run without a GPU.
    import numpy as np
    import pytest

    import workloads

    empty = lbm.Geometry.from_labels(np.zeros((4, 4, 4), np.uint8), np.zeros(3), 1.0, normals=False)
    assert empty.n == 0
    with pytest.raises(lbm.LbmError) as e:
        lbm.Solver(empty, tau=0.8)
    assert e.value.code == lbm.ERR_ARG and "no nodes" in str(e.value)
    lab, origin, _ = workloads.tube_mask(4.0, 6)
    g = lbm.Geometry.from_labels(lab, origin, 1.0, periodic=(0, 0, 1), normals=False)
    for kw, code in ((dict(tau=0.5), lbm.ERR_ARG), (dict(tau=0.8, precision=7), lbm.ERR_ARG),
                     (dict(tau=0.8, wall_mode=5), lbm.ERR_ARG)):
        with pytest.raises(lbm.LbmError) as e:
            lbm.Solver(g, **kw)
        assert e.value.code == code, kw
    g27 = lbm.Geometry.from_labels(lab, origin, 1.0, periodic=(0, 0, 1), normals=False, q=27)
    assert g27.n == g.n
    # copy probe: sizes below one element, no iterations, or a grid beyond 2^31 blocks
    for nbytes, iters in ((8, 1), (1 << 20, 0), (1 << 47, 1)):
        with pytest.raises(lbm.LbmError) as e:
            lbm.copy_probe(nbytes, iters=iters)
        assert e.value.code == lbm.ERR_ARG, (nbytes, iters)


def test_geometry_edge_cases(tmp_path):
    import numpy as np
    import pytest

    # An STL with no facets is an input error (the reference reads an empty list).
    stl = tmp_path / "empty.stl"
    stl.write_text("solid empty\nendsolid empty\n")
    cfg = tmp_path / "config.txt"
    text = open(os.path.join(ROOT, "tests", "golden", "reference_config.txt")).read()
    assert '"JT.stl"' in text
    cfg.write_text(text.replace('"JT.stl"', '"empty.stl"'))
    with pytest.raises(lbm.LbmError) as e:
        lbm.Geometry.from_config(str(cfg))
    assert e.value.code == lbm.ERR_ARG and "no facets" in str(e.value)
    # One fluid column of a periodic box: every node is regular, none near a wall.
    lab = np.zeros((3, 3, 5), np.uint8)
    lab[1, 1, :] = 1
    g = lbm.Geometry.from_labels(lab, np.zeros(3), 1.0, periodic=(0, 0, 1), normals=False)
    assert g.n == 5 and g.info.n_inlet == 0 and g.info.n_outlet == 0
    nb = g.neighbours()
    c, _, _ = lbm.lattice(19)
    ijk, _, _ = g.nodes()
    for q in range(19):
        for n in range(5):
            x = ijk[n] + c[q]
            inside = x[0] == 1 and x[1] == 1
            assert (nb[q, n] >= 0) == bool(inside), (q, n)
    with pytest.raises(lbm.LbmError):
        lbm.Geometry.from_labels(np.zeros((0, 3, 3), np.uint8), np.zeros(3), 1.0, normals=False)
