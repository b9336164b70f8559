"""Ray-parity voxeliser on the GPU (Geometry::labelPoints,
ComplexGeometry.cpp:440-500): the same +z ray test as the host voxeliser,
which the compiled reference pins (tests/test_reference_build.py), over the
same candidate triangles; the labels and every table built from them must be
identical."""
import time

import numpy as np
import pytest

import lbm
import workloads

pytestmark = pytest.mark.gpu


def _both(cfg, **kw):
    t0 = time.perf_counter()
    host = lbm.Geometry.from_config(cfg, normals=False, **kw)
    t1 = time.perf_counter()
    dev = lbm.Geometry.from_config(cfg, normals=False, label_device=True, **kw)
    t2 = time.perf_counter()
    print(f"{host.info.grid[0]}x{host.info.grid[1]}x{host.info.grid[2]} points: host build {t1 - t0:.2f}s, "
          f"device labels {t2 - t1:.2f}s", flush=True)
    return host, dev


def _same(host, dev):
    assert np.array_equal(host.labels(), dev.labels())
    assert host.n == dev.n
    a, b = host.nodes(), dev.nodes()
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert np.array_equal(host.neighbours(), dev.neighbours())


def test_tube_labels_equal_host(tube_config):
    _same(*_both(tube_config))


def test_stenosis_labels_equal_host(tmp_path):
    cfg, _ = workloads.stenosed_tube(str(tmp_path), radius=24.0, length=256.0, z_center=96.0, sten_length=48.0)
    _same(*_both(cfg, q=27))


def test_bifurcation_labels_equal_host(tmp_path):
    cfg, _ = workloads.y_bifurcation(str(tmp_path), mesh_h=4.0, parent_radius=12.0, daughter_radius=9.0,
                                      parent_length=120.0, daughter_length=160.0)
    host, dev = _both(cfg)
    _same(host, dev)
    assert host.n > 10000
