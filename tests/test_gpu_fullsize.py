"""Full-size workloads on the GPU, through size-independent properties:
FP32-shifted storage tracks FP64 within the precision target (relative L2
error of the wall shear stress and the velocity at most 1e-4), on the
2.85M-node bifurcation (D3Q19) and the 12.7M-node stenosis (D3Q27)."""
import os
import sys
import tempfile

import numpy as np
import pytest

import lbm
import workloads

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402

pytestmark = pytest.mark.gpu

TOL = 1e-4  # relative L2, FP32-shifted vs FP64


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_bifurcation_fp32_tracks_fp64():
    with tempfile.TemporaryDirectory() as wd:
        geom, basis, _ = bench.build_geometry(wd)
    out = {}
    for prec in (lbm.PREC_F64, lbm.PREC_F32):
        s, _, _, _, _ = bench.run_precision(geom, basis, prec, 1000, 0)
        out[prec] = (s.wss()[0], s.macroscopic()[1])
        s.close()
    ew = rel(out[lbm.PREC_F32][0], out[lbm.PREC_F64][0])
    eu = rel(out[lbm.PREC_F32][1], out[lbm.PREC_F64][1])
    print(f"{geom.n} nodes, 1000 steps: WSS rel L2 {ew:.2e}, velocity rel L2 {eu:.2e}", flush=True)
    assert ew <= TOL and eu <= TOL


def test_stenosis_d3q27_fp32_tracks_fp64():
    with tempfile.TemporaryDirectory() as wd:
        cfg, info = workloads.stenosed_tube(wd, tau=bench.STENOSIS_TAU)
        geom = lbm.Geometry.from_config(cfg, q=27, normals=2, label_device=True)
    ijk, kind, _ = geom.nodes()
    b, h = geom.info.bbox, geom.info.spacing
    inl = np.where(kind == lbm.KIND_INLET)[0]
    x = b[0] + ijk[inl, 0] * h - info["axis"][0]
    y = b[2] + ijk[inl, 1] * h - info["axis"][1]
    prof = np.clip(1.0 - (x * x + y * y) / info["radius"] ** 2, 0.0, None)
    prof /= prof.mean()
    u_mean = 200.0 * (bench.STENOSIS_TAU - 0.5) / 3.0 / (2.0 * info["radius"])
    out = {}
    for prec in (lbm.PREC_F64, lbm.PREC_F32):
        s = lbm.Solver(geom, precision=prec, wall_mode=lbm.WALL_HALFWAY_BB, inlet_mode=lbm.INLET_NEE_VELOCITY,
                       outlet_mode=lbm.OUTLET_NEE_PRESSURE, tau=bench.STENOSIS_TAU, dt=1e-5, parity_checks=0, wss_every=1)
        s.set_inlet_profile(prof[None, :])
        s.step_coefficients(np.full((300, 1), u_mean))
        out[prec] = (s.wss()[0], s.macroscopic()[1])
        s.close()
    ew = rel(out[lbm.PREC_F32][0], out[lbm.PREC_F64][0])
    eu = rel(out[lbm.PREC_F32][1], out[lbm.PREC_F64][1])
    print(f"{geom.n} nodes, 300 steps: WSS rel L2 {ew:.2e}, velocity rel L2 {eu:.2e}", flush=True)
    assert ew <= TOL and eu <= TOL
