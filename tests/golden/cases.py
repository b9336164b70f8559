"""Inputs shared by the fixture script and the tests: the reference's own
vessel (JT.stl, kept here gzip-compressed, with its config.txt) at a chosen
grid size, and seeded blobs with two outlets."""
import gzip
import os
import shutil

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def jt_case(workdir, grid_size):
    """The reference's config.txt and JT.stl in `workdir`, grid_size replaced.
    Returns the config path."""
    os.makedirs(workdir, exist_ok=True)
    with gzip.open(os.path.join(HERE, "JT.stl.gz"), "rb") as src, open(os.path.join(workdir, "JT.stl"), "wb") as dst:
        shutil.copyfileobj(src, dst)
    text = open(os.path.join(HERE, "reference_config.txt")).read()
    text = text.replace("grid_size: 0.025", f"grid_size: {grid_size!r}")
    cfg = os.path.join(workdir, "config.txt")
    with open(cfg, "w") as f:
        f.write(text)
    return cfg


def blob_case(workdir, seed):
    """A star-shaped blob (seeded icosphere deformation) with an x inlet and
    two overlapping z outlet blocks, pitch 0.8. Returns the config path."""
    import workloads

    os.makedirs(workdir, exist_ok=True)
    rng = np.random.default_rng(seed)
    a = rng.normal(size=3) * 0.12

    def radial(d):
        return 1.0 + a[0] * d[:, 0] * d[:, 2] + a[1] * d[:, 1] ** 2 + a[2] * d[:, 0]

    tris, nrm = workloads.icosphere((0.11, 0.07, -0.3), 6.1, subdiv=3, rotation_seed=seed, radial=radial)
    workloads.write_stl(os.path.join(workdir, "s.stl"), tris, nrm)
    zmin = tris[..., 2].min()
    cfg = os.path.join(workdir, "config.txt")
    workloads.write_config(cfg, "s.stl", tau=0.8, grid_size=0.4,
                           inlets=[dict(plane="x", coordinate=float(tris[..., 0].min()), normal=(-1, 0, 0))],
                           outlets=[dict(plane="z", coordinate=float(zmin), normal=(0, 0, -1), range_x=(-100.0, 0.5)),
                                    dict(plane="z", coordinate=float(zmin), normal=(0, 0, -1))])
    return cfg


CASES = {"jt_0.2": lambda d: jt_case(d, 0.2), "blob_4": lambda d: blob_case(d, 4), "blob_5": lambda d: blob_case(d, 5)}
