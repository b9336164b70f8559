"""Writes reference_tables.npz: the node tables of the reference's own
compiled Geometry + SubDomain (oracle/_ref/libref.so, built from the
reference sources by oracle/ref/Makefile) for the cases of cases.py, so that
tests on a machine without the reference sources can compare against them.

Per case: grid, spacing, bbox, labels (bit-packed), ijk, kind, outlet_id,
the push table nb, the raw BoundaryType_ of plain fluid nodes (fluid_bt),
the near-wall nodes' positions in nextToTheWall order (nw_ijk), their normals
and the inlet normal.

    python tests/golden/make_reference_tables.py
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [HERE, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "reza-seti_amd")]

import cases  # noqa: E402
import refbind  # noqa: E402


def main():
    out = {}
    for name, make in cases.CASES.items():
        with tempfile.TemporaryDirectory() as d:
            r = refbind.build(make(d))
        for k, v in r.items():
            if k == "labels":
                out[f"{name}/labels_shape"] = np.array(v.shape, np.int64)
                v = np.packbits(v.reshape(-1))
            out[f"{name}/{k}"] = np.asarray(v)
        print(f"{name}: grid {r['grid']}, {len(r['kind'])} nodes, {len(r['nw_ijk'])} near-wall", flush=True)
    np.savez_compressed(os.path.join(HERE, "reference_tables.npz"), **out)


if __name__ == "__main__":
    main()
