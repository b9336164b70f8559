"""Writes zouhe_reference.npz: Boundary::ApplyBc of the reference's own
compiled Boundary class (oracle/_ref/libref.so, built from the reference
sources by oracle/ref/Makefile) on seeded D3Q19 population vectors, for the
six axis normals and two oblique ones (every branch, alone and combined).

    python tests/golden/make_zouhe_vectors.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle")]

import oracle  # noqa: E402
import refbind  # noqa: E402

NORMALS = [(0, 0, -1), (0, 0, 1), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (1, 0, -1), (-0.5, 0.3, 0.8)]


def main():
    rng = np.random.default_rng(21)
    _, _, w = oracle.lattice(19)
    f_in, vel, out = [], [], []
    for n in NORMALS:
        f = w[None, :] * (1 + 0.1 * rng.uniform(-1, 1, (32, 19)))
        u = 0.08 * rng.uniform(-1, 1, 32)
        f_in.append(f)
        vel.append(u)
        out.append(np.stack([refbind.zouhe(f[i], u[i], n) for i in range(32)]))
    np.savez_compressed(os.path.join(HERE, "zouhe_reference.npz"), normals=np.array(NORMALS, np.float64),
                        f=np.stack(f_in), vel=np.stack(vel), out=np.stack(out))


if __name__ == "__main__":
    main()
