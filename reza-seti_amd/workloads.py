"""Synthetic vessel workloads: ASCII STL surfaces and config.txt files.

The reference ships one clinical STL (JT.stl) and no test inputs, so the
workloads are generated: each writes an STL in the line format
Reader::readSTLASCII accepts (GeometryReader.cpp:74-158) and a config.txt in
the format parseInputFile reads (LBM.h:81-188). Coordinates are in lattice
units with grid_size 0.5, i.e. a lattice pitch of 1 (the reference doubles
grid_size, ComplexGeometry.cpp:157).

Every surface is closed and oriented outward, because the voxeliser counts
+z ray crossings. Surfaces are placed so that no grid column passes through a
triangle edge or vertex (the reference's inclusive inside test counts such a
crossing twice).
"""
import math
import os

import numpy as np


def _num(v):
    return repr(float(v))


def write_stl(path, tris, normals, name="vessel"):
    """tris: (T, 3, 3) vertices, normals: (T, 3). Written with 17 significant
    digits so that the reader recovers the exact doubles."""
    tris = np.asarray(tris, dtype=np.float64)
    normals = np.asarray(normals, dtype=np.float64)
    lines = [f"solid {name}"]
    for t, n in zip(tris, normals):
        lines.append(f"  facet normal {_num(n[0])} {_num(n[1])} {_num(n[2])}")
        lines.append("    outer loop")
        for v in t:
            lines.append(f"      vertex {_num(v[0])} {_num(v[1])} {_num(v[2])}")
        lines.append("    endloop")
        lines.append("  endfacet")
    lines.append(f"endsolid {name}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def _unit(v):
    n = np.linalg.norm(v, axis=-1, keepdims=True)
    return v / np.where(n > 0, n, 1.0)


def tube_surface(radius, length, segments=64, center=(0.37, 0.21), z0=0.0):
    """Closed cylinder along z with fan caps. Rim vertices sit at half-segment
    angles so the bounding box edge is not a rim vertex and no grid column
    meets the cap centre or a spoke."""
    cx, cy = center
    th = 2.0 * np.pi * (np.arange(segments) + 0.5) / segments
    rim = np.stack([cx + radius * np.cos(th), cy + radius * np.sin(th)], axis=1)
    z1 = z0 + length
    tris, nrm = [], []
    for k in range(segments):
        a, b = rim[k], rim[(k + 1) % segments]
        pa0, pb0 = [a[0], a[1], z0], [b[0], b[1], z0]
        pa1, pb1 = [a[0], a[1], z1], [b[0], b[1], z1]
        mid = 0.5 * (th[k] + th[k] + 2 * np.pi / segments)
        side_n = [math.cos(mid), math.sin(mid), 0.0]
        tris += [[pa0, pb0, pb1], [pa0, pb1, pa1]]
        nrm += [side_n, side_n]
        tris.append([[cx, cy, z1], pa1, pb1])
        nrm.append([0.0, 0.0, 1.0])
        tris.append([[cx, cy, z0], pb0, pa0])
        nrm.append([0.0, 0.0, -1.0])
    return np.array(tris), np.array(nrm)


def write_config(path, stl_file, *, tau, cte_vel=0.05, max_iter=2000, dt=1e-5, grid_size=0.5,
                 flow_type="Constant", inlets=(), outlets=(), iter_to_write=100000):
    """inlets/outlets: dicts with plane, coordinate, normal and optional ranges."""
    def block(b):
        out = [f"    plane: {b['plane']}", f"    coordinate: {_num(b['coordinate'])}"]
        for key in ("range_x", "range_y", "range_z"):
            if key in b:
                out.append(f"    {key}: [{_num(b[key][0])}, {_num(b[key][1])}]")
        n = b["normal"]
        out.append(f"    normal: [{n[0]}, {n[1]}, {n[2]}]")
        return "\n".join(out)

    text = ["parameters:", "",
            f'  stl_file: "{stl_file}"', f'  FlowType: "{flow_type}"',
            f"  time_resolution: {dt!r}", "  density: 1060", "  kinematic_viscosity: 3.5e-6",
            f"  relaxation_factor: {tau!r}", f"  grid_size: {grid_size!r}", "  pressure: 1",
            f"  CteVel: {cte_vel!r}", "  reynolds_number: 100", f"  number_of_outlets: {len(outlets)}",
            f"  MaxIter: {max_iter}", f"  IterToWrite: {iter_to_write}", "  Systole: 0.15",
            "  Diastole: 0.9", "  period: 0.9", "  NumCycle: 5", "", "", "inlets:", ""]
    for b in inlets:
        text += [block(b), ""]
    text += ["outlets:", ""]
    for b in outlets:
        text += [block(b), ""]
    with open(path, "w") as f:
        f.write("\n".join(text) + "\n")


def c0_tube(workdir, radius=10.0, length=80.0, tau=0.8, cte_vel=0.05, max_iter=2000, segments=64):
    """Straight tube: 64-segment cylinder, radius 10,
    length 80, inlet cap at z = 0 (normal -z, reference Zou-He), outlet cap at
    z = length (normal +z, reference copy rule). Returns the config path."""
    os.makedirs(workdir, exist_ok=True)
    tris, nrm = tube_surface(radius, length, segments=segments)
    stl = os.path.join(workdir, "tube.stl")
    write_stl(stl, tris, nrm, "tube")
    cfg = os.path.join(workdir, "config.txt")
    write_config(cfg, "tube.stl", tau=tau, cte_vel=cte_vel, max_iter=max_iter,
                 inlets=[dict(plane="z", coordinate=0.0, normal=(0, 0, -1))],
                 outlets=[dict(plane="z", coordinate=float(length), normal=(0, 0, 1))])
    return cfg


def tube_mask(radius, length, pad=1, center_offset=(0.37, 0.21)):
    """Fluid labels of a z-periodic tube sampled on the lattice directly (for
    the Poiseuille check): point (i, j) is fluid when its distance to the axis
    is below radius. Returns (labels[nx, ny, nz], origin, axis_xy)."""
    n = int(math.ceil(2 * radius)) + 2 * pad + 1
    ax = (n - 1) / 2.0 + center_offset[0] * 0.0
    x = np.arange(n, dtype=np.float64)
    X, Y = np.meshgrid(x, x, indexing="ij")
    r2 = (X - ax) ** 2 + (Y - ax) ** 2
    disc = r2 < radius * radius
    lab = np.repeat(disc[:, :, None], int(length), axis=2).astype(np.uint8)
    return lab, (0.0, 0.0, 0.0), (ax, ax)


def icosphere(center, radius, subdiv=2, rotation_seed=None, radial=None):
    """Closed triangulated sphere (outward normals). `radial(dirs) -> radii`
    deforms it into a star-shaped blob. A seeded random rotation keeps edges
    off grid-aligned positions."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t],
                  [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
         [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11],
         [6, 2, 10], [8, 6, 7], [9, 8, 1]]
    v = _unit(v)
    verts = list(v)
    for _ in range(subdiv):
        cache, nf = {}, []

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        f = nf
    d = np.array(verts)
    if rotation_seed is not None:
        q, _ = np.linalg.qr(np.random.default_rng(rotation_seed).normal(size=(3, 3)))
        d = d @ q.T
    r = radius * (radial(d) if radial is not None else np.ones(len(d)))
    pts = np.asarray(center, dtype=np.float64)[None, :] + d * r[:, None]
    tris = pts[np.array(f)]
    n = _unit(np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0]))
    return tris, n
