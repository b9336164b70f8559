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
    tris = np.asarray(tris, dtype=np.float64).reshape(-1, 9)
    normals = np.asarray(normals, dtype=np.float64)
    facet = ("  facet normal %r %r %r\n    outer loop\n" + "      vertex %r %r %r\n" * 3 +
             "    endloop\n  endfacet\n")
    rows = np.concatenate([normals, tris], 1).tolist()
    with open(path, "w") as f:
        f.write(f"solid {name}\n")
        f.write("".join(facet % tuple(r) for r in rows))
        f.write(f"endsolid {name}\n")


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


def straight_tube(workdir, radius=10.0, length=80.0, tau=0.8, cte_vel=0.05, max_iter=2000, segments=64):
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


_AXES = {"x": (2, 0, 1), "y": (1, 2, 0), "z": (0, 1, 2)}  # new coordinate a takes old coordinate _AXES[ax][a]


def axis_tube(workdir, axis="z", inlet_end="min", radius=5.0, length=24.0, tau=0.8, cte_vel=0.02, max_iter=200,
              dt=1e-5, flow_type="Constant", segments=64):
    """Straight tube along `axis` (the z tube of tube_surface under a cyclic
    coordinate permutation, which keeps the triangles' outward orientation).
    The inlet is the cap at the `inlet_end` ("min" or "max") of the axis, with
    its outward normal, and the outlet the opposite cap, so the six
    (axis, end) pairs select each branch of the reference's Zou-He rule
    (Boundary.cpp:111-289) and of its outlet copy (Simulation.cpp:650-741)
    once. The end caps lie parallel to the voxeliser's +z rays for the x and y
    tubes; such triangles never count (ComplexGeometry.cpp:291). Returns the
    config path."""
    os.makedirs(workdir, exist_ok=True)
    perm = _AXES[axis]
    tris, nrm = tube_surface(radius, length, segments=segments)
    tris, nrm = tris[:, :, perm], nrm[:, perm]
    a = "xyz".index(axis)
    lo = [0.0, 0.0, 0.0]
    lo[a] = -1.0
    hi = [0.0, 0.0, 0.0]
    hi[a] = 1.0
    ends = {"min": dict(plane=axis, coordinate=0.0, normal=tuple(int(v) for v in lo)),
            "max": dict(plane=axis, coordinate=float(length), normal=tuple(int(v) for v in hi))}
    stl = os.path.join(workdir, f"tube_{axis}.stl")
    write_stl(stl, tris, nrm, f"tube_{axis}")
    cfg = os.path.join(workdir, "config.txt")
    write_config(cfg, os.path.basename(stl), tau=tau, cte_vel=cte_vel, max_iter=max_iter, dt=dt, flow_type=flow_type,
                 inlets=[ends[inlet_end]], outlets=[ends["max" if inlet_end == "min" else "min"]])
    return cfg


def revolution_surface(radius_of_z, z_levels, segments=128, center=(0.37, 0.21)):
    """Closed surface of revolution about the z axis through `center`:
    radius_of_z(z) sampled at the increasing z_levels, rims at half-segment
    angles (as tube_surface), fan caps at both ends. Outward normals from
    each triangle's vertices."""
    cx, cy = center
    z = np.asarray(z_levels, dtype=np.float64)
    th = 2.0 * np.pi * (np.arange(segments) + 0.5) / segments
    r = np.asarray([radius_of_z(v) for v in z], dtype=np.float64)
    P = np.stack([cx + r[:, None] * np.cos(th)[None, :], cy + r[:, None] * np.sin(th)[None, :],
                  np.repeat(z[:, None], segments, axis=1)], axis=-1)  # (levels, segments, 3)
    tris = []
    for i in range(len(z) - 1):
        for k in range(segments):
            a, b = P[i, k], P[i, (k + 1) % segments]
            c, d = P[i + 1, (k + 1) % segments], P[i + 1, k]
            tris += [[a, b, c], [a, c, d]]
    bot = np.array([cx, cy, z[0]])
    top = np.array([cx, cy, z[-1]])
    for k in range(segments):
        tris.append([bot, P[0, (k + 1) % segments], P[0, k]])
        tris.append([top, P[-1, k], P[-1, (k + 1) % segments]])
    tris = np.array(tris)
    nrm = _unit(np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0]))
    return tris, nrm


def stenosis_radius(radius=64.0, z_center=384.0, length=128.0, area_reduction=0.5):
    """Cosine stenosis: r(z) = R (1 - d (1 + cos(2 pi (z - zc) / L)) / 2) for
    |z - zc| <= L/2, R elsewhere, with d = 1 - sqrt(1 - area_reduction) so the
    throat area is (1 - area_reduction) of the tube's."""
    d = 1.0 - math.sqrt(1.0 - area_reduction)

    def r(z):
        if abs(z - z_center) <= 0.5 * length:
            return radius * (1.0 - 0.5 * d * (1.0 + math.cos(2.0 * math.pi * (z - z_center) / length)))
        return radius

    return r


def stenosed_tube(workdir, radius=64.0, length=1024.0, z_center=384.0, sten_length=128.0, area_reduction=0.5,
                segments=256, tau=0.55, max_iter=20000):
    """Stenosed tube: radius 64, length 1024, cosine stenosis of 50% area
    reduction centred at z = 384 over 128 lu (a 130 x 130 x 1025 grid at
    pitch 1). Inlet cap at z = 0 (normal -z), outlet cap at z = length
    (normal +z). Returns (config path, info)."""
    os.makedirs(workdir, exist_ok=True)
    r = stenosis_radius(radius, z_center, sten_length, area_reduction)
    z0, z1 = z_center - 0.5 * sten_length, z_center + 0.5 * sten_length
    levels = sorted(set([0.0, float(length)] + list(np.arange(0.0, z0, 16.0)) + list(np.arange(z0, z1 + 1e-9, 2.0)) +
                        list(np.arange(z1, length, 16.0))))
    tris, nrm = revolution_surface(r, levels, segments=segments)
    stl = os.path.join(workdir, "stenosis.stl")
    write_stl(stl, tris, nrm, "stenosis")
    cfg = os.path.join(workdir, "config.txt")
    write_config(cfg, "stenosis.stl", tau=tau, max_iter=max_iter,
                 inlets=[dict(plane="z", coordinate=0.0, normal=(0, 0, -1))],
                 outlets=[dict(plane="z", coordinate=float(length), normal=(0, 0, 1))])
    return cfg, dict(radius=radius, axis=(0.37, 0.21), n_triangles=len(tris), throat_radius=r(z_center))


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


# ----------------------------------------------------------------- meshing
# Kuhn split of a cube (corner index = x + 2y + 4z) into six tetrahedra that
# share the 0-7 diagonal; neighbouring cubes split shared faces the same way,
# so the extracted surface is closed.
_KUHN = np.array([[0, 1, 3, 7], [0, 3, 2, 7], [0, 2, 6, 7], [0, 6, 4, 7], [0, 4, 5, 7], [0, 5, 1, 7]])


def marching_tetrahedra(func, lo, hi, h, offset=(0.2113, 0.3571, 0.1237), chunk=48):
    """Zero level set of func(x, y, z) (negative inside) on a grid of pitch h
    covering [lo, hi]. Edge points are interpolated from the endpoint with the
    smaller global index, so a shared edge gives bit-identical vertices in
    every tetrahedron that meets it. Returns (tris[T,3,3], normals[T,3])."""
    lo = np.asarray(lo, np.float64) - h + np.asarray(offset) * h
    n = np.ceil((np.asarray(hi, np.float64) + h - lo) / h).astype(int) + 1
    nx, ny, nz = (int(v) for v in n)
    xs = lo[0] + np.arange(nx) * h
    ys = lo[1] + np.arange(ny) * h
    zs = lo[2] + np.arange(nz) * h
    out_t, out_n = [], []
    for i0 in range(0, nx - 1, chunk):
        i1 = min(nx, i0 + chunk + 1)
        X, Y, Z = np.meshgrid(xs[i0:i1], ys, zs, indexing="ij")
        F = func(X, Y, Z)
        F[F == 0.0] = 1e-12
        pos = F > 0
        cx, cy, cz = F.shape[0] - 1, ny - 1, nz - 1
        corners = [pos[a:a + cx, b:b + cy, c:c + cz] for c in (0, 1) for b in (0, 1) for a in (0, 1)]
        anyp = np.zeros_like(corners[0])
        allp = np.ones_like(corners[0])
        for cc in corners:
            anyp |= cc
            allp &= cc
        act = np.argwhere(anyp & ~allp)
        if len(act) == 0:
            continue
        off = np.array([[a, b, c] for c in (0, 1) for b in (0, 1) for a in (0, 1)])
        for tet in _KUHN:
            vi = act[:, None, :] + off[tet][None, :, :]          # local (i, j, k) of the 4 corners
            gi = (vi[..., 0] + i0) * (ny * nz) + vi[..., 1] * nz + vi[..., 2]
            val = F[vi[..., 0], vi[..., 1], vi[..., 2]]
            pt = np.stack([xs[vi[..., 0] + i0], ys[vi[..., 1]], zs[vi[..., 2]]], -1)
            sp = val > 0
            npos = sp.sum(1)

            def cut(a, b, sel):
                ga, gb = gi[sel, a], gi[sel, b]
                sw = ga > gb
                A = np.where(sw[:, None], pt[sel, b], pt[sel, a])
                B = np.where(sw[:, None], pt[sel, a], pt[sel, b])
                fa = np.where(sw, val[sel, b], val[sel, a])
                fb = np.where(sw, val[sel, a], val[sel, b])
                t = fa / (fa - fb)
                return A + t[:, None] * (B - A)

            grad_sel = []
            for k in (1, 3):
                sel = np.where(npos == k)[0]
                if len(sel) == 0:
                    continue
                lone_is_pos = (k == 1)
                order = np.argsort(~sp[sel] if lone_is_pos else sp[sel], axis=1, kind="stable")
                L, o1, o2, o3 = (order[:, m] for m in range(4))
                p1 = _cut_idx(cut, sel, L, o1)
                p2 = _cut_idx(cut, sel, L, o2)
                p3 = _cut_idx(cut, sel, L, o3)
                grad_sel.append((np.stack([p1, p2, p3], 1), _grad(pt[sel], sp[sel])))
            sel = np.where(npos == 2)[0]
            if len(sel):
                order = np.argsort(~sp[sel], axis=1, kind="stable")
                P1, P2, N1, N2 = (order[:, m] for m in range(4))
                a = _cut_idx(cut, sel, P1, N1)
                b = _cut_idx(cut, sel, P1, N2)
                c = _cut_idx(cut, sel, P2, N2)
                d = _cut_idx(cut, sel, P2, N1)
                g = _grad(pt[sel], sp[sel])
                grad_sel.append((np.stack([a, b, c], 1), g))
                grad_sel.append((np.stack([a, c, d], 1), g))
            for tri, g in grad_sel:
                nrm = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
                flip = (nrm * g).sum(1) < 0
                tri[flip] = tri[flip][:, [0, 2, 1]]
                nrm[flip] = -nrm[flip]
                ln = np.linalg.norm(nrm, axis=1)
                keep = ln > 1e-12
                out_t.append(tri[keep])
                out_n.append(nrm[keep] / ln[keep, None])
    return np.concatenate(out_t), np.concatenate(out_n)


def _cut_idx(cut, sel, a, b):
    # per-row endpoints a[i], b[i] (tetra-local corner numbers)
    out = np.empty((len(sel), 3))
    for ca in range(4):
        for cb in range(4):
            m = (a == ca) & (b == cb)
            if m.any():
                out[m] = cut(ca, cb, sel[m])
    return out


def _grad(pt, sp):
    w = sp.astype(np.float64)
    npos = w.sum(1, keepdims=True)
    pc = (pt * w[..., None]).sum(1) / npos
    nc = (pt * (1 - w)[..., None]).sum(1) / (4 - npos)
    return pc - nc


def _segment_distance(X, Y, Z, p0, p1):
    d = np.asarray(p1, np.float64) - np.asarray(p0, np.float64)
    L = np.linalg.norm(d)
    d = d / L
    rx, ry, rz = X - p0[0], Y - p0[1], Z - p0[2]
    s = rx * d[0] + ry * d[1] + rz * d[2]
    sc = np.clip(s, 0.0, L)
    qx, qy, qz = rx - sc * d[0], ry - sc * d[1], rz - sc * d[2]
    return np.sqrt(qx * qx + qy * qy + qz * qz), s, (qx, qy, qz), d


def bifurcation_function(parent_radius=24.0, daughter_radius=18.0, half_angle_deg=35.0, parent_length=600.0,
                         daughter_length=900.0, amplitude=2.0, seed=1):
    """Implicit Y-bifurcation (negative inside): parent along +z from z = 0,
    two daughters at +-half_angle in the xz plane, cut by z = 0 and by a
    horizontal plane through the daughters; radii carry a seeded
    three-mode sinusoidal perturbation with total amplitude <= `amplitude`."""
    rng = np.random.default_rng(seed)
    amps = rng.uniform(0.2, 1.0, 3)
    amps *= amplitude / amps.sum()
    kz = rng.uniform(2 * np.pi / 400, 2 * np.pi / 80, 3)
    nth = rng.integers(0, 4, 3)
    ph = rng.uniform(0, 2 * np.pi, 3)
    a = np.deg2rad(half_angle_deg)
    P0 = np.array([0.0, 0.0, -20.0])
    P1 = np.array([0.0, 0.0, parent_length])
    D = [P1 + daughter_length * np.array([np.sin(a), 0.0, np.cos(a)]),
         P1 + daughter_length * np.array([-np.sin(a), 0.0, np.cos(a)])]
    z_top = parent_length + (daughter_length - daughter_radius - 4.0) * np.cos(a)

    def radius(R0, s, q):
        th = np.arctan2(q[1], q[0])
        r = np.full_like(s, R0)
        for m in range(3):
            r = r + amps[m] * np.sin(kz[m] * s + nth[m] * th + ph[m])
        return r

    def F(X, Y, Z):
        dist, s, q, _ = _segment_distance(X, Y, Z, P0, P1)
        f = dist - radius(parent_radius, s, q)
        for Pd in D:
            dd, sd, qd, _ = _segment_distance(X, Y, Z, P1, Pd)
            f = np.minimum(f, dd - radius(daughter_radius, sd + parent_length, qd))
        f = np.maximum(f, -Z)
        f = np.maximum(f, Z - z_top)
        return f

    ext = daughter_length * np.sin(a) + daughter_radius + amplitude + 2
    lo = (-ext, -(parent_radius + amplitude + 2), 0.0)
    hi = (ext, parent_radius + amplitude + 2, z_top)
    return F, lo, hi, dict(z_top=z_top, parent_axis=(0.0, 0.0), parent_radius=parent_radius,
                           branch_point=tuple(P1), daughter_dirs=[tuple((Pd - P1) / daughter_length) for Pd in D],
                           daughter_radius=daughter_radius)


def y_bifurcation(workdir, mesh_h=3.5, **kw):
    """Precision-study Y-bifurcation (parent r 24, daughters r 18 at 35
    degrees, 3-mode wall perturbation of amplitude 2, seed 1). Inlet: the z = 0
    cap (normal -z); outlet: the top cut plane (normal +z), set on the last
    grid layer inside the cap. Returns (config path, info dict)."""
    os.makedirs(workdir, exist_ok=True)
    F, lo, hi, info = bifurcation_function(**kw)
    tris, nrm = marching_tetrahedra(F, lo, hi, mesh_h)
    stl = os.path.join(workdir, "bifurcation.stl")
    write_stl(stl, tris, nrm, "bifurcation")
    zmin, zmax = float(tris[..., 2].min()), float(tris[..., 2].max())
    k_top = math.floor((zmax - zmin) / 1.0)
    z_out = zmin + k_top * 1.0
    cfg = os.path.join(workdir, "config.txt")
    write_config(cfg, "bifurcation.stl", tau=0.67, max_iter=8000,
                 inlets=[dict(plane="z", coordinate=zmin, normal=(0, 0, -1))],
                 outlets=[dict(plane="z", coordinate=z_out, normal=(0, 0, 1))])
    info.update(n_triangles=len(tris), z_in=zmin, z_out=z_out)
    return cfg, info


def zernike_basis(rho, phi, n_modes=16):
    """Zernike polynomials Z_n^m in OSA/ANSI order (j = 0, 1, ...) on the unit
    disk: radial part R_n^|m|(rho) times cos(m phi) for m >= 0, sin(|m| phi)
    for m < 0. Returns (n_modes, len(rho))."""
    out = []
    n = 0
    while len(out) < n_modes:
        for m in range(-n, n + 1, 2):
            if len(out) == n_modes:
                break
            am = abs(m)
            R = np.zeros_like(rho)
            for k in range((n - am) // 2 + 1):
                c = (-1) ** k * math.factorial(n - k) / (
                    math.factorial(k) * math.factorial((n + am) // 2 - k) * math.factorial((n - am) // 2 - k))
                R = R + c * rho ** (n - 2 * k)
            out.append(R * (np.cos(m * phi) if m >= 0 else np.sin(am * phi)))
        n += 1
    return np.stack(out)


def branch_segment_mask(points, origin, direction, s0, length, radius):
    """Points within `radius` of the axis origin + s * direction with
    s0 <= s < s0 + length (a wall segment of one branch)."""
    d = np.asarray(direction, float)
    d = d / np.linalg.norm(d)
    rel = np.asarray(points, float) - np.asarray(origin, float)
    s = rel @ d
    dist = np.linalg.norm(rel - s[:, None] * d[None, :], axis=1)
    return (s >= s0) & (s < s0 + length) & (dist <= radius)


# ------------------------------------------------------ large single-GPU share
def _catmull_rom(P, samples_per_seg):
    """Uniform Catmull-Rom curve through the rows of P (phantom end points
    reflected), sampled samples_per_seg times per segment; returns points."""
    P = np.asarray(P, dtype=np.float64)
    Q = np.vstack([2 * P[0] - P[1], P, 2 * P[-1] - P[-2]])
    out = []
    t = np.linspace(0.0, 1.0, samples_per_seg, endpoint=False)[:, None]
    for i in range(1, len(Q) - 2):
        p0, p1, p2, p3 = Q[i - 1], Q[i], Q[i + 1], Q[i + 2]
        out.append(0.5 * ((2 * p1) + (-p0 + p2) * t + (2 * p0 - 5 * p1 + 4 * p2 - p3) * t * t +
                          (-p0 + 3 * p1 - 3 * p2 + p3) * t * t * t))
    out.append(P[-1][None, :])
    return np.vstack(out)


def _vessel_shape(scale, seed, sample_step):
    """Centreline samples C (every sample_step lu of arc), radii R, the bulge
    centre and radius, the walk's control points and the arc length of the
    tortuous vessel (tortuous_vessel describes the shape)."""
    rng = np.random.default_rng(seed)
    step, rmin, rmax, rb = 80.0 * scale, 20.0 * scale, 40.0 * scale, 60.0 * scale
    P = [np.zeros(3), np.array([0.0, 0.0, step])]
    d = np.array([0.0, 0.0, 1.0])
    for _ in range(9):
        d = _unit(d + 0.45 * rng.normal(size=3))
        d[2] = max(d[2], 0.55)
        d = _unit(d)
        P.append(P[-1] + step * d)
    P.append(P[-1] + np.array([0.0, 0.0, step]))
    P = np.array(P)
    dense = _catmull_rom(P, 400)
    seg = np.linalg.norm(np.diff(dense, axis=0), axis=1)
    s_dense = np.concatenate([[0.0], np.cumsum(seg)])
    length = float(s_dense[-1])
    s = np.arange(0.0, length, sample_step)
    s = np.append(s, length)
    C = np.stack([np.interp(s, s_dense, dense[:, a]) for a in range(3)], axis=1)
    phase = rng.uniform(0, 2 * np.pi)
    R = rmin + (rmax - rmin) * (0.5 + 0.5 * np.sin(2 * np.pi * 1.5 * s / length + phase))
    # Bulge: centre off the axis, perpendicular to the tangent, so that the
    # sphere overlaps the wall (a neck) and bulges out of the side.
    kb = int(np.searchsorted(s, 0.55 * length))
    tb = _unit(C[min(kb + 1, len(C) - 1)] - C[max(kb - 1, 0)])
    nb = _unit(np.cross(tb, rng.normal(size=3)))
    cb = C[kb] + nb * (R[kb] + 0.35 * rb)
    return C, R, cb, rb, P, length


def vessel_function(C, R, cb, rb, margin, band=8, device="cpu"):
    """Implicit function of the tortuous vessel (negative inside): the union
    of the balls (C[k], R[k]) and the bulge ball, cut by the planes through
    the end points' z. Evaluated in float64 with torch on `device`, on
    meshgrid slabs (indexing "ij") band by band of z levels, each against the
    balls within `margin` lu of the band and of the slab's x range; a point
    with no ball in reach gets a large positive value (it lies further than
    `margin` from the surface)."""
    import torch

    z0, z1 = float(C[0, 2]), float(C[-1, 2])
    big = 1e6
    dev = torch.device(device)
    Ct = torch.tensor(C, dtype=torch.float64, device=dev)
    Rt = torch.tensor(R, dtype=torch.float64, device=dev)

    def F(X, Y, Z):
        xs, zs = X[:, 0, 0], Z[0, 0, :]
        xa, xb = float(xs.min()), float(xs.max())
        nearx = (C[:, 0] + R + margin >= xa) & (C[:, 0] - R - margin <= xb)
        Xt, Yt, Zt = (torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in (X, Y, Z))
        out = torch.empty_like(Xt)
        for k0 in range(0, len(zs), band):
            k1 = min(len(zs), k0 + band)
            za, zb = float(zs[k0:k1].min()), float(zs[k0:k1].max())
            sel = np.where(nearx & (C[:, 2] + R + margin >= za) & (C[:, 2] - R - margin <= zb))[0]
            Xb, Yb, Zb = Xt[:, :, k0:k1], Yt[:, :, k0:k1], Zt[:, :, k0:k1]
            f = torch.sqrt((Xb - cb[0]) ** 2 + (Yb - cb[1]) ** 2 + (Zb - cb[2]) ** 2) - rb
            for c0 in range(0, len(sel), 32):
                idx = torch.as_tensor(sel[c0:c0 + 32], device=dev)
                c, r = Ct[idx], Rt[idx]
                d = torch.sqrt((Xb[..., None] - c[:, 0]) ** 2 + (Yb[..., None] - c[:, 1]) ** 2 +
                               (Zb[..., None] - c[:, 2]) ** 2) - r
                f = torch.minimum(f, d.amin(-1))
            f = torch.clamp(f, max=big)
            f = torch.maximum(f, z0 - Zb)
            out[:, :, k0:k1] = torch.maximum(f, Zb - z1)
        return out.cpu().numpy()

    return F


def tortuous_vessel_stl(workdir, scale=4.3, seed=5, sample_step=2.0, mesh_h=8.0, tau=0.8, max_iter=1000,
                        device="cpu"):
    """The tortuous vessel (tortuous_vessel) as a closed STL surface and a
    config.txt, for the set-up chain from the surface: the zero level set of
    vessel_function (on `device`) by marching tetrahedra at pitch mesh_h,
    written with 17 digits; grid_size 0.5 (a lattice pitch of 1: the reader doubles it). Inlet: the z min cap (normal -z); outlet: the last
    grid layer inside the z max cap (normal +z). Returns (config path, info)."""
    os.makedirs(workdir, exist_ok=True)
    C, R, cb, rb, _, length = _vessel_shape(scale, seed, sample_step)
    lo = np.minimum((C - R[:, None]).min(0), cb - rb)
    hi = np.maximum((C + R[:, None]).max(0), cb + rb)
    tris, nrm = marching_tetrahedra(vessel_function(C, R, cb, rb, 3.0 * mesh_h, device=device), lo, hi, mesh_h)
    stl = os.path.join(workdir, "vessel.stl")
    write_stl(stl, tris, nrm, "vessel")
    zmin, zmax = float(tris[..., 2].min()), float(tris[..., 2].max())
    z_out = zmin + math.floor(zmax - zmin)
    cfg = os.path.join(workdir, "config.txt")
    write_config(cfg, "vessel.stl", tau=tau, max_iter=max_iter, grid_size=0.5,
                 inlets=[dict(plane="z", coordinate=zmin, normal=(0, 0, -1))],
                 outlets=[dict(plane="z", coordinate=z_out, normal=(0, 0, 1))])
    return cfg, dict(scale=scale, length=length, centreline=C, radius=R, bulge_center=cb, bulge_radius=rb,
                     n_triangles=len(tris), z_in=zmin, z_out=z_out)


def tortuous_vessel(scale=4.5, seed=5, device=None, sample_step=2.0, slab=32):
    """Tortuous vessel of the large single-GPU share: a Catmull-Rom
    centreline through a 12-point random walk (step 80 scale lu, a persistent
    walk heading +z whose first and last steps run along z), radius varying
    between 20 and 40 scale lu along the arc, and a spherical saccular bulge
    of radius 60 scale lu on the side of the vessel at 55% of its length.
    scale = 4.5 gives about 300M nodes at pitch 1 (the 80/20-40/60 lu sizes
    as such give about 3.3M).

    Labels are the union of balls of radius r(s) centred on the centreline
    sampled every sample_step lu (a swept tube), and the bulge ball, on the
    lattice x = origin + (i, j, k); the box is cut at the end points' z, so
    the vessel's ends lie on the z faces (inlet at z min, normal -z; outlet
    at z max, normal +z). Computed slab by slab with torch on `device`
    (default: the GPU when there is one). Returns (fluid uint8 [nx, ny, nz],
    origin, info); info holds the triangles of a swept-ring surface mesh and
    the bulge sphere (for the near-wall normals), the plane blocks and sizes."""
    import torch

    if device is None:
        device = "cuda" if torch.cuda.is_available() else "cpu"
    C, R, cb, rb, P, length = _vessel_shape(scale, seed, sample_step)
    z0, z1 = float(C[0, 2]), float(C[-1, 2])
    lo = np.minimum((C - R[:, None]).min(0), cb - rb) - 2.0
    hi = np.maximum((C + R[:, None]).max(0), cb + rb) + 2.0
    origin = np.array([np.floor(lo[0]) + 0.37, np.floor(lo[1]) + 0.21, z0])
    nx = int(np.ceil(hi[0] - origin[0])) + 1
    ny = int(np.ceil(hi[1] - origin[1])) + 1
    nz = int(np.floor(z1 - z0)) + 1
    fluid = np.zeros((nx, ny, nz), np.uint8)
    dev = torch.device(device)
    Ct = torch.tensor(C, dtype=torch.float32, device=dev)
    Rt = torch.tensor(R, dtype=torch.float32, device=dev)
    cbt = torch.tensor(cb, dtype=torch.float32, device=dev)
    for k0 in range(0, nz, slab):
        k1 = min(nz, k0 + slab)
        za, zb = origin[2] + k0, origin[2] + k1 - 1
        sel = np.where((C[:, 2] + R >= za) & (C[:, 2] - R <= zb))[0]
        if sel.size:
            blo = (C[sel, :2] - R[sel, None]).min(0)
            bhi = (C[sel, :2] + R[sel, None]).max(0)
        else:
            blo, bhi = None, None
        if cb[2] + rb >= za and cb[2] - rb <= zb:
            blo = cb[:2] - rb if blo is None else np.minimum(blo, cb[:2] - rb)
            bhi = cb[:2] + rb if bhi is None else np.maximum(bhi, cb[:2] + rb)
        if blo is None:
            continue
        i0 = max(0, int(np.floor(blo[0] - origin[0]))); i1 = min(nx, int(np.ceil(bhi[0] - origin[0])) + 1)
        j0 = max(0, int(np.floor(blo[1] - origin[1]))); j1 = min(ny, int(np.ceil(bhi[1] - origin[1])) + 1)
        xs = torch.arange(i0, i1, device=dev, dtype=torch.float32) + float(origin[0])
        ys = torch.arange(j0, j1, device=dev, dtype=torch.float32) + float(origin[1])
        zs = torch.arange(k0, k1, device=dev, dtype=torch.float32) + float(origin[2])
        X, Y, Z = torch.meshgrid(xs, ys, zs, indexing="ij")
        pts = torch.stack([X.reshape(-1), Y.reshape(-1), Z.reshape(-1)], dim=1)
        inside = ((pts - cbt) ** 2).sum(1) < rb * rb
        if sel.size:
            cs, rs = Ct[torch.as_tensor(sel, device=dev)], Rt[torch.as_tensor(sel, device=dev)]
            r2 = (rs * rs)[None, :]
            step_pts = 1 << 18
            for a in range(0, pts.shape[0], step_pts):
                # exact squared distances (not the matrix-product form): the same labels on every run
                d2 = ((pts[a:a + step_pts, None, :] - cs[None, :, :]) ** 2).sum(-1)
                inside[a:a + step_pts] |= (d2 < r2).any(dim=1)
        fluid[i0:i1, j0:j1, k0:k1] = inside.reshape(i1 - i0, j1 - j0, k1 - k0).to(torch.uint8).cpu().numpy()
    # Surface mesh for the near-wall normals: rings around the centreline
    # (parallel-transported frames, outward triangle order) and the bulge sphere.
    ring_every, nseg = max(1, int(round(8.0 / sample_step))), 64
    idx = np.arange(0, len(C), ring_every)
    T = np.gradient(C, axis=0)
    T = T / np.linalg.norm(T, axis=1)[:, None]
    u = _unit(np.cross(T[0], [1.0, 0.0, 0.0]) if abs(T[0, 0]) < 0.9 else np.cross(T[0], [0.0, 1.0, 0.0]))
    rings = []
    for k in idx:
        t = T[k]
        u = _unit(u - np.dot(u, t) * t)
        v = np.cross(t, u)
        th = 2 * np.pi * (np.arange(nseg) + 0.5) / nseg
        rings.append(C[k] + R[k] * (np.cos(th)[:, None] * u + np.sin(th)[:, None] * v))
    rings = np.array(rings)
    a = rings[:-1, :, :]
    b = np.roll(rings[:-1], -1, axis=1)
    c = rings[1:, :, :]
    dd = np.roll(rings[1:], -1, axis=1)
    tube = np.concatenate([np.stack([a, b, c], axis=2).reshape(-1, 3, 3), np.stack([b, dd, c], axis=2).reshape(-1, 3, 3)])
    sph, _ = icosphere(cb, rb, subdiv=3, rotation_seed=seed)
    tris = np.concatenate([tube, sph])
    info = dict(scale=scale, length=length, centreline=C, radius=R, bulge_center=cb, bulge_radius=rb,
                control_points=P, grid=(nx, ny, nz), z_range=(z0, z1), triangles=tris.reshape(-1, 9),
                n_triangles=len(tris), inlet_normal=(0.0, 0.0, -1.0), outlet_normal=(0.0, 0.0, 1.0))
    return fluid, origin, info
