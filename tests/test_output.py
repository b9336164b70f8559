"""Output files of Simulation::WriteOutput (OutputFile.cpp): the fluid
variables VTU and the smoothing of the near-wall CSV, checked on the host
against direct restatements of the reference loops (the VTK writer itself is
not in the image; the VTU is parsed back here)."""
import math
import re

import numpy as np

import lbm


def read_vtu(path):
    raw = open(path, "rb").read()
    head, _, rest = raw.partition(b"<AppendedData encoding=\"raw\">")
    head = head.decode()
    data = rest[rest.index(b"_") + 1:]
    piece = re.search(r'NumberOfPoints="(\d+)" NumberOfCells="(\d+)"', head)
    arrays = {}
    for m in re.finditer(r'<DataArray type="(\w+)"(?: Name="(\w+)")?(?: NumberOfComponents="(\d+)")? '
                         r'format="appended" offset="(\d+)"/>', head):
        typ, name, comps, off = m.group(1), m.group(2) or "Points", int(m.group(3) or 1), int(m.group(4))
        nbytes = int(np.frombuffer(data[off:off + 8], np.uint64)[0])
        dt = {"Float64": np.float64, "Int64": np.int64, "UInt8": np.uint8}[typ]
        a = np.frombuffer(data[off + 8:off + 8 + nbytes], dt)
        arrays[name] = a.reshape(-1, comps) if comps > 1 else a
    return int(piece.group(1)), int(piece.group(2)), arrays


def test_vtu_matches_the_reference_writer_loops(tmp_path, tube_geometry):
    g = tube_geometry
    rng = np.random.default_rng(2)
    rho = 1.0 + 0.01 * rng.standard_normal(g.n)
    u = 0.01 * rng.standard_normal((g.n, 3))
    path = tmp_path / "0.vtu"
    g.write_vtu(rho, u, path)
    npts, ncells, a = read_vtu(path)
    ijk, _, _ = g.nodes()
    # Points: the nodes in the writer's z-y-x loop order, at their grid indices.
    order = np.lexsort((ijk[:, 0], ijk[:, 1], ijk[:, 2]))
    assert npts == g.n
    assert np.array_equal(a["Points"], ijk[order].astype(np.float64))
    assert np.array_equal(a["Velocity"], u[order])
    assert np.array_equal(a["Density"], rho[order])
    assert np.array_equal(a["VelocityMagnitude"], np.sqrt((u[order] ** 2).sum(1)))
    # Cells: every cube x < NX-2, y < NY-2, z < NZ-2 whose 8 corners are nodes,
    # corners in VTK hexahedron order, in z-y-x loop order.
    NX, NY, NZ = g.info.grid
    pid = -np.ones((NX, NY, NZ), np.int64)
    pid[ijk[order, 0], ijk[order, 1], ijk[order, 2]] = np.arange(g.n)
    corners = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    cells = []
    for z in range(NZ - 2):
        for y in range(NY - 2):
            for x in range(NX - 2):
                ids = [pid[x + dx, y + dy, z + dz] for dx, dy, dz in corners]
                if min(ids) >= 0:
                    cells.append(ids)
    cells = np.array(cells, np.int64)
    assert ncells == len(cells) > 0
    assert np.array_equal(a["connectivity"], cells.reshape(-1))
    assert np.array_equal(a["offsets"], 8 * np.arange(1, len(cells) + 1))
    assert np.all(a["types"] == 12)


def _smooth_reference(rows, radius, pulsatile):
    # OutputFile::smoothData / calculateMeanNeighboringValues (OutputFile.cpp:
    # 321-393), a full scan in row order.
    out = []
    for p in rows:
        acc = [0.0, 0.0, 0.0]
        cnt = 0
        for q in rows:
            d = math.sqrt(math.pow(q[0] - p[0], 2) + math.pow(q[1] - p[1], 2) + math.pow(q[2] - p[2], 2))
            if d > 0 and d <= radius:
                acc[0] += q[3]
                if pulsatile:
                    acc[1] += q[4]
                    acc[2] += q[5]
                cnt += 1
        r = list(p)
        if cnt:
            r[3] = acc[0] / cnt
            if pulsatile:
                r[4], r[5] = acc[1] / cnt, acc[2] / cnt
        out.append(r)
    return out


def _fmt(v):
    return "%g" % v  # the default ostream format of a double


def test_smoothing_matches_a_full_scan(tmp_path):
    rng = np.random.default_rng(3)
    for pulsatile in (False, True):
        n = 300
        # Points on a 0.1 grid (radius 0.25 reaches 2 spacings), rounded as printed.
        pts = np.round(rng.integers(0, 12, (n, 3)) * 0.1, 6)
        pts = np.unique(pts, axis=0)
        vals = rng.uniform(0, 2, (len(pts), 4))
        cols = "X,Y,Z,WSS" + (",OSI,TAWSS" if pulsatile else "") + ",VelMag"
        src = tmp_path / f"raw{int(pulsatile)}.csv"
        with open(src, "w") as fh:
            fh.write(cols + "\n")
            for p, v in zip(pts, vals):
                row = list(p) + ([v[0], v[1], v[2]] if pulsatile else [v[0]]) + [v[3]]
                fh.write(",".join(_fmt(x) for x in row) + "\n")
        dst = tmp_path / f"smooth{int(pulsatile)}.csv"
        lbm.smooth_surface_csv(src, dst, 0.25, pulsatile)
        rows = [[float(x) for x in line.split(",")] for line in open(src).read().splitlines()[1:]]
        want = _smooth_reference(rows, 0.25, pulsatile)
        got_lines = open(dst).read().splitlines()
        assert got_lines[0] == cols
        assert got_lines[1:] == [",".join(_fmt(x) for x in r) for r in want]
