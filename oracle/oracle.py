"""ORACLE binding — test infrastructure only.

ctypes wrapper of oracle/liboracle.so (a plain-C restatement of the
reference solver, see oracle.h). Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may import this module, and only as the checker
or the CPU baseline; the product never calls it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_P = C.c_void_p
_lib = None


class Setup(C.Structure):
    _fields_ = [("Q", C.c_int32), ("N", C.c_int64), ("grid", C.c_int32 * 3), ("periodic", C.c_int32 * 3),
                ("ijk", _P), ("kind", _P), ("outlet_id", _P), ("n_outlets", C.c_int32),
                ("outlet_normals", _P), ("inlet_normal", C.c_double * 3), ("n_nearwall", C.c_int64),
                ("nearwall", _P), ("normals", _P), ("tau", C.c_double), ("wss_C", C.c_double),
                ("wall_mode", C.c_int32), ("inlet_mode", C.c_int32), ("outlet_mode", C.c_int32),
                ("outlet_rho", C.c_double), ("force", C.c_double * 3), ("inlet_profile", _P)]


class Triangle(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("x1", "y1", "z1", "x2", "y2", "z2", "x3", "y3", "z3", "nx", "ny", "nz")]


class Block(C.Structure):
    _fields_ = [("plane", C.c_char * 16), ("coordinate", C.c_double), ("range_x", C.c_double * 2),
                ("range_y", C.c_double * 2), ("range_z", C.c_double * 2), ("normal", C.c_double * 3)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing; build with `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        sig = {
            "oracle_lattice": ([C.c_int, _P, _P, _P], C.c_int),
            "oracle_create": ([_P], _P),
            "oracle_destroy": ([_P], None),
            "oracle_reset": ([_P], None),
            "oracle_set_f": ([_P, _P], None),
            "oracle_get_f": ([_P, _P], None),
            "oracle_get_fs": ([_P, _P], None),
            "oracle_get_feq": ([_P, _P], None),
            "oracle_get_macro": ([_P, _P, _P], None),
            "oracle_get_wss": ([_P, _P, _P, _P], None),
            "oracle_step": ([_P, C.c_int32, _P, C.c_int32], C.c_int),
            "oracle_collide": ([_P], None),
            "oracle_zouhe": ([_P, C.c_double, _P], None),
            "oracle_inlet_velocity": ([C.c_double, C.c_double], C.c_double),
            "oracle_wss_tau": ([C.c_int, _P, _P, _P, C.c_double, _P], None),
            "oracle_grid": ([_P, C.c_int64, C.c_double, _P, _P, _P], None),
            "oracle_label_points": ([_P, C.c_int64, _P, _P, C.c_double, _P, C.c_int32], None),
            "oracle_classify": ([_P, _P, _P, C.c_double, _P, C.c_int32, _P, C.c_int32, C.c_int32, _P, _P, _P, _P],
                                C.c_int64),
            "oracle_normal": ([_P, C.c_int64, C.c_double, C.c_double, C.c_double, _P], None),
        }
        for k, (a, r) in sig.items():
            f = getattr(L, k)
            f.argtypes = a
            f.restype = r
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def lattice(q):
    c = np.zeros((q, 3), np.int32)
    opp = np.zeros(q, np.int32)
    w = np.zeros(q)
    assert lib().oracle_lattice(q, _p(c), _p(opp), _p(w)) == 0
    return c, opp, w


def zouhe(f, vel, normal):
    f = np.array(f, dtype=np.float64)
    n = np.ascontiguousarray(normal, dtype=np.float64)
    lib().oracle_zouhe(_p(f), float(vel), _p(n))
    return f


def wss_tau(f, feq, normal, C_wss):
    f = np.ascontiguousarray(f, np.float64)
    feq = np.ascontiguousarray(feq, np.float64)
    n = np.ascontiguousarray(normal, np.float64)
    tau = np.zeros(3)
    lib().oracle_wss_tau(len(f), _p(f), _p(feq), _p(n), float(C_wss), _p(tau))
    return tau


def inlet_velocity(x, period=0.9):
    return lib().oracle_inlet_velocity(float(x), float(period))


def wss_constant(dt, spacing):
    """Simulation.cpp:768-771."""
    phys_den = 1060.0
    dtdt = dt * dt
    dxdx = spacing * spacing * 1e-6
    return phys_den * dtdt / dxdx


class Reference:
    """Reference solver state on node tables given in reference order."""

    def __init__(self, q, grid, ijk, kind, outlet_id, outlet_normals, inlet_normal, nearwall, normals, tau,
                 wss_C, wall_mode=0, inlet_mode=0, outlet_mode=0, outlet_rho=1.0, force=(0, 0, 0),
                 periodic=(0, 0, 0), inlet_profile=None):
        self._keep = [np.ascontiguousarray(ijk, np.int32), np.ascontiguousarray(kind, np.uint8),
                      np.ascontiguousarray(outlet_id, np.int32),
                      np.ascontiguousarray(outlet_normals, np.float64).reshape(-1),
                      np.ascontiguousarray(nearwall, np.int32), np.ascontiguousarray(normals, np.float64)]
        s = Setup()
        s.Q = q
        s.N = len(kind)
        for a in range(3):
            s.grid[a] = int(grid[a])
            s.periodic[a] = int(periodic[a])
            s.inlet_normal[a] = float(inlet_normal[a])
            s.force[a] = float(force[a])
        s.ijk, s.kind, s.outlet_id, s.outlet_normals, s.nearwall, s.normals = [_p(a) for a in self._keep]
        s.n_outlets = len(self._keep[3]) // 3
        s.n_nearwall = len(self._keep[4])
        s.tau = tau
        s.wss_C = wss_C
        s.wall_mode, s.inlet_mode, s.outlet_mode = wall_mode, inlet_mode, outlet_mode
        s.outlet_rho = outlet_rho
        if inlet_profile is not None:
            self._keep.append(np.ascontiguousarray(inlet_profile, np.float64))
            s.inlet_profile = _p(self._keep[-1])
        self.q, self.n, self.nnw = q, len(kind), len(self._keep[4])
        self._h = C.c_void_p(lib().oracle_create(C.byref(s)))

    @classmethod
    def from_geometry(cls, geom, tau, dt, **kw):
        ijk, kind, oid = geom.nodes()
        nw, nrm = geom.nearwall()
        return cls(geom.q, tuple(geom.info.grid), ijk, kind, oid, geom.outlet_normals(),
                   tuple(geom.info.inlet_normal), nw, nrm, tau, wss_constant(dt, geom.info.spacing), **kw)

    def step(self, u_in, nthreads=1):
        u = np.ascontiguousarray(np.atleast_1d(u_in), np.float64)
        lib().oracle_step(self._h, len(u), _p(u), int(nthreads))

    def collide(self):
        lib().oracle_collide(self._h)

    def reset(self):
        lib().oracle_reset(self._h)

    def f(self):
        out = np.zeros((self.n, self.q))
        lib().oracle_get_f(self._h, _p(out))
        return out

    def fs(self):
        out = np.zeros((self.n, self.q))
        lib().oracle_get_fs(self._h, _p(out))
        return out

    def feq(self):
        out = np.zeros((self.n, self.q))
        lib().oracle_get_feq(self._h, _p(out))
        return out

    def macro(self):
        """(rho[N], u[N, 3]) set by the last collision."""
        rho, u = np.zeros(self.n), np.zeros((self.n, 3))
        lib().oracle_get_macro(self._h, _p(rho), _p(u))
        return rho, u

    def set_f(self, f):
        f = np.ascontiguousarray(f, np.float64)
        lib().oracle_set_f(self._h, _p(f))

    def wss(self):
        w = np.zeros(self.nnw)
        acc = np.zeros((self.nnw, 3))
        mag = np.zeros(self.nnw)
        lib().oracle_get_wss(self._h, _p(w), _p(acc), _p(mag))
        return w, acc, mag

    def __del__(self):
        try:
            if self._h:
                lib().oracle_destroy(self._h)
                self._h = None
        except Exception:
            pass


def brute_force_geometry(tris, normals, grid_size, inlets, outlets, q=19, nthreads=1):
    """labels + classification + push table, every point against every triangle."""
    T = len(tris)
    arr = (Triangle * max(T, 1))()
    for i in range(T):
        v = np.asarray(tris[i], np.float64).reshape(9)
        n = np.asarray(normals[i], np.float64)
        arr[i] = Triangle(*(list(v) + list(n)))
    bbox = np.zeros(6)
    dims = np.zeros(3, np.int32)
    sp = C.c_double()
    lib().oracle_grid(arr, T, grid_size, _p(bbox), _p(dims), C.byref(sp))
    fluid = np.zeros(tuple(dims), np.uint8)
    lib().oracle_label_points(arr, T, _p(bbox), _p(dims), sp.value, _p(fluid), int(nthreads))

    def blocks(lst):
        out = (Block * max(1, len(lst)))()
        inf = float("inf")
        for i, b in enumerate(lst):
            out[i].plane = b["plane"].encode()
            out[i].coordinate = b["coordinate"]
            for key in ("range_x", "range_y", "range_z"):
                r = b.get(key, (-inf, inf))
                getattr(out[i], key)[0], getattr(out[i], key)[1] = r
            for a in range(3):
                out[i].normal[a] = b["normal"][a]
        return out

    N = int(fluid.sum())
    ijk = np.zeros((N, 3), np.int32)
    kind = np.zeros(N, np.uint8)
    oid = np.zeros(N, np.int32)
    nb = np.zeros((q, N), np.int32)
    n2 = lib().oracle_classify(_p(fluid), _p(dims), _p(bbox), sp.value, blocks(inlets), len(inlets),
                               blocks(outlets), len(outlets), q, _p(ijk), _p(kind), _p(oid), _p(nb))
    assert n2 == N
    return dict(fluid=fluid, bbox=bbox, dims=dims, spacing=sp.value, ijk=ijk, kind=kind, outlet_id=oid, nb=nb,
                triangles=arr, ntri=T)


def normal_of(geom_bf, x, y, z):
    out = np.zeros(3)
    lib().oracle_normal(geom_bf["triangles"], geom_bf["ntri"], x, y, z, _p(out))
    return out
