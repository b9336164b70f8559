"""ORACLE binding to oracle/_ref/libref.so — test infrastructure only.

libref.so is the reference's own set-up and boundary code (compiled from
/root/reference by oracle/ref/Makefile) behind a small C harness. Tests use it
to pin the product's voxelisation, node classification, connectivity,
near-wall normals, Zou-He rule, STL reader and config parser against the real
implementation. It exists only where the reference sources are present.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_ref", "libref.so")
_P = C.c_void_p
_lib = None


class Tables(C.Structure):
    _fields_ = [("grid", C.c_int32 * 3), ("spacing", C.c_double), ("bbox", C.c_double * 6),
                ("labels", C.POINTER(C.c_uint8)), ("n", C.c_int64), ("ijk", C.POINTER(C.c_int32)),
                ("kind", C.POINTER(C.c_uint8)), ("outlet_id", C.POINTER(C.c_int32)),
                ("nb", C.POINTER(C.c_int32)), ("fluid_bt", C.POINTER(C.c_int32)), ("n_nw", C.c_int64),
                ("nw_ijk", C.POINTER(C.c_int32)), ("nw_normal", C.POINTER(C.c_double)),
                ("inlet_normal", C.c_double * 3)]


class Block(C.Structure):
    _fields_ = [("plane", C.c_char * 16), ("coordinate", C.c_double), ("range_x", C.c_double * 2),
                ("range_y", C.c_double * 2), ("range_z", C.c_double * 2), ("normal", C.c_double * 3),
                ("normal_len", C.c_int32)]


class Config(C.Structure):
    _fields_ = [("stl_file", C.c_char * 512), ("flow_type", C.c_char * 32)] + \
               [(k, C.c_double) for k in ("dt", "rho", "nu_blood", "tau", "dx", "pressure", "reynolds_number",
                                          "cte_vel", "systole", "diastole")] + \
               [(k, C.c_int32) for k in ("num_out", "num_cycle", "max_iter", "iter_to_write", "n_inlets",
                                         "n_outlets", "n_equations")] + \
               [("inlets", Block * 16), ("outlets", Block * 16)]


def available():
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(LIB_PATH)
        L.ref_build.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(Tables)]
        L.ref_build.restype = C.c_int
        L.ref_free.argtypes = [C.POINTER(Tables)]
        L.ref_free.restype = None
        L.ref_zouhe.argtypes = [C.c_double, _P, _P]
        L.ref_zouhe.restype = C.c_int
        L.ref_read_stl.argtypes = [C.c_char_p, _P, C.c_int64]
        L.ref_read_stl.restype = C.c_int64
        L.ref_parse_config.argtypes = [C.c_char_p, C.POINTER(Config)]
        L.ref_parse_config.restype = C.c_int
        L.ref_lattice.argtypes = [_P, _P, _P]
        L.ref_lattice.restype = C.c_int
        L.ref_inlet_velocity.argtypes = [C.c_double, C.c_double]
        L.ref_inlet_velocity.restype = C.c_double
        L.ref_pulsatile_piece.argtypes = [C.c_int, C.c_double]
        L.ref_pulsatile_piece.restype = C.c_double
        _lib = L
    return _lib


def _arr(ptr, n, dtype):
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True) if n else np.zeros(0, dtype)


def build(config_path):
    """Run the reference Geometry + SubDomain constructors on a config."""
    d, name = os.path.split(os.path.abspath(config_path))
    t = Tables()
    rc = lib().ref_build(d.encode(), name.encode(), C.byref(t))
    if rc != 0:
        raise RuntimeError(f"reference build failed ({rc})")
    n = t.n
    g = tuple(t.grid)
    out = dict(grid=g, spacing=t.spacing, bbox=np.array(t.bbox),
               labels=_arr(t.labels, g[0] * g[1] * g[2], np.uint8).reshape(g),
               ijk=_arr(t.ijk, 3 * n, np.int32).reshape(n, 3), kind=_arr(t.kind, n, np.uint8),
               outlet_id=_arr(t.outlet_id, n, np.int32), nb=_arr(t.nb, 19 * n, np.int32).reshape(19, n),
               fluid_bt=_arr(t.fluid_bt, n, np.int32), nw_ijk=_arr(t.nw_ijk, 3 * t.n_nw, np.int32).reshape(-1, 3),
               nw_normal=_arr(t.nw_normal, 3 * t.n_nw, np.float64).reshape(-1, 3),
               inlet_normal=np.array(t.inlet_normal))
    lib().ref_free(C.byref(t))
    return out


def zouhe(f, vel, normal):
    f = np.array(f, dtype=np.float64)
    n = np.ascontiguousarray(normal, dtype=np.float64)
    lib().ref_zouhe(float(vel), n.ctypes.data_as(_P), f.ctypes.data_as(_P))
    return f


def read_stl(path, max_tris=10_000_000):
    buf = np.zeros((max_tris, 12))
    n = lib().ref_read_stl(str(path).encode(), buf.ctypes.data_as(_P), max_tris)
    return buf[: min(n, max_tris)]


def parse_config(path):
    c = Config()
    if lib().ref_parse_config(str(path).encode(), C.byref(c)) != 0:
        raise RuntimeError("reference parser failed")
    return c


def inlet_velocity(x, period=0.9):
    """Pulsatile::InletVelequation1..3 of the reference header, selected as
    Simulation::getInletVelocity does."""
    return lib().ref_inlet_velocity(float(x), float(period))


def pulsatile_piece(piece, x):
    return lib().ref_pulsatile_piece(int(piece), float(x))


def lattice():
    c = np.zeros((19, 3), np.int32)
    opp = np.zeros(19, np.int32)
    w = np.zeros(19)
    lib().ref_lattice(c.ctypes.data_as(_P), opp.ctypes.data_as(_P), w.ctypes.data_as(_P))
    return c, opp, w
