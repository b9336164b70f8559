"""ctypes binding of the C ABI in include/lbm_c_api.h.

Plumbing for the tests and the benchmark: it loads the in-tree
liblbm_mi355x.so and wraps its handles. The product is the C ABI itself;
nothing here computes on the CPU in place of the HIP kernels, and loading
fails loudly when the library has not been built.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblbm_mi355x.so")

OK, ERR_ARG, ERR_IO, ERR_HIP, ERR_UNSUPPORTED, ERR_PARITY, ERR_STATE = range(7)
KIND_FLUID, KIND_WALL, KIND_INLET, KIND_OUTLET = range(4)
PREC_F64, PREC_F32, PREC_F16 = 0, 1, 2
WALL_REFERENCE, WALL_HALFWAY_BB = 0, 1
INLET_REFERENCE_ZOUHE, INLET_NEE_VELOCITY, INLET_NONE = 0, 1, 2
OUTLET_REFERENCE, OUTLET_NEE_PRESSURE, OUTLET_NONE = 0, 1, 2
FLOW_CONSTANT, FLOW_PULSATILE = 0, 1
MAX_BLOCKS = 16


class LbmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[lbm error {code}] {msg}")
        self.code = code


class PlaneBlock(C.Structure):
    _fields_ = [("plane", C.c_char * 16), ("coordinate", C.c_double), ("range_x", C.c_double * 2),
                ("range_y", C.c_double * 2), ("range_z", C.c_double * 2), ("normal", C.c_double * 3),
                ("normal_len", C.c_int32)]


class Config(C.Structure):
    _fields_ = [("stl_file", C.c_char * 512), ("flow_type", C.c_char * 32),
                ("dt", C.c_double), ("rho", C.c_double), ("nu_blood", C.c_double), ("tau", C.c_double),
                ("dx", C.c_double), ("pressure", C.c_double), ("reynolds_number", C.c_double),
                ("cte_vel", C.c_double), ("systole", C.c_double), ("diastole", C.c_double),
                ("num_out", C.c_int32), ("num_cycle", C.c_int32), ("max_iter", C.c_int32),
                ("iter_to_write", C.c_int32), ("n_inlets", C.c_int32), ("n_outlets", C.c_int32),
                ("n_equations", C.c_int32), ("inlets", PlaneBlock * MAX_BLOCKS),
                ("outlets", PlaneBlock * MAX_BLOCKS)]


class GeometryOptions(C.Structure):
    _fields_ = [("q", C.c_int32), ("periodic", C.c_int32 * 3), ("nthreads", C.c_int32),
                ("verify_stride", C.c_int32), ("compute_normals", C.c_int32), ("device", C.c_int32),
                ("label_device", C.c_int32), ("classify_device", C.c_int32)]


class GeometryInfo(C.Structure):
    _fields_ = [("n_nodes", C.c_int64), ("n_fluid", C.c_int64), ("n_wall", C.c_int64),
                ("n_inlet", C.c_int64), ("n_outlet", C.c_int64), ("n_nearwall", C.c_int64),
                ("n_boundary_points", C.c_int64), ("n_triangles", C.c_int64),
                ("upstream_out_of_box", C.c_int64), ("grid", C.c_int32 * 3), ("q", C.c_int32),
                ("n_outlets", C.c_int32), ("has_inlet_normal", C.c_int32), ("spacing", C.c_double),
                ("bbox", C.c_double * 6), ("inlet_normal", C.c_double * 3)]


class TablesIn(C.Structure):
    _fields_ = [("q", C.c_int32), ("n_nodes", C.c_int64), ("grid", C.c_int32 * 3), ("origin", C.c_double * 3),
                ("spacing", C.c_double), ("ijk", C.c_void_p), ("kind", C.c_void_p), ("outlet_id", C.c_void_p),
                ("nb", C.c_void_p), ("has_inlet_normal", C.c_int32), ("inlet_normal", C.c_double * 3),
                ("n_outlets", C.c_int32), ("outlet_normals", C.c_void_p), ("n_nearwall", C.c_int64),
                ("nearwall", C.c_void_p), ("normals", C.c_void_p)]


class Params(C.Structure):
    _fields_ = [("precision", C.c_int32), ("wall_mode", C.c_int32), ("inlet_mode", C.c_int32),
                ("outlet_mode", C.c_int32), ("flow_type", C.c_int32), ("wss_every", C.c_int32),
                ("parity_checks", C.c_int32), ("device", C.c_int32), ("tau", C.c_double),
                ("dt", C.c_double), ("cte_vel", C.c_double), ("spacing", C.c_double),
                ("outlet_rho", C.c_double), ("force", C.c_double * 3)]


class Stats(C.Structure):
    _fields_ = [("mass", C.c_double), ("momentum", C.c_double * 3), ("wss_l2", C.c_double),
                ("wss_max", C.c_double), ("iteration", C.c_int64), ("time", C.c_double)]


class SolverInfo(C.Structure):
    _fields_ = [("n_nodes", C.c_int64), ("n_special", C.c_int64), ("n_bulk", C.c_int64),
                ("n_pad", C.c_int64), ("device_bytes", C.c_int64), ("q", C.c_int32),
                ("precision", C.c_int32), ("bytes_per_update", C.c_double)]


class Objective(C.Structure):
    _fields_ = [("wss_weight", C.c_double), ("dp_weight", C.c_double), ("wss_mask", C.c_void_p)]


_lib = None

# (name, argtypes, restype) of every exported symbol; also the list the
# symbol-export test checks against include/lbm_c_api.h.
_P = C.c_void_p
SIGNATURES = [
    ("lbm_lattice", [C.c_int32, _P, _P, _P], C.c_int),
    ("lbm_inlet_velocity", [C.c_double, C.c_double], C.c_double),
    ("lbm_config_parse", [C.c_char_p, _P], C.c_int),
    ("lbm_geometry_options_default", [_P], None),
    ("lbm_geometry_from_config", [C.c_char_p, _P, _P], C.c_int),
    ("lbm_geometry_from_tables", [_P, _P], C.c_int),
    ("lbm_geometry_from_labels", [_P, C.c_int32, C.c_int32, C.c_int32, _P, C.c_double, _P, C.c_int32, _P,
                                  C.c_int32, _P, C.c_int64, _P, _P], C.c_int),
    ("lbm_geometry_get_info", [_P, _P], C.c_int),
    ("lbm_geometry_get_labels", [_P, _P], C.c_int),
    ("lbm_geometry_get_nodes", [_P, _P, _P, _P], C.c_int),
    ("lbm_geometry_get_neighbours", [_P, _P], C.c_int),
    ("lbm_geometry_get_nearwall", [_P, _P, _P], C.c_int),
    ("lbm_geometry_get_bc_neighbour", [_P, _P], C.c_int),
    ("lbm_geometry_get_outlet_normals", [_P, _P], C.c_int),
    ("lbm_geometry_get_triangles", [_P, _P], C.c_int),
    ("lbm_geometry_destroy", [_P], None),
    ("lbm_params_default", [_P], None),
    ("lbm_create", [_P, _P, _P], C.c_int),
    ("lbm_set_inlet_profile", [_P, C.c_int32, _P], C.c_int),
    ("lbm_set_porosity", [_P, _P], C.c_int),
    ("lbm_step", [_P, C.c_int32], C.c_int),
    ("lbm_step_coefficients", [_P, C.c_int32, _P], C.c_int),
    ("lbm_get_populations", [_P, _P], C.c_int),
    ("lbm_get_post_collision", [_P, _P], C.c_int),
    ("lbm_set_populations", [_P, _P], C.c_int),
    ("lbm_get_wss", [_P, _P, _P, _P], C.c_int),
    ("lbm_get_osi_tawss", [_P, _P, _P], C.c_int),
    ("lbm_get_macroscopic", [_P, _P, _P], C.c_int),
    ("lbm_get_post_collision_moments", [_P, _P, C.c_int64, _P, _P], C.c_int),
    ("lbm_reduce", [_P, _P], C.c_int),
    ("lbm_zouhe", [C.c_int32, C.c_int64, _P, _P, _P], C.c_int),
    ("lbm_reduce_noalloc", [_P, _P], C.c_int),
    ("lbm_step_timed", [_P, C.c_int32, _P], C.c_int),
    ("lbm_last_batch_time", [_P, _P, _P], C.c_int),
    ("lbm_set_kernel_timing", [_P, C.c_int32], C.c_int),
    ("lbm_set_wide_offsets", [_P, C.c_int32], C.c_int),
    ("lbm_write_output", [_P, _P, C.c_int32, C.c_char_p, C.c_char_p, C.c_int32], C.c_int),
    ("lbm_write_vtu", [_P, _P, _P, C.c_char_p], C.c_int),
    ("lbm_smooth_surface_csv", [C.c_char_p, C.c_char_p, C.c_double, C.c_int32], C.c_int),
    ("lbm_last_batch_kernel_time", [_P, _P, _P], C.c_int),
    ("lbm_synchronize", [_P], C.c_int),
    ("lbm_destroy", [_P], None),
    ("lbm_get_solver_info", [_P, _P], C.c_int),
    ("lbm_copy_probe", [C.c_int32, C.c_int64, C.c_int32, _P], C.c_int),
    ("lbm_adjoint_forward", [_P, C.c_int32, _P, _P, C.c_int32, _P], C.c_int),
    ("lbm_adjoint_backward", [_P, C.c_int32, _P, _P], C.c_int),
    ("lbm_adjoint_set_tape", [_P, C.c_int32], C.c_int),
    ("lbm_checkpoint_schedule", [C.c_int32, C.c_int32, _P, C.c_int64, _P], C.c_int),
    ("lbm_last_error", [], C.c_char_p),
]


def copy_probe(nbytes, iters=20, device=0):
    """Achievable copy bandwidth in GB/s (read + write)."""
    g = C.c_double()
    _check(lib().lbm_copy_probe(device, int(nbytes), int(iters), C.byref(g)))
    return g.value


def checkpoint_schedule(nsteps, slots):
    """Binomial reversal schedule of one checkpoint interval: rows of
    (op 0 advance / 1 reverse, from, to, buffer from, buffer to)."""
    n = C.c_int64()
    _check(lib().lbm_checkpoint_schedule(int(nsteps), int(slots), None, 0, C.byref(n)))
    out = np.zeros((max(n.value, 1), 5), np.int32)
    _check(lib().lbm_checkpoint_schedule(int(nsteps), int(slots), _ptr(out), n.value, C.byref(n)))
    return out[:n.value]


def lib():
    """Load liblbm_mi355x.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing; build it with __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, args, res in SIGNATURES:
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _check(rc):
    if rc != OK:
        raise LbmError(rc, lib().lbm_last_error().decode())


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def lattice(q=19):
    c = np.zeros((q, 3), np.int32)
    opp = np.zeros(q, np.int32)
    w = np.zeros(q, np.float64)
    _check(lib().lbm_lattice(q, _ptr(c), _ptr(opp), _ptr(w)))
    return c, opp, w


def inlet_velocity(t, period=0.9):
    """Simulation::getInletVelocity: the pulsatile inlet speed lbm_step uses."""
    return lib().lbm_inlet_velocity(float(t), float(period))


def zouhe(f, vel, normal, device=0):
    """Boundary::ApplyBc on D3Q19 population vectors (rows of f), on the GPU."""
    f = np.array(f, dtype=np.float64, order="C").reshape(-1, 19)
    v = np.ascontiguousarray(np.broadcast_to(np.asarray(vel, np.float64), (f.shape[0],)))
    n = np.ascontiguousarray(normal, np.float64)
    _check(lib().lbm_zouhe(int(device), f.shape[0], _ptr(f), _ptr(v), _ptr(n)))
    return f


def parse_config(path):
    cfg = Config()
    _check(lib().lbm_config_parse(str(path).encode(), C.byref(cfg)))
    return cfg


def smooth_surface_csv(in_path, out_path, radius=0.25, pulsatile=False):
    """OutputFile::processFlowData's smoothing of a near-wall CSV."""
    _check(lib().lbm_smooth_surface_csv(str(in_path).encode(), str(out_path).encode(), float(radius),
                                        1 if pulsatile else 0))


def plane_block(plane, coordinate, normal, range_x=None, range_y=None, range_z=None):
    inf = float("inf")
    b = PlaneBlock()
    b.plane = plane.encode()
    b.coordinate = coordinate
    for name, r in (("range_x", range_x), ("range_y", range_y), ("range_z", range_z)):
        arr = getattr(b, name)
        lo, hi = (-inf, inf) if r is None else r
        arr[0], arr[1] = lo, hi
    for i, v in enumerate(normal):
        b.normal[i] = v
    b.normal_len = len(normal)
    return b


class Geometry:
    """Node tables built on the host (reference Geometry + SubDomain)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)
        self.info = GeometryInfo()
        _check(lib().lbm_geometry_get_info(self._h, C.byref(self.info)))

    @staticmethod
    def _opts(q, periodic, nthreads, verify_stride, normals, label_device=False, classify_device=False):
        o = GeometryOptions()
        lib().lbm_geometry_options_default(C.byref(o))
        o.label_device = int(label_device)
        o.classify_device = int(classify_device)
        o.q = q
        for a in range(3):
            o.periodic[a] = int(periodic[a])
        o.nthreads = nthreads
        o.verify_stride = verify_stride
        o.compute_normals = int(normals)
        return o

    @classmethod
    def from_config(cls, path, q=19, periodic=(0, 0, 0), nthreads=0, verify_stride=0, normals=True,
                    label_device=False, classify_device=False):
        """label_device=True casts the voxeliser's rays on the GPU (same labels);
        classify_device=True builds the node tables there (same tables)."""
        o = cls._opts(q, periodic, nthreads, verify_stride, normals, label_device, classify_device)
        h = C.c_void_p()
        _check(lib().lbm_geometry_from_config(str(path).encode(), C.byref(o), C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_labels(cls, fluid, origin, spacing, inlets=(), outlets=(), triangles=None, q=19,
                    periodic=(0, 0, 0), nthreads=0, normals=True, classify_device=False):
        fluid = np.ascontiguousarray(fluid, dtype=np.uint8)
        nx, ny, nz = fluid.shape
        o = cls._opts(q, periodic, nthreads, 0, normals, classify_device=classify_device)
        org = np.ascontiguousarray(origin, dtype=np.float64)
        ins = (PlaneBlock * max(1, len(inlets)))(*inlets)
        outs = (PlaneBlock * max(1, len(outlets)))(*outlets)
        if triangles is None:
            tri = np.zeros((0, 12), np.float64)
        else:
            tri = np.ascontiguousarray(triangles, dtype=np.float64).reshape(-1, 12)
        h = C.c_void_p()
        _check(lib().lbm_geometry_from_labels(_ptr(fluid), nx, ny, nz, _ptr(org), float(spacing), ins,
                                              len(inlets), outs, len(outlets), _ptr(tri), tri.shape[0],
                                              C.byref(o), C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_tables(cls, q, grid, origin, spacing, ijk, kind, nb, outlet_id=None, inlet_normal=None,
                    outlet_normals=None, nearwall=None, normals=None):
        """Node tables as a reference SubDomain holds them (lbm_geometry_from_tables)."""
        keep = [np.ascontiguousarray(ijk, np.int32), np.ascontiguousarray(kind, np.uint8),
                np.ascontiguousarray(nb, np.int32)]
        t = TablesIn()
        t.q = q
        t.n_nodes = len(keep[1])
        for a in range(3):
            t.grid[a] = int(grid[a])
            t.origin[a] = float(origin[a])
        t.spacing = float(spacing)
        t.ijk, t.kind, t.nb = [k.ctypes.data for k in keep]
        if outlet_id is not None:
            keep.append(np.ascontiguousarray(outlet_id, np.int32))
            t.outlet_id = keep[-1].ctypes.data
        if inlet_normal is not None:
            t.has_inlet_normal = 1
            for a in range(3):
                t.inlet_normal[a] = float(inlet_normal[a])
        if outlet_normals is not None and len(outlet_normals):
            keep.append(np.ascontiguousarray(outlet_normals, np.float64).reshape(-1, 3))
            t.n_outlets = keep[-1].shape[0]
            t.outlet_normals = keep[-1].ctypes.data
        if nearwall is not None and len(nearwall):
            keep.append(np.ascontiguousarray(nearwall, np.int32))
            keep.append(np.ascontiguousarray(normals, np.float64).reshape(-1, 3))
            t.n_nearwall = keep[-2].shape[0]
            t.nearwall, t.normals = keep[-2].ctypes.data, keep[-1].ctypes.data
        h = C.c_void_p()
        _check(lib().lbm_geometry_from_tables(C.byref(t), C.byref(h)))
        return cls(h.value)

    @property
    def n(self):
        return self.info.n_nodes

    @property
    def q(self):
        return self.info.q

    def write_vtu(self, rho, u, path):
        """OutputFile::WriteVTK_Collective's VTU of node fields (reference node order)."""
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        u = np.ascontiguousarray(u, dtype=np.float64)
        assert rho.shape == (self.n,) and u.shape == (self.n, 3)
        _check(lib().lbm_write_vtu(self._h, _ptr(rho), _ptr(u), str(path).encode()))

    def labels(self):
        g = tuple(self.info.grid)
        out = np.zeros(g, np.uint8)
        _check(lib().lbm_geometry_get_labels(self._h, _ptr(out)))
        return out

    def nodes(self):
        ijk = np.zeros((self.n, 3), np.int32)
        kind = np.zeros(self.n, np.uint8)
        oid = np.zeros(self.n, np.int32)
        _check(lib().lbm_geometry_get_nodes(self._h, _ptr(ijk), _ptr(kind), _ptr(oid)))
        return ijk, kind, oid

    def neighbours(self):
        nb = np.zeros((self.q, self.n), np.int32)
        _check(lib().lbm_geometry_get_neighbours(self._h, _ptr(nb)))
        return nb

    def nearwall(self):
        m = self.info.n_nearwall
        nodes = np.zeros(m, np.int32)
        nrm = np.zeros((m, 3), np.float64)
        _check(lib().lbm_geometry_get_nearwall(self._h, _ptr(nodes), _ptr(nrm)))
        return nodes, nrm

    def bc_neighbour(self):
        out = np.zeros(self.n, np.int32)
        _check(lib().lbm_geometry_get_bc_neighbour(self._h, _ptr(out)))
        return out

    def triangles(self):
        out = np.zeros((self.info.n_triangles, 12), np.float64)
        if self.info.n_triangles:
            _check(lib().lbm_geometry_get_triangles(self._h, _ptr(out)))
        return out

    def outlet_normals(self):
        out = np.zeros((self.info.n_outlets, 3), np.float64)
        if self.info.n_outlets:
            _check(lib().lbm_geometry_get_outlet_normals(self._h, _ptr(out)))
        return out

    def close(self):
        if self._h:
            lib().lbm_geometry_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_params(**kw):
    p = Params()
    lib().lbm_params_default(C.byref(p))
    for k, v in kw.items():
        if k == "force":
            for i in range(3):
                p.force[i] = v[i]
        else:
            setattr(p, k, v)
    return p


class Solver:
    """Device solver handle (reference Simulation state + time loop)."""

    def __init__(self, geom, params=None, **kw):
        self.geom = geom
        self.params = params if params is not None else default_params(**kw)
        h = C.c_void_p()
        _check(lib().lbm_create(geom._h, C.byref(self.params), C.byref(h)))
        self._h = h
        self.q = geom.q
        self.n = geom.n
        self.n_modes = 1

    def step(self, nsteps=1):
        _check(lib().lbm_step(self._h, int(nsteps)))

    def step_coefficients(self, coef):
        coef = np.ascontiguousarray(coef, dtype=np.float64)
        _check(lib().lbm_step_coefficients(self._h, coef.shape[0], _ptr(coef)))

    def step_timed(self, nsteps):
        ms = C.c_double()
        _check(lib().lbm_step_timed(self._h, int(nsteps), C.byref(ms)))
        return ms.value

    def last_batch_ms(self):
        ms = C.c_double()
        n = C.c_int32()
        _check(lib().lbm_last_batch_time(self._h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def write_output(self, geom, iteration, phase_name, directory, pulsatile=False):
        """Simulation::WriteOutput: the VTU, the near-wall CSV and its smoothed copy."""
        _check(lib().lbm_write_output(geom._h, self._h, int(iteration), phase_name.encode(), str(directory).encode(),
                                      1 if pulsatile else 0))

    def set_wide_offsets(self, on=True):
        """Take the 64-bit-offset step path whatever the state size."""
        _check(lib().lbm_set_wide_offsets(self._h, 1 if on else 0))

    def set_kernel_timing(self, on=True):
        """Time the stress-flush launches of later batches separately."""
        _check(lib().lbm_set_kernel_timing(self._h, 1 if on else 0))

    def last_batch_kernel_ms(self):
        """(step kernel ms, stress-flush ms) of the last batch (needs set_kernel_timing)."""
        a, b = C.c_double(), C.c_double()
        _check(lib().lbm_last_batch_kernel_time(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_inlet_profile(self, basis):
        basis = np.ascontiguousarray(np.atleast_2d(basis), dtype=np.float64)
        _check(lib().lbm_set_inlet_profile(self._h, basis.shape[0], _ptr(basis)))
        self.n_modes = basis.shape[0]

    def set_porosity(self, alpha):
        if alpha is None:
            _check(lib().lbm_set_porosity(self._h, None))
        else:
            a = np.ascontiguousarray(alpha, dtype=np.float32)
            _check(lib().lbm_set_porosity(self._h, _ptr(a)))

    def populations(self):
        f = np.zeros((self.n, self.q), np.float64)
        _check(lib().lbm_get_populations(self._h, _ptr(f)))
        return f

    def post_collision(self):
        f = np.zeros((self.n, self.q), np.float64)
        _check(lib().lbm_get_post_collision(self._h, _ptr(f)))
        return f

    def set_populations(self, f=None):
        """f: (n, q) FP64 populations; None = the rest state f = w_q."""
        if f is None:
            _check(lib().lbm_set_populations(self._h, None))
            return
        f = np.ascontiguousarray(f, dtype=np.float64)
        assert f.shape == (self.n, self.q)
        _check(lib().lbm_set_populations(self._h, _ptr(f)))

    def wss(self):
        m = self.geom.info.n_nearwall
        w = np.zeros(m)
        acc = np.zeros((m, 3))
        mag = np.zeros(m)
        _check(lib().lbm_get_wss(self._h, _ptr(w), _ptr(acc), _ptr(mag)))
        return w, acc, mag

    def osi_tawss(self):
        m = self.geom.info.n_nearwall
        osi = np.zeros(m)
        tawss = np.zeros(m)
        _check(lib().lbm_get_osi_tawss(self._h, _ptr(osi), _ptr(tawss)))
        return osi, tawss

    def macroscopic(self):
        rho = np.zeros(self.n)
        u = np.zeros((self.n, 3))
        _check(lib().lbm_get_macroscopic(self._h, _ptr(rho), _ptr(u)))
        return rho, u

    def post_collision_moments(self, nodes):
        """(rho, u) of the post-collision state at the given nodes."""
        nodes = np.ascontiguousarray(nodes, dtype=np.int32)
        rho = np.zeros(len(nodes))
        u = np.zeros((len(nodes), 3))
        _check(lib().lbm_get_post_collision_moments(self._h, _ptr(nodes), len(nodes), _ptr(rho), _ptr(u)))
        return rho, u

    def reduce(self):
        s = Stats()
        _check(lib().lbm_reduce(self._h, C.byref(s)))
        return s

    def adjoint_forward(self, nsteps, coef=None, wss_weight=1.0, dp_weight=0.0, wss_mask=None,
                        checkpoint_every=200):
        """Forward pass with checkpoints; returns the objective J."""
        obj = Objective(float(wss_weight), float(dp_weight), None)
        if wss_mask is not None:
            self._mask = np.ascontiguousarray(wss_mask, dtype=np.uint8)
            obj.wss_mask = self._mask.ctypes.data
        cp = None
        if coef is not None:
            self._coef = np.ascontiguousarray(coef, dtype=np.float64).reshape(nsteps, self.n_modes)
            cp = _ptr(self._coef)
        J = C.c_double()
        _check(lib().lbm_adjoint_forward(self._h, int(nsteps), cp, C.byref(obj), int(checkpoint_every),
                                         C.byref(J)))
        self._adj_steps = int(nsteps)
        self._adj_modes = self.n_modes
        return J.value

    def adjoint_backward(self, slots=0, tape=True):
        """Reverse pass: (dJ/dcoef [steps, modes], dJ/dalpha [nodes]).
        tape=False re-pulls every reverse step's moments instead of reading
        the recomputation's moments tape (same result bit for bit)."""
        _check(lib().lbm_adjoint_set_tape(self._h, 1 if tape else 0))
        gc = np.zeros((self._adj_steps, self._adj_modes))
        ga = np.zeros(self.n)
        _check(lib().lbm_adjoint_backward(self._h, int(slots), _ptr(gc), _ptr(ga)))
        return gc, ga

    def info(self):
        i = SolverInfo()
        _check(lib().lbm_get_solver_info(self._h, C.byref(i)))
        return i

    def synchronize(self):
        _check(lib().lbm_synchronize(self._h))

    def close(self):
        if self._h:
            lib().lbm_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
