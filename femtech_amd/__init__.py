"""femtech_amd -- MI355X explicit C3D8 engine with FemTech's call surface.

The compute path is the HIP library built by ``femtech_amd.build`` (C ABI in
include/femtech_amd.h).  This module binds it with ctypes; there is no CPU
fallback: without the built library every call raises.
"""
import ctypes
import os

import numpy as np

from .build import LIB_PATH

FT_FP64 = 0
FT_FP32 = 1
MOTION_STRIDE = 16
MOTION_FIXED = 0
MOTION_LINEAR = 1
MOTION_RAMP_HOLD = 2
MOTION_ROT_HALFSINE = 3
STATUS_NONFINITE = 1
STATUS_INVERTED = 2
STATUS_DT_COLLAPSE = 4
STATUS_DT_FP32 = 16
STATUS_DT_BOUND = 32
# force-gather layouts (ft_engine_layout / ft_engine_set_layout)
LAYOUT_AUTO, LAYOUT_CORNER, LAYOUT_NODE_SLOTS, LAYOUT_ELL, LAYOUT_CORNER_PAIRS = 0, 1, 2, 3, 4
# error codes (include/femtech_amd.h): FemTech's own, then the engine's
FT_ERR_MATERIAL = 1
FT_ERR_INPUT = 3
FT_ERR_ALLOC = 12
FT_ERR_DT = 19
FT_ERR_HIP = 50
FT_ERR_ARG = 51
FT_ERR_STATE = 52
FT_ERR_NUMERIC = 53

# FemTech's termination codes (InitFinalizeFemTech.cpp:82-86 callers) and the
# engine's own.
ERR_MATERIAL, ERR_INPUT, ERR_ALLOC, ERR_DT = 1, 3, 12, 19
ERR_HIP, ERR_ARG, ERR_STATE, ERR_NUMERIC = 50, 51, 52, 53

_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_VP = ctypes.c_void_p
_F = ctypes.POINTER(ctypes.c_float)

# name -> argtypes (all return int unless listed in _RET)
_SIGS = {
    "ft_last_error": [],
    "ft_version": [],
    "ft_device_count": [_I],
    "ft_engine_create": [ctypes.c_int, ctypes.c_int, _D, _I, _I, ctypes.c_int, _I, _D,
                         ctypes.c_int, ctypes.POINTER(_VP)],
    "ft_engine_destroy": [_VP],
    "ft_engine_info": [_VP, _I, _I, _I],
    "ft_engine_layout": [_VP, _I, _I],
    "ft_engine_set_layout": [_VP, ctypes.c_int],
    "ft_snapshot_begin": [_VP],
    "ft_snapshot_end": [_VP, _D, _D, _I],
    "ft_snapshot_view": [_VP],
    "ft_set_boundary": [_VP, _I, ctypes.c_int, _D],
    "ft_set_state": [_VP, _D, _D],
    "ft_reset_history": [_VP],
    "ft_assemble_lumped_mass": [_VP, _D],
    "ft_stable_time_step": [_VP, _D, _I, ctypes.c_double, _D],
    "ft_get_force": [_VP, _D, ctypes.c_double, _D],
    "ft_element_forces": [_VP, _D, ctypes.c_double, ctypes.c_int, _D],
    "ft_run_init": [_VP, ctypes.c_double, ctypes.c_double, ctypes.c_double],
    "ft_run_steps": [_VP, ctypes.c_int, _I],
    "ft_get_time": [_VP, _D, _D, _I],
    "ft_get_state": [_VP, _D, _D, _D, _D],
    "ft_get_energy": [_VP, _D],
    "ft_set_history": [_VP, ctypes.c_int],
    "ft_get_history": [_VP, _D, _D, _D, _I],
    "ft_get_status": [_VP, _I],
    "ft_get_visco_history": [_VP, _D],
    "ft_set_timing": [_VP, ctypes.c_int],
    "ft_set_graph": [_VP, ctypes.c_int],
    "ft_set_persistent": [_VP, ctypes.c_int],
    "ft_autotune": [_VP, ctypes.c_int, _D, _I],
    "ft_get_segments": [_VP, _I, _I],
    "ft_set_segment_shape": [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int],
    "ft_get_timing": [_VP, _D, _I],
    "ft_traffic_model": [_VP, _D],
    "ft_copy_probe": [ctypes.c_double, ctypes.c_int, _D],
    "ft_point_eval": [ctypes.c_int, ctypes.c_int, _D, _D],
    "ft_get_force_fields": [_VP, _D, ctypes.c_double, _D, _D, _D, _D],
    "ft_set_visco_history": [_VP, _D],
    "ft_mesh_ellipsoid": [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, _D, _I, _D],
    "ft_set_instances": [_VP, ctypes.c_int, _I, _I],
    "ft_set_instance_end_times": [_VP, _D],
    "ft_get_instance": [_VP, ctypes.c_int, _D],
    "ft_set_fibres": [_VP, _D],
    "ft_injury_enable": [_VP, _I],
    "ft_injury_disable": [_VP],
    "ft_injury_update": [_VP, _D, ctypes.c_double, ctypes.c_double],
    "ft_injury_results": [_VP, _D, _I],
    "ft_element_strain": [_VP, _D, _D, _D],
    "ft_mesh_block": [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                      ctypes.c_double, ctypes.c_double, ctypes.c_double,
                      ctypes.c_ulonglong, _D, _I],
    "ft_write_inp": [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _D, _I, _I],
    "ft_read_mesh": [ctypes.c_char_p, _I, _I, _D, _I, _I, _I],
    "ft_read_materials": [ctypes.c_char_p, ctypes.c_int, _I, _D],
    "ft_surrogate_reserve": [_VP, ctypes.c_long],
    "ft_surrogate_collect": [_VP, ctypes.POINTER(ctypes.c_long)],
    "ft_surrogate_samples": [_VP, ctypes.c_long, ctypes.c_long, _F, _F],
    "ft_surrogate_train": [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_ulonglong, _D],
    "ft_surrogate_get": [_VP, _F, _F],
    "ft_surrogate_set": [_VP, _F, _F],
    "ft_surrogate_enable": [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int],
    "ft_surrogate_stats": [_VP, _D],
    "ft_mlp_param_count": [],
    "ft_mlp_create": [ctypes.c_ulonglong, ctypes.POINTER(_VP)],
    "ft_mlp_destroy": [_VP],
    "ft_mlp_get_params": [_VP, _F],
    "ft_mlp_set_params": [_VP, _F],
    "ft_mlp_forward": [_VP, _F, ctypes.c_int, _F],
    "ft_mlp_gradients": [_VP, _F, _F, ctypes.c_int, _D, _F],
    "ft_mlp_train": [_VP, _F, _F, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                     ctypes.c_ulonglong, _D],
}
_RET = {"ft_last_error": ctypes.c_char_p, "ft_snapshot_view": ctypes.c_void_p}


class FemTechError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("error %d: %s" % (code, msg))
        self.code = code


def library():
    """Loads the built HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("femtech_amd: %s is missing; run femtech_amd.build.build() "
                               "(the engine has no CPU fallback)" % LIB_PATH)
        lib = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = _RET.get(name, ctypes.c_int)
        _lib = lib
    return _lib


def _check(rc):
    if rc != 0:
        raise FemTechError(rc, library().ft_last_error().decode())


def _dp(a):
    return None if a is None else a.ctypes.data_as(_D)


def _ip(a):
    return None if a is None else a.ctypes.data_as(_I)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def device_count():
    c = ctypes.c_int(0)
    _check(library().ft_device_count(ctypes.byref(c)))
    return c.value


# ft_point_eval operations (include/femtech_amd.h) and their record sizes
PT_SHAPE, PT_DEFGRAD, PT_DETF, PT_INVF, PT_STRESS, PT_BMAT, PT_FINT, PT_CHARLEN, PT_TIMESTEP, PT_INV3 = range(10)
PT_SIZES = {PT_SHAPE: (27, 33), PT_DEFGRAD: (48, 9), PT_DETF: (9, 1), PT_INVF: (10, 9), PT_STRESS: (20, 34),
            PT_BMAT: (12, 18), PT_FINT: (64, 192), PT_CHARLEN: (24, 1), PT_TIMESTEP: (27, 1), PT_INV3: (9, 10)}


def point_eval(op, records, out=None):
    """FemTech's per-Gauss-point routines on the device: `records` is an
    (n, in_size) array of input records; returns the (n, out_size) outputs
    (`out` supplies their initial contents, e.g. the Prony history of the
    stress operation)."""
    ni, no = PT_SIZES[op]
    rec = _f64(np.asarray(records, np.float64).reshape(-1, ni))
    n = rec.shape[0]
    res = np.zeros((n, no)) if out is None else _f64(np.array(out, np.float64).reshape(n, no))
    _check(library().ft_point_eval(op, n, _dp(rec), _dp(res)))
    return res


def copy_probe(megabytes=256.0, repeats=10):
    """GB/s of a plain device copy (read + write) of `megabytes`."""
    out = ctypes.c_double(0)
    _check(library().ft_copy_probe(float(megabytes), int(repeats), ctypes.byref(out)))
    return out.value


def mesh_block(nx, ny, nz, h=(1e-3, 1e-3, 1e-3), jitter=0.0, seed=0):
    """Structured hex block (nodes x-fastest), interior nodes jittered by
    +-jitter*h from a seeded 64-bit Mersenne twister."""
    if np.isscalar(h):
        h = (h, h, h)
    nn = (nx + 1) * (ny + 1) * (nz + 1)
    ne = nx * ny * nz
    X = np.empty(3 * nn, np.float64)
    conn = np.empty(8 * ne, np.int32)
    _check(library().ft_mesh_block(nx, ny, nz, h[0], h[1], h[2], jitter, seed, _dp(X), _ip(conn)))
    return X, conn


def mesh_ellipsoid(n, a, b, c):
    """Ellipsoid of semi-axes a, b, c from an n^3 cube mapped onto the ball;
    returns (X, conn, radius) with radius the normalised radius of each
    element's centroid (0 centre, 1 surface)."""
    nn, ne = (n + 1) ** 3, n ** 3
    X = np.empty(3 * nn, np.float64)
    conn = np.empty(8 * ne, np.int32)
    r = np.empty(ne, np.float64)
    _check(library().ft_mesh_ellipsoid(n, a, b, c, _dp(X), _ip(conn), _dp(r)))
    return X, conn, r


def write_inp(path, X, conn, pid=None):
    X, conn = _f64(X), _i32(conn)
    ne = conn.size // 8
    pid = None if pid is None else _i32(pid)
    _check(library().ft_write_inp(path.encode(), X.size // 3, ne, _dp(X), _ip(conn), _ip(pid)))


def read_mesh(path):
    nn, ne = ctypes.c_int(0), ctypes.c_int(0)
    lib = library()
    _check(lib.ft_read_mesh(path.encode(), ctypes.byref(nn), ctypes.byref(ne), None, None, None, None))
    X = np.empty(3 * nn.value, np.float64)
    conn = np.empty(8 * ne.value, np.int32)
    pid = np.empty(ne.value, np.int32)
    gid = np.empty(nn.value, np.int32)
    _check(lib.ft_read_mesh(path.encode(), ctypes.byref(nn), ctypes.byref(ne), _dp(X), _ip(conn),
                            _ip(pid), _ip(gid)))
    return X, conn, pid, gid


def read_materials(path, n_parts):
    mat = np.zeros(n_parts, np.int32)
    props = np.zeros(9 * n_parts, np.float64)
    _check(library().ft_read_materials(path.encode(), n_parts, _ip(mat), _dp(props)))
    return mat, props


def motion(kind, *params):
    m = np.zeros(MOTION_STRIDE, np.float64)
    m[0] = kind
    m[1:1 + len(params)] = params
    return m


class Engine:
    """One mesh resident on the current HIP device (see include/femtech_amd.h)."""

    def __init__(self, X, conn, pid, material_id, props, precision=FT_FP64):
        self._lib = library()
        self.X = _f64(X)
        self.conn = _i32(conn)
        self.pid = _i32(pid)
        self.material_id = _i32(material_id)
        self.props = _f64(props)
        self.nn = self.X.size // 3
        self.ne = self.conn.size // 8
        self.precision = precision
        h = _VP()
        _check(self._lib.ft_engine_create(self.nn, self.ne, _dp(self.X), _ip(self.conn), _ip(self.pid),
                                          self.material_id.size, _ip(self.material_id), _dp(self.props),
                                          precision, ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.ft_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def layout(self):
        """(layout, ell_width): LAYOUT_CORNER_PAIRS (corner slots, face
        neighbours along xi of one wavefront summing their shared corners in
        registers), LAYOUT_CORNER (per-node corner slots), LAYOUT_NODE_SLOTS
        (per-node slots by element rank) or LAYOUT_ELL (element-major forces
        gathered through the node-to-element table), and that table's
        width."""
        c, w = ctypes.c_int(0), ctypes.c_int(0)
        _check(self._lib.ft_engine_layout(self._h, ctypes.byref(c), ctypes.byref(w)))
        return c.value, w.value

    def set_layout(self, layout):
        """Choose the force-gather layout before run_init (LAYOUT_AUTO picks
        the first of corner pairs, node slots, ELL that applies)."""
        _check(self._lib.ft_engine_set_layout(self._h, int(layout)))

    # --- set-up -----------------------------------------------------------
    def set_boundary(self, dof_motion, motions):
        dm = _i32(dof_motion)
        mt = _f64(np.asarray(motions, np.float64).reshape(-1))
        nm = mt.size // MOTION_STRIDE
        _check(self._lib.ft_set_boundary(self._h, _ip(dm), nm, _dp(mt) if nm else None))

    def set_state(self, u=None, v=None):
        u = None if u is None else _f64(u)
        v = None if v is None else _f64(v)
        _check(self._lib.ft_set_state(self._h, _dp(u), _dp(v)))

    def lumped_mass(self):
        m = np.empty(3 * self.nn, np.float64)
        _check(self._lib.ft_assemble_lumped_mass(self._h, _dp(m)))
        return m

    # --- per-routine ------------------------------------------------------
    def stable_time_step(self, u=None, boundary=None, failure_dt=0.0):
        out = ctypes.c_double(0)
        u = None if u is None else _f64(u)
        b = None if boundary is None else _i32(boundary)
        _check(self._lib.ft_stable_time_step(self._h, _dp(u), _ip(b), failure_dt, ctypes.byref(out)))
        return out.value

    def get_force(self, u=None, dt=0.0):
        fi = np.empty(3 * self.nn, np.float64)
        u = None if u is None else _f64(u)
        _check(self._lib.ft_get_force(self._h, _dp(u), dt, _dp(fi)))
        return fi

    def force_fields(self, u=None, dt=0.0):
        """GetForce with the per-Gauss-point fields: (fi, F [ne, 8, 9],
        detF [ne, 8], pk2 [ne, 8, 6])."""
        fi = np.empty(3 * self.nn)
        F = np.empty(72 * self.ne)
        dF = np.empty(8 * self.ne)
        S = np.empty(48 * self.ne)
        u = None if u is None else _f64(u)
        _check(self._lib.ft_get_force_fields(self._h, _dp(u), dt, _dp(fi), _dp(F), _dp(dF), _dp(S)))
        return fi, F.reshape(-1, 8, 9), dF.reshape(-1, 8), S.reshape(-1, 8, 6)

    def set_visco_history(self, hist):
        h = _f64(hist)
        _check(self._lib.ft_set_visco_history(self._h, _dp(h)))

    def element_forces(self, u=None, dt=0.0, update_history=False):
        fe = np.empty(24 * self.ne, np.float64)
        u = None if u is None else _f64(u)
        _check(self._lib.ft_element_forces(self._h, _dp(u), dt, int(update_history), _dp(fe)))
        return fe

    # --- resident loop ----------------------------------------------------
    def run_init(self, tmax=float("inf"), dt_factor=0.8, failure_dt=1e-11):
        _check(self._lib.ft_run_init(self._h, tmax, dt_factor, failure_dt))

    def run_steps(self, n):
        done = ctypes.c_int(0)
        _check(self._lib.ft_run_steps(self._h, int(n), ctypes.byref(done)))
        return done.value

    def snapshot_begin(self):
        """Queue a copy of the current displacement; it reaches the host
        while later run_steps calls compute."""
        _check(self._lib.ft_snapshot_begin(self._h))

    def snapshot_end(self, copy=True):
        """(u, time, step) of the queued snapshot (waits for its copy); with
        copy=False, u is None and snapshot_view() reads it in place."""
        u = np.empty(3 * self.nn, np.float64) if copy else None
        t, k = ctypes.c_double(0), ctypes.c_int(0)
        _check(self._lib.ft_snapshot_end(self._h, _dp(u) if copy else None, ctypes.byref(t), ctypes.byref(k)))
        return u, t.value, k.value

    def snapshot_view(self):
        """The last snapshot's displacement in pinned host memory (read-only,
        valid until the next snapshot_begin), or None."""
        p = self._lib.ft_snapshot_view(self._h)
        if not p:
            return None
        a = np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_double)), shape=(3 * self.nn,))
        a.flags.writeable = False
        return a

    def time(self):
        t, dt, k = ctypes.c_double(0), ctypes.c_double(0), ctypes.c_int(0)
        _check(self._lib.ft_get_time(self._h, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(k)))
        return t.value, dt.value, k.value

    def state(self):
        u, v, a, fi = (np.empty(3 * self.nn, np.float64) for _ in range(4))
        _check(self._lib.ft_get_state(self._h, _dp(u), _dp(v), _dp(a), _dp(fi)))
        return u, v, a, fi

    def energy(self):
        out = np.empty(4, np.float64)
        _check(self._lib.ft_get_energy(self._h, _dp(out)))
        return out

    def set_history(self, capacity):
        _check(self._lib.ft_set_history(self._h, int(capacity)))
        self._hist_cap = int(capacity)

    def history(self):
        cap = getattr(self, "_hist_cap", 0)
        t, dt = np.empty(max(cap, 1)), np.empty(max(cap, 1))
        en = np.empty(4 * max(cap, 1))
        n = ctypes.c_int(0)
        _check(self._lib.ft_get_history(self._h, _dp(t), _dp(dt), _dp(en), ctypes.byref(n)))
        k = n.value
        return t[:k], dt[:k], en[:4 * k].reshape(k, 4)

    def status(self):
        s = ctypes.c_int(0)
        _check(self._lib.ft_get_status(self._h, ctypes.byref(s)))
        return s.value

    def set_instances(self, node_offsets, elem_offsets):
        """Declare independent instances (batched simulations) by their node
        and element ranges; each then has its own dt, time and energy."""
        no = np.ascontiguousarray(node_offsets, np.int32)
        eo = np.ascontiguousarray(elem_offsets, np.int32)
        _check(self._lib.ft_set_instances(self._h, no.size - 1, _ip(no), _ip(eo)))
        self.n_inst = no.size - 1

    def set_instance_end_times(self, tmax):
        _check(self._lib.ft_set_instance_end_times(self._h, _dp(_f64(tmax))))

    def instance(self, i):
        """(time, next dt, step, state, W_int, W_ext, KE, |balance|); state 0
        running, 1 reached its end time, 2 stopped without a stable step."""
        out = np.empty(8, np.float64)
        _check(self._lib.ft_get_instance(self._h, int(i), _dp(out)))
        return out

    def set_fibres(self, directions):
        """Fibre direction per element (3 each) for the anisotropic HGO law."""
        d = _f64(directions)
        _check(self._lib.ft_set_fibres(self._h, _dp(d)))

    def injury_enable(self, include_part=None):
        """Track the ex5 injury criteria over the following steps."""
        inc = None if include_part is None else np.ascontiguousarray(include_part, np.int32)
        _check(self._lib.ft_injury_enable(self._h, _ip(inc)))

    def injury_disable(self):
        _check(self._lib.ft_injury_disable(self._h))

    def injury_update(self, u, time, dt):
        u = None if u is None else _f64(u)
        _check(self._lib.ft_injury_update(self._h, _dp(u), float(time), float(dt)))

    def injury_results(self):
        """(out[21], flags[ne]) as documented for ft_injury_results."""
        out = np.empty(21, np.float64)
        flags = np.empty(self.ne, np.int32)
        _check(self._lib.ft_injury_results(self._h, _dp(out), _ip(flags)))
        return out, flags

    def element_strain(self, u=None):
        """Element-averaged Green-Lagrange strain (9 per element, column-major)
        and principal summary (max principal, min principal, max shear)."""
        E = np.empty(9 * self.ne, np.float64)
        P = np.empty(3 * self.ne, np.float64)
        u = None if u is None else _f64(u)
        _check(self._lib.ft_element_strain(self._h, _dp(u), _dp(E), _dp(P)))
        return E, P

    def visco_history(self):
        h = np.empty(144 * self.ne, np.float64)
        _check(self._lib.ft_get_visco_history(self._h, _dp(h)))
        return h

    def set_graph(self, steps):
        """Steps per captured HIP graph in run_steps (0: separate launches)."""
        _check(self._lib.ft_set_graph(self._h, int(steps)))

    def set_persistent(self, enable=True):
        """run_steps in one cooperative launch (grid-wide barriers between
        the phases of every step) instead of launches per step."""
        _check(self._lib.ft_set_persistent(self._h, 1 if enable else 0))

    def segments(self):
        """Material segments: rows (material, begin, end, waves, tile)."""
        n = ctypes.c_int(0)
        _check(self._lib.ft_get_segments(self._h, ctypes.byref(n), None))
        d = np.zeros(5 * n.value, np.int32)
        _check(self._lib.ft_get_segments(self._h, ctypes.byref(n), _ip(d)))
        return d.reshape(n.value, 5)

    def set_segment_shape(self, segment, waves, tile):
        _check(self._lib.ft_set_segment_shape(self._h, int(segment), int(waves), int(tile)))

    def autotune(self, repeats=5):
        """Times every element-kernel launch shape per material segment and
        keeps the fastest.  Returns rows (material, elements, waves, tile,
        best us, default-shape us)."""
        n = len(self.segments())
        rep = np.zeros(6 * n)
        cnt = ctypes.c_int(0)
        _check(self._lib.ft_autotune(self._h, int(repeats), _dp(rep), ctypes.byref(cnt)))
        return rep.reshape(n, 6)

    # ---- surrogate constitutive path (materials 7 and 8) ----
    def surrogate_reserve(self, capacity):
        _check(self._lib.ft_surrogate_reserve(self._h, int(capacity)))

    def surrogate_collect(self):
        """Appends the training samples of the current state; returns the
        number held."""
        n = ctypes.c_long(0)
        _check(self._lib.ft_surrogate_collect(self._h, ctypes.byref(n)))
        return n.value

    def surrogate_samples(self, first, count):
        x = np.empty((count, 16), np.float32)
        t = np.empty((count, 8), np.float32)
        _check(self._lib.ft_surrogate_samples(self._h, int(first), int(count), _fp(x), _fp(t)))
        return x, t

    def surrogate_train(self, steps, batch=8192, lr=2e-3, seed=1):
        """Trains on the collected samples; returns (final loss, samples,
        relative L2 error of S_iso on a spread subset, steps)."""
        out = np.zeros(4)
        _check(self._lib.ft_surrogate_train(self._h, int(steps), int(batch), float(lr), ctypes.c_ulonglong(seed),
                                            _dp(out)))
        return out

    def surrogate_get(self):
        p = np.empty(self._lib.ft_mlp_param_count(), np.float32)
        nrm = np.empty(80, np.float32)
        _check(self._lib.ft_surrogate_get(self._h, _fp(p), _fp(nrm)))
        return p, nrm

    def surrogate_set(self, params, norm):
        p = np.ascontiguousarray(params, np.float32)
        nrm = np.ascontiguousarray(norm, np.float32)
        _check(self._lib.ft_surrogate_set(self._h, _fp(p), _fp(nrm)))

    def surrogate_enable(self, enable=True, audit_every=100, audit_permille=10):
        _check(self._lib.ft_surrogate_enable(self._h, int(enable), int(audit_every), int(audit_permille)))

    def surrogate_stats(self):
        """(audited rel. L2 error of S_iso, the same relative to the full
        PK2, fallback share, audited, surrogate and fallback evaluations)."""
        out = np.zeros(6)
        _check(self._lib.ft_surrogate_stats(self._h, _dp(out)))
        return out

    def reset_history(self):
        """Prony history back to zero (a new run from the same engine)."""
        _check(self._lib.ft_reset_history(self._h))

    def set_timing(self, enable):
        _check(self._lib.ft_set_timing(self._h, int(enable)))

    def timing(self):
        """Summed ms and launch counts per class (element, nodal, other)."""
        ms = np.zeros(3)
        n = np.zeros(3, np.int32)
        _check(self._lib.ft_get_timing(self._h, _dp(ms), _ip(n)))
        return ms, n

    def traffic_model(self):
        b = np.zeros(2)
        _check(self._lib.ft_traffic_model(self._h, _dp(b)))
        return b


def _fp(a):
    return a.ctypes.data_as(_F)


class Mlp:
    """The constitutive surrogate network on its own (BF16 matrix-core
    kernels): 16 inputs (12 used) -> 4 x 128 SiLU -> 8 outputs (6 used)."""

    def __init__(self, seed=1):
        self._lib = library()
        h = _VP()
        _check(self._lib.ft_mlp_create(ctypes.c_ulonglong(seed), ctypes.byref(h)))
        self._h = h
        self.n_params = self._lib.ft_mlp_param_count()

    def close(self):
        if self._h:
            self._lib.ft_mlp_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def params(self):
        p = np.empty(self.n_params, np.float32)
        _check(self._lib.ft_mlp_get_params(self._h, _fp(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params
        _check(self._lib.ft_mlp_set_params(self._h, _fp(p)))

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 16)
        y = np.empty((x.shape[0], 8), np.float32)
        _check(self._lib.ft_mlp_forward(self._h, _fp(x), x.shape[0], _fp(y)))
        return y

    def gradients(self, x, t):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 16)
        t = np.ascontiguousarray(t, np.float32).reshape(-1, 8)
        g = np.empty(self.n_params, np.float32)
        loss = ctypes.c_double(0)
        _check(self._lib.ft_mlp_gradients(self._h, _fp(x), _fp(t), x.shape[0], ctypes.byref(loss), _fp(g)))
        return loss.value, g

    def train(self, x, t, steps, batch=4096, lr=2e-3, seed=1):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 16)
        t = np.ascontiguousarray(t, np.float32).reshape(-1, 8)
        loss = ctypes.c_double(0)
        _check(self._lib.ft_mlp_train(self._h, _fp(x), _fp(t), x.shape[0], int(steps), int(batch), float(lr),
                                      ctypes.c_ulonglong(seed), ctypes.byref(loss)))
        return loss.value

