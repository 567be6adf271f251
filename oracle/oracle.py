"""TEST INFRASTRUCTURE ONLY: ctypes binding of the FP64 CPU oracle.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module; the product package never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.oracle_create.restype = ctypes.c_void_p
        L.oracle_create.argtypes = [ctypes.c_int, ctypes.c_int, _D, _I, _I, ctypes.c_int, _I, _D]
        L.oracle_destroy.argtypes = [ctypes.c_void_p]
        L.oracle_set_threads.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.oracle_shape_functions.argtypes = [ctypes.c_void_p, _D, _D]
        L.oracle_lumped_mass.argtypes = [ctypes.c_void_p, _D]
        L.oracle_get_force.argtypes = [ctypes.c_void_p, _D, ctypes.c_double, _D]
        L.oracle_element_forces.argtypes = [ctypes.c_void_p, _D, ctypes.c_double, ctypes.c_int, _D]
        L.oracle_stable_dt.restype = ctypes.c_double
        L.oracle_stable_dt.argtypes = [ctypes.c_void_p, _D, _I]
        L.oracle_get_history.argtypes = [ctypes.c_void_p, _D]
        L.oracle_set_history.argtypes = [ctypes.c_void_p, _D]
        L.oracle_set_bc.argtypes = [ctypes.c_void_p, _I, ctypes.c_int, _D]
        L.oracle_set_initial.argtypes = [ctypes.c_void_p, _D, _D]
        L.oracle_run.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_double, _D, _D, _D, _D, _D, _D, _D, ctypes.c_int, _I, _D]
        L.oracle_gauss_c3d8.argtypes = [_D, _D]
        L.oracle_shape_c3d8.argtypes = [_D, _D, _D]
        L.oracle_stress.argtypes = [ctypes.c_int, _D, _D, ctypes.c_double, ctypes.c_double, _D, _D]
        L.oracle_det3.restype = ctypes.c_double
        L.oracle_det3.argtypes = [_D]
        L.oracle_hex_volume.restype = ctypes.c_double
        L.oracle_hex_volume.argtypes = [_D]
        L.oracle_face_area.restype = ctypes.c_double
        L.oracle_face_area.argtypes = [_D, _I]
        L.oracle_char_length.restype = ctypes.c_double
        L.oracle_char_length.argtypes = [_D]
        L.oracle_element_mass.argtypes = [_D, ctypes.c_double, _D]
        L.oracle_motion_eval.argtypes = [_D, _D, ctypes.c_double, _D, _D, _D]
        L.oracle_element_strain.argtypes = [ctypes.c_void_p, _D, _D, _D]
        L.oracle_injury_init.argtypes = [ctypes.c_void_p, _I]
        L.oracle_stress_fibre.argtypes = [ctypes.c_int, _D, _D, _D, ctypes.c_double, ctypes.c_double, _D, _D]
        L.oracle_set_fibres.argtypes = [ctypes.c_void_p, _D]
        L.oracle_injury_update.argtypes = [ctypes.c_void_p, _D, ctypes.c_double, ctypes.c_double]
        L.oracle_injury_results.argtypes = [ctypes.c_void_p, _D, _I]
        _lib = L
    return _lib


def _d(a):
    return None if a is None else a.ctypes.data_as(_D)


def _i(a):
    return None if a is None else a.ctypes.data_as(_I)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def stress(material, props, F, dt=0.0, hist=None, fibre=None):
    """PK2 (Voigt 6) of StressUpdate for column-major F; hist (27) updated;
    fibre: direction for material 8."""
    F = f64(F).reshape(9)
    p = f64(props).reshape(9)
    S = np.zeros(6)
    J = lib().oracle_det3(_d(F))
    fb = None if fibre is None else f64(fibre).reshape(3)
    rc = lib().oracle_stress_fibre(int(material), _d(p), _d(fb), _d(F), J, dt, _d(hist), _d(S))
    if rc:
        raise RuntimeError("oracle_stress error %d" % rc)
    return S, J


def shape_c3d8(xe):
    xe = f64(xe).reshape(24)
    dN = np.zeros(192)
    dJ = np.zeros(8)
    lib().oracle_shape_c3d8(_d(xe), _d(dN), _d(dJ))
    return dN.reshape(8, 8, 3), dJ


def hex_volume(xe):
    return lib().oracle_hex_volume(_d(f64(xe).reshape(24)))


FACES = np.array([0, 1, 2, 3, 4, 5, 6, 7, 0, 3, 7, 4, 1, 2, 6, 5, 0, 1, 5, 4, 3, 2, 6, 7], np.int32)


def face_area(xe, f):
    idx = i32(FACES[4 * f:4 * f + 4])
    return lib().oracle_face_area(_d(f64(xe).reshape(24)), _i(idx))


def char_length(xe):
    return lib().oracle_char_length(_d(f64(xe).reshape(24)))


def element_mass(xe, rho):
    m = np.zeros(8)
    lib().oracle_element_mass(_d(f64(xe).reshape(24)), rho, _d(m))
    return m


def motion_eval(motion, X, t):
    u, v, a = np.zeros(3), np.zeros(3), np.zeros(3)
    lib().oracle_motion_eval(_d(f64(motion)), _d(f64(X)), t, _d(u), _d(v), _d(a))
    return u, v, a


class Model:
    """Mesh-level oracle (FemTech's one-rank step semantics)."""

    def __init__(self, X, conn, pid, material_id, props, threads=1):
        self.X, self.conn, self.pid = f64(X), i32(conn), i32(pid)
        self.material_id, self.props = i32(material_id), f64(props)
        self.nn, self.ne = self.X.size // 3, self.conn.size // 8
        self.h = lib().oracle_create(self.nn, self.ne, _d(self.X), _i(self.conn), _i(self.pid),
                                     self.material_id.size, _i(self.material_id), _d(self.props))
        lib().oracle_set_threads(self.h, threads)

    def close(self):
        if self.h:
            lib().oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_threads(self, n):
        lib().oracle_set_threads(self.h, int(n))

    def shape_functions(self):
        dN = np.zeros(192 * self.ne)
        dJ = np.zeros(8 * self.ne)
        lib().oracle_shape_functions(self.h, _d(dN), _d(dJ))
        return dN, dJ

    def lumped_mass(self):
        m = np.zeros(3 * self.nn)
        lib().oracle_lumped_mass(self.h, _d(m))
        return m

    def get_force(self, u, dt=0.0):
        fi = np.zeros(3 * self.nn)
        rc = lib().oracle_get_force(self.h, _d(f64(u)), dt, _d(fi))
        if rc:
            raise RuntimeError("oracle error %d" % rc)
        return fi

    def element_forces(self, u, dt=0.0, update_history=False):
        fe = np.zeros(24 * self.ne)
        lib().oracle_element_forces(self.h, _d(f64(u)), dt, int(update_history), _d(fe))
        return fe

    def stable_dt(self, u, boundary):
        return lib().oracle_stable_dt(self.h, _d(f64(u)), _i(i32(boundary)))

    def element_strain(self, u):
        E = np.zeros(9 * self.ne)
        P = np.zeros(3 * self.ne)
        lib().oracle_element_strain(self.h, _d(f64(u)), _d(E), _d(P))
        return E, P

    def set_fibres(self, directions):
        lib().oracle_set_fibres(self.h, _d(f64(directions)))

    def injury_init(self, include_part=None):
        lib().oracle_injury_init(self.h, _i(None if include_part is None else i32(include_part)))

    def injury_update(self, u, time, dt):
        lib().oracle_injury_update(self.h, _d(f64(u)), float(time), float(dt))

    def injury_results(self):
        out = np.zeros(21)
        flags = np.zeros(self.ne, np.int32)
        lib().oracle_injury_results(self.h, _d(out), _i(flags))
        return out, flags

    def history(self):
        h = np.zeros(216 * self.ne)
        lib().oracle_get_history(self.h, _d(h))
        return h

    def set_bc(self, dof_motion, motions):
        m = f64(np.asarray(motions, np.float64).reshape(-1))
        lib().oracle_set_bc(self.h, _i(i32(dof_motion)), m.size // 16, _d(m))

    def set_initial(self, u0=None, v0=None):
        lib().oracle_set_initial(self.h, _d(None if u0 is None else f64(u0)),
                                 _d(None if v0 is None else f64(v0)))

    def run(self, max_steps, tmax=float("inf"), dt_factor=0.8, failure_dt=1e-11, snap_steps=()):
        nd = 3 * self.nn
        u, v, a, fi = (np.zeros(nd) for _ in range(4))
        ht = np.zeros(max_steps + 1)
        hd = np.zeros(max_steps + 1)
        he = np.zeros(4 * (max_steps + 1))
        snaps = i32(np.array(snap_steps, np.int32))
        su = np.zeros(max(1, len(snap_steps)) * nd)
        k = lib().oracle_run(self.h, int(max_steps), tmax, dt_factor, failure_dt, _d(u), _d(v), _d(a),
                             _d(fi), _d(ht), _d(hd), _d(he), len(snap_steps), _i(snaps), _d(su))
        if k < 0:
            raise RuntimeError("oracle run terminated with code %d" % -k)
        return dict(steps=k, u=u, v=v, a=a, fi=fi, time=ht[:k + 1], dt=hd[:k + 1],
                    energy=he[:4 * (k + 1)].reshape(k + 1, 4),
                    snap_u=su[:len(snap_steps) * nd].reshape(len(snap_steps), nd))
