"""ft_energy_terms_noalloc (device arrays, static scratch, no allocation)
against ft_energy_terms (host arrays) through the C ABI, on two calls in a
row on one GPU, the second on a smaller mesh.  Both launch the same kernels
with the same fixed grid, so the sums agree bitwise.  The device inputs are
allocated here by the test through the HIP runtime the engine itself uses."""
import ctypes

import numpy as np
import pytest

import femtech_amd as ft
from conftest import golden

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        ft.library()  # loads libamdhip64 as a dependency of the engine
        h = ctypes.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipDeviceSynchronize.argtypes = []
        _hip = h
    return _hip


class DeviceArray:
    def __init__(self, host):
        self.host = np.ascontiguousarray(host)
        self.ptr = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(self.ptr), self.host.nbytes) == 0
        assert hip().hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, 1) == 0  # H2D

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None


def energy_inputs(nn, seed):
    rng = np.random.default_rng(seed)
    nd = 3 * nn
    arrs = [rng.uniform(-1e-3, 1e-3, nd) for _ in range(9)]  # u, u_prev, v, a, a_prev, fi, fi_prev, fe, fe_prev
    boundary = (rng.uniform(size=nd) < 0.2).astype(np.int32)
    return arrs, boundary


def both_ways(mesh, seed):
    X, conn = mesh
    nn, ne = X.size // 3, conn.size // 8
    props = np.array([1000.0, 1.0e6, 9.3333e6, 0, 0, 0, 0, 0, 0])
    arrs, boundary = energy_inputs(nn, seed)
    L = ft.library()
    D = ctypes.POINTER(ctypes.c_double)
    out_host = np.zeros(3)
    out_dev = np.full(3, np.nan)
    with ft.Engine(X, conn, np.zeros(ne, np.int32), np.array([1], np.int32), props, ft.FT_FP64) as e:
        e.lumped_mass()
        ptrs = [a.ctypes.data_as(D) for a in arrs]
        rc = L.ft_energy_terms(e._h, *ptrs, boundary.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                               out_host.ctypes.data_as(D))
        assert rc == 0, L.ft_last_error()
        dev = [DeviceArray(a) for a in arrs] + [DeviceArray(boundary)]
        try:
            fn = L.ft_energy_terms_noalloc
            fn.argtypes = [ctypes.c_void_p, ctypes.c_long] + [ctypes.c_void_p] * 10 + [D]
            fn.restype = ctypes.c_int
            rc = fn(e._h, 3 * nn, *[d.ptr for d in dev], out_dev.ctypes.data_as(D))
            assert rc == 0, L.ft_last_error()
            # a wrong DOF count is refused before any launch
            assert fn(e._h, 3 * nn - 3, *[d.ptr for d in dev], out_dev.ctypes.data_as(D)) == ft.FT_ERR_ARG
        finally:
            for d in dev:
                d.free()
    return out_host, out_dev


def test_noalloc_energy_terms_equal_host_path_twice(gpu_ready):
    z = golden("run_impact")
    big = (z["X"], z["conn"])
    small = ft.mesh_block(3, 2, 2, 1e-3, jitter=0.1, seed=3)
    assert small[0].size < big[0].size
    for mesh, seed in ((big, 1), (small, 2)):
        print("noalloc energy terms: %d nodes" % (mesh[0].size // 3), flush=True)
        host, dev = both_ways(mesh, seed)
        assert np.all(np.isfinite(host))
        np.testing.assert_array_equal(dev, host)
