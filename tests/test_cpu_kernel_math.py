"""The element kernel's math, built for the host from the same header the
HIP kernel compiles (femtech_amd/csrc/ft_device.h: element_eval), against the
oracle on the golden meshes: forces for every material, the stable-step
length, and multi-step Prony history."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from conftest import HERE, golden
from oracle import oracle as O

SRC = os.path.join(HERE, "native", "kernel_math_host.cpp")
OUT = os.path.join(HERE, "native", "_build", "libkm.so")
_D = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(scope="module")
def km():
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    subprocess.run(["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "-fPIC", "-shared", "-x", "c++",
                    "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", SRC, "-o", OUT], check=True)
    lib = ctypes.CDLL(OUT)
    lib.km_element.argtypes = [ctypes.c_int, _D, _D, ctypes.c_int, _D, ctypes.c_double, _D,
                               ctypes.c_int, _D, _D, _D]
    return lib


def dp(a):
    return None if a is None else a.ctypes.data_as(_D)


def elem(km, fp32, x24, u24, mat, p9, dt=1e-5, hist=None, update=1, fibre=None):
    f = np.zeros(24)
    dte = np.zeros(1)
    fb = None if fibre is None else np.ascontiguousarray(fibre, np.float64)
    flags = km.km_element(int(fp32), dp(np.ascontiguousarray(x24)), dp(np.ascontiguousarray(u24)),
                          int(mat), dp(np.ascontiguousarray(p9)), dt, dp(hist), update, dp(f), dp(dte), dp(fb))
    return f, dte[0], flags


def field(X, scale, seed):
    rng = np.random.default_rng(seed)
    return (scale * 1e-3 * (np.sin(X / 3e-3) + 0.3 * rng.uniform(-1, 1, X.shape))).reshape(-1)


@pytest.mark.parametrize("name", ["bench10", "impact", "materials"])
@pytest.mark.parametrize("fp32", [0, 1])
def test_element_forces_and_dt(km, name, fp32):
    z = golden("run_" + name)
    X = z["X"].reshape(-1, 3)
    conn = z["conn"].reshape(-1, 8)
    props = z["props"].reshape(-1, 9)
    mats = z["material_id"][z["pid"]]
    o = O.Model(z["X"], z["conn"], z["pid"], z["material_id"], z["props"])
    for scale, seed in ((1e-3, 1), (5e-2, 2)):
        u = field(X, scale, seed)
        fo = o.element_forces(u, 1e-5).reshape(-1, 24)
        num = den = 0.0
        worst = 0.0
        for e in range(len(conn)):
            x24 = X[conn[e]].reshape(24)
            u24 = u.reshape(-1, 3)[conn[e]].reshape(24)
            f, dte, flags = elem(km, fp32, x24, u24, mats[e], props[z["pid"][e]], hist=np.zeros(144))
            assert flags == 0
            num += np.sum((f - fo[e]) ** 2)
            den += np.sum(fo[e] ** 2)
            if mats[e] != 0:
                p = props[z["pid"][e]]
                nu = 0.5 * p[2] / (p[2] + p[1])
                ref_dt = O.char_length(x24 + u24) / np.sqrt(p[2] * (1 / nu - 1) / p[0])
                worst = max(worst, abs(dte - ref_dt) / ref_dt)
        err = np.sqrt(num / den)
        assert err < (1e-6 if fp32 else 1e-12), (name, scale, err)
        assert worst < (1e-6 if fp32 else 1e-13), (name, scale, worst)


def test_prony_history_multi_step(km):
    """Material 5 history over a deformation path, against the oracle's
    HGOIsotropicViscoelastic restatement (pinned by the stress KATs)."""
    z = golden("run_materials")
    X = z["X"].reshape(-1, 3)
    conn = z["conn"].reshape(-1, 8)
    e = int(np.nonzero(z["material_id"][z["pid"]] == 5)[0][0])
    p = z["props"].reshape(-1, 9)[z["pid"][e]]
    x24 = X[conn[e]].reshape(24)
    o = O.Model(x24, np.arange(8, dtype=np.int32), np.zeros(1, np.int32), np.array([5], np.int32), p)
    hist = np.zeros(144)
    for k in range(8):
        u24 = field(x24.reshape(8, 3), 2e-2 * (k + 1), 30 + k)
        f, _, flags = elem(km, 0, x24, u24, 5, p, dt=3e-5, hist=hist)
        fo = o.element_forces(u24, 3e-5, update_history=True)
        assert np.linalg.norm(f - fo) <= 1e-12 * np.linalg.norm(fo)
    ho = o.history().reshape(8, 3, 9)[..., [0, 4, 8, 7, 6, 3]]
    assert np.abs(hist.reshape(8, 3, 6) - ho).max() <= 1e-12 * np.abs(ho).max()


def test_rigid_rotation_gives_zero_force(km):
    """Objectivity: a finite rigid rotation produces no internal force."""
    z = golden("run_materials")
    X = z["X"].reshape(-1, 3)
    conn = z["conn"].reshape(-1, 8)
    c, s = np.cos(0.7), np.sin(0.7)
    R = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])
    for mat, row in ((1, 0), (2, 1), (4, 3), (5, 4)):
        e = int(np.nonzero(z["pid"] == row)[0][0])
        x = X[conn[e]]
        u = (x @ R.T - x).reshape(24)
        p = z["props"].reshape(-1, 9)[row]
        f, _, _ = elem(km, 0, x.reshape(24), u, mat, p, hist=np.zeros(144))
        f_ref, _, _ = elem(km, 0, x.reshape(24), 1e-3 * u, mat, p, hist=np.zeros(144))
        assert np.abs(f).max() <= 1e-9 * max(np.abs(f_ref).max(), p[1] * 1e-6)


def new_law_block():
    """3x3x3 jittered block in three parts: 1-term Ogden (alpha = -4.4),
    Ogden + Prony, anisotropic HGO (kappa = 0.1, seeded fibres)."""
    import femtech_amd as ft
    X, conn = ft.mesh_block(3, 3, 3, 1e-3, jitter=0.15, seed=17)
    ne = conn.size // 8
    pid = (np.arange(ne) % 3).astype(np.int32)
    mu, lam = 2.0e4, 2.0e5
    props = np.zeros((3, 9))
    props[0, :4] = [1040.0, mu, lam, -4.4]
    props[1, :9] = [1040.0, mu, lam, 3.0, 0.0, 0.5, 2e-3, 0.3, 2e-2]
    props[2, :6] = [1000.0, mu, lam, 3.0e4, 8.0, 0.1]
    rng = np.random.default_rng(5)
    fib = rng.normal(size=(ne, 3))
    fib /= np.linalg.norm(fib, axis=1)[:, None]
    return X, conn, pid, np.array([6, 7, 8], np.int32), props, fib


@pytest.mark.parametrize("fp32", [0, 1])
def test_new_laws_element_forces(km, fp32):
    """Materials 6-8 (no reference counterpart) in the element kernel's math
    against the oracle's independent restatement (itself pinned by
    reductions, energy derivatives and objectivity)."""
    X, conn, pid, mid, props, fib = new_law_block()
    o = O.Model(X, conn, pid, mid, props)
    o.set_fibres(fib)
    Xr, cr = X.reshape(-1, 3), conn.reshape(-1, 8)
    for scale, seed in ((1e-3, 3), (8e-2, 4)):
        u = field(Xr, scale, seed)
        fo = o.element_forces(u, 1e-5).reshape(-1, 24)
        num = den = 0.0
        for e in range(len(cr)):
            f, _, flags = elem(km, fp32, Xr[cr[e]].reshape(24), u.reshape(-1, 3)[cr[e]].reshape(24),
                               mid[pid[e]], props[pid[e]], hist=np.zeros(144), fibre=fib[e])
            assert flags == 0
            num += np.sum((f - fo[e]) ** 2)
            den += np.sum(fo[e] ** 2)
        err = np.sqrt(num / den)
        assert err < (1e-6 if fp32 else 1e-11), (scale, err)


@pytest.mark.parametrize("ne", [1, 63, 64, 65, 200, 1728])
def test_history_layout_blocks(km, ne):
    """The Prony history layout the element kernel loads and stores: every
    (component, element) pair has its own slot inside the allocation; the 64
    elements of a block read one contiguous run per component; a block's 144
    runs fill one contiguous range of 144 x 64 slots."""
    km.km_hist_index.restype = ctypes.c_longlong
    km.km_hist_len.restype = ctypes.c_longlong
    n = km.km_hist_len(ne)
    assert n == 144 * (-(-ne // 64) * 64)
    idx = np.array([[km.km_hist_index(c, e) for c in range(144)] for e in range(ne)])
    assert idx.min() == 0 and idx.max() < n
    assert len(np.unique(idx)) == idx.size
    for b in range(0, ne, 64):
        blk = idx[b:b + 64]
        assert np.all(np.diff(blk, axis=0) == 1)  # consecutive elements, one run per component
        assert blk.min() == 144 * b and blk.max() < 144 * (b + 64)
