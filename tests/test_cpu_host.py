"""CPU-side checks: the C ABI library loads and exports every declared entry
point, and the host mesh tools (generator, .inp/.k readers, materials.dat)
follow the reference's conventions."""
import ctypes
import os
import re

import numpy as np
import pytest

import femtech_amd as ft
from conftest import ROOT, golden
from oracle import oracle as O


def declared_symbols():
    names = set()
    for hdr in os.listdir(os.path.join(ROOT, "include")):
        if not hdr.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", hdr)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(ft_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(ft.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_device_count_without_gpu_is_harmless():
    assert ft.device_count() >= 0


def test_engine_create_without_device_fails_loudly():
    if ft.device_count() > 0:
        pytest.skip("a GPU is present")
    X, conn = ft.mesh_block(1, 1, 1, 1e-3)
    with pytest.raises(ft.FemTechError) as ei:
        ft.Engine(X, conn, np.zeros(1, np.int32), np.array([1], np.int32), np.zeros(9))
    assert ei.value.code == ft.ERR_HIP


def test_engine_arguments_checked_before_the_device():
    """Bad meshes are rejected with the reference's input code (3) or the
    argument code (51) on any machine, before a device is touched."""
    X, conn = ft.mesh_block(2, 1, 1, 1e-3)
    pid = np.zeros(2, np.int32)
    mid = np.array([1], np.int32)
    props = np.zeros(9)
    lib = ft.library()
    for bad, code in (
            (dict(conn=np.where(conn == conn.max(), 999, conn).astype(np.int32)), ft.FT_ERR_INPUT),
            (dict(pid=np.array([0, 3], np.int32)), ft.FT_ERR_INPUT),
            (dict(precision=7), ft.FT_ERR_ARG)):
        a = dict(conn=conn, pid=pid, precision=ft.FT_FP64)
        a.update(bad)
        with pytest.raises(ft.FemTechError) as ei:
            ft.Engine(X, a["conn"], a["pid"], mid, props, a["precision"])
        assert ei.value.code == code, (bad, ei.value)
    # an empty mesh
    h = ctypes.c_void_p()
    d = np.zeros(3)
    rc = lib.ft_engine_create(0, 0, d.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                              np.zeros(8, np.int32).ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                              np.zeros(1, np.int32).ctypes.data_as(ctypes.POINTER(ctypes.c_int)), 1,
                              mid.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                              props.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ft.FT_FP64, ctypes.byref(h))
    assert rc == ft.FT_ERR_ARG and not h.value


def test_mesh_size_limit():
    """Node counts past the 32-bit byte offsets of the kernels' gathers are
    rejected up front (the arrays are not read)."""
    lib = ft.library()
    h = ctypes.c_void_p()
    d = np.zeros(3)
    one = np.zeros(8, np.int32)
    mid = np.array([1], np.int32)
    props = np.zeros(9)
    rc = lib.ft_engine_create(100_000_000, 1, d.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                              one.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                              np.zeros(1, np.int32).ctypes.data_as(ctypes.POINTER(ctypes.c_int)), 1,
                              mid.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                              props.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ft.FT_FP64, ctypes.byref(h))
    assert rc == ft.FT_ERR_ARG and not h.value
    assert b"too large" in lib.ft_last_error()


def test_null_engine_is_an_argument_error():
    """Engine entry points reject a null handle with FT_ERR_ARG."""
    lib = ft.library()
    out = ctypes.c_int(0)
    d = ctypes.c_double(0.0)
    assert lib.ft_set_persistent(None, 1) == ft.FT_ERR_ARG
    assert lib.ft_set_graph(None, 4) == ft.FT_ERR_ARG
    assert lib.ft_run_steps(None, 1, ctypes.byref(out)) == ft.FT_ERR_ARG
    assert lib.ft_get_time(None, ctypes.byref(d), ctypes.byref(d), ctypes.byref(out)) == ft.FT_ERR_ARG
    assert lib.ft_autotune(None, 1, None, ctypes.byref(out)) == ft.FT_ERR_ARG


def test_network_without_device_fails_loudly():
    if ft.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(ft.FemTechError) as ei:
        ft.Mlp()
    assert ei.value.code == ft.ERR_HIP


def test_block_generator_geometry():
    X, conn = ft.mesh_block(4, 3, 2, (1e-3, 2e-3, 3e-3))
    X = X.reshape(-1, 3)
    assert X.shape == (5 * 4 * 3, 3)
    assert conn.size == 8 * 24
    # every element is a positive-volume box of the requested size
    for e in range(24):
        xe = X[conn[8 * e:8 * e + 8]]
        assert O.hex_volume(xe) == pytest.approx(6e-9, rel=1e-12)
    # node numbering x-fastest
    assert np.allclose(X[1], [1e-3, 0, 0]) and np.allclose(X[5], [0, 2e-3, 0])


def test_block_jitter_is_seeded_interior_only():
    X1, _ = ft.mesh_block(5, 5, 5, 1e-3, jitter=0.15, seed=42)
    X2, _ = ft.mesh_block(5, 5, 5, 1e-3, jitter=0.15, seed=42)
    X3, _ = ft.mesh_block(5, 5, 5, 1e-3, jitter=0.15, seed=43)
    X0, _ = ft.mesh_block(5, 5, 5, 1e-3)
    assert np.array_equal(X1, X2) and not np.array_equal(X1, X3)
    d = (X1 - X0).reshape(-1, 3)
    assert np.abs(d).max() <= 0.15e-3 + 1e-18
    Xr = X0.reshape(-1, 3)
    on_face = np.any((Xr < 1e-9) | (Xr > 5e-3 - 1e-9), axis=1)
    assert np.all(d[on_face] == 0) and np.all(np.abs(d[~on_face]).max(axis=1) > 0)


def test_inp_round_trip(tmp_path):
    X, conn = ft.mesh_block(3, 2, 2, 1e-3, jitter=0.1, seed=3)
    pid = np.array([0] * 4 + [1] * 4 + [2] * 4, np.int32)
    p = str(tmp_path / "m.inp")
    ft.write_inp(p, X, conn, pid)
    X2, conn2, pid2, gid = ft.read_mesh(p)
    assert np.array_equal(X, X2) and np.array_equal(conn, conn2) and np.array_equal(pid, pid2)
    assert np.array_equal(gid, np.arange(1, X.size // 3 + 1))


def test_reader_renumbers_like_partition_mesh(tmp_path):
    """Node ids need not be contiguous: local node i is the i-th smallest id
    used by the connectivity (PartitionMesh.cpp:489-530)."""
    p = tmp_path / "gap.inp"
    ids = [10, 3, 7, 40, 12, 5, 9, 1, 99]  # 99 is unused
    coords = {i: (0.001 * k, 0.002 * k, 0.003 * k) for k, i in enumerate(ids)}
    lines = ["*Node"] + ["%d, %r, %r, %r" % ((i,) + coords[i]) for i in ids]
    lines += ["*Element, type=C3D8, elset=A", "1, 10, 3, 7, 40, 12, 5, 9, 1"]
    p.write_text("\n".join(lines) + "\n")
    X, conn, pid, gid = ft.read_mesh(str(p))
    used = sorted([10, 3, 7, 40, 12, 5, 9, 1])
    assert list(gid) == used
    assert list(conn) == [used.index(i) for i in [10, 3, 7, 40, 12, 5, 9, 1]]
    assert np.allclose(X.reshape(-1, 3)[used.index(40)], coords[40])


def test_reader_skips_output_requests_and_assembly_nodes(tmp_path):
    """Abaqus/CAE decks (the 27-element benchmark deck is one) end with
    assembly reference nodes and "*Node Output" / "*Element Output"
    requests: only "*Node" / "*Node," and "*Element, type=..." open sections,
    and only the first node section is read (ReadAbaqus.cpp:41-49, 359-412)."""
    X, conn = ft.mesh_block(2, 1, 1, 1e-3)
    nn = X.size // 3
    Xr = X.reshape(-1, 3)
    lines = ["*Heading", "*Part, name=P", "*Node"]
    lines += ["%d, %r, %r, %r" % ((i + 1,) + tuple(float(x) for x in Xr[i])) for i in range(nn)]
    lines += ["*Element, type=C3D8"]
    lines += [", ".join(str(v) for v in [e + 1] + list(conn.reshape(-1, 8)[e] + 1)) for e in range(2)]
    lines += ["*End Part", "*Assembly, name=A", "*Node", "      1, 0., 0.5, 0.25", "*Nset, nset=x",
              " 1, 2", "*End Assembly", "*Step, name=S", "*Element Output, directions=YES",
              "LE, S", "*Node Output", "RF, U", "*End Step"]
    p = tmp_path / "cae.inp"
    p.write_text("\n".join(lines) + "\n")
    X2, conn2, pid2, gid = ft.read_mesh(str(p))
    assert np.array_equal(X2, X) and np.array_equal(conn2, conn) and list(pid2) == [0, 0]


def test_reader_matches_reference_on_its_meshes():
    for name in ("ex9", "bench10"):
        z = golden("run_" + name)
        assert np.array_equal(z["X"], z["ref_coords"])
        assert np.array_equal(z["conn"], z["ref_conn"])


def test_read_materials(tmp_path):
    p = tmp_path / "materials.dat"
    p.write_text("1 5 1000 1 2 3 4 5 6 7 8\n0 1 1040 100.0 100.00\n\n# comment\n")
    mat, props = ft.read_materials(str(p), 2)
    assert list(mat) == [1, 5]
    assert np.allclose(props.reshape(2, 9)[0, :3], [1040, 100, 100])
    assert np.allclose(props.reshape(2, 9)[1], [1000, 1, 2, 3, 4, 5, 6, 7, 8])
    p.write_text("0 9 1 2 3\n")
    with pytest.raises(ft.FemTechError) as ei:
        ft.read_materials(str(p), 1)
    assert ei.value.code == ft.ERR_INPUT
    p.write_text("0 1 1 2 3\n")
    with pytest.raises(ft.FemTechError):
        ft.read_materials(str(p), 2)  # part 1 missing


@pytest.mark.parametrize("n", [8, 24])
def test_ellipsoid_generator(n):
    """Mapped-cube ellipsoid: every Gauss-point Jacobian positive, smallest
    element not degenerate, volume converging to 4/3 pi abc, surface nodes on
    the ellipsoid, regions by normalised radius."""
    from femtech_amd import workloads as W
    a, b, c = 0.09, 0.075, 0.065
    X, conn, r = ft.mesh_ellipsoid(n, a, b, c)
    ne = conn.size // 8
    assert X.size == 3 * (n + 1) ** 3 and ne == n ** 3
    m = O.Model(X, conn, np.zeros(ne, np.int32), np.array([1], np.int32),
                np.array([[1000.0, 1.0, 1.0, 0, 0, 0, 0, 0, 0]]))
    _, dJ = m.shape_functions()
    dJ = dJ.reshape(-1, 8)
    vol = dJ.sum(axis=1)
    assert dJ.min() > 0 and vol.min() > 0.3 * vol.mean()
    exact = 4.0 / 3.0 * np.pi * a * b * c
    assert abs(vol.sum() / exact - 1) < 2.0 / n ** 2
    w = W.ellipsoid_head(n=n)
    Xr = w["X"].reshape(-1, 3)
    surf = w["dof_motion"].reshape(-1, 3)[:, 0] == 0
    np.testing.assert_allclose(((Xr[surf] / np.array([a, b, c])) ** 2).sum(axis=1), 1.0, atol=1e-14)
    assert np.all((r >= 0) & (r <= 1.0 + 1e-12))
    assert set(np.unique(w["pid"])) <= {0, 1, 2}
    np.testing.assert_array_equal(w["pid"], np.where(r > 0.93, 0, np.where(r > 0.3, 1, 2)))


def test_batch_generator():
    """The batched workload: log-uniform sizes within range (n^3 elements),
    instance-local renumbering, moduli within +-30 %, pulses within their
    ranges, one rotational motion per instance acting on its own nodes."""
    from femtech_amd import workloads as W
    m, no, eo, tmax, info = W.batch_instances(count=24)
    ne = np.diff(eo)
    assert len(info) == 24 and np.all(tmax == 0.020)
    assert np.all(ne == np.array([i[1] ** 3 for i in info]))
    assert ne.min() >= 12 ** 3 and ne.max() <= 44 ** 3
    assert {i[0] for i in info} == {"block", "ellipsoid"}
    conn = m["conn"].reshape(-1, 8)
    for k in range(24):
        c = conn[eo[k]:eo[k + 1]]
        assert c.min() >= no[k] and c.max() < no[k + 1]
        d = m["dof_motion"][3 * no[k]:3 * no[k + 1]]
        assert set(np.unique(d)) == {-1, k}
        mo = m["motions"][k]
        assert mo[0] == ft.MOTION_ROT_HALFSINE and 2000 <= mo[7] <= 8000 and 0.005 <= mo[8] <= 0.020
    props = m["props"]
    mid = m["material_id"]
    nh = props[mid == 1]
    assert np.all((nh[:, 1] >= 0.7e6) & (nh[:, 1] <= 1.3e6))
    K = nh[:, 2] + 2 * nh[:, 1] / 3
    assert np.all((K >= 7e6 * (1 - 1e-12)) & (K <= 13e6 * (1 + 1e-12)))
    og = props[mid == 7]
    assert np.all((og[:, 1] >= 0.7 * 530) & (og[:, 1] <= 1.3 * 530)) and np.all(og[:, 3] == -4.4)
    hgo = props[mid == 8]
    assert np.all((hgo[:, 3] >= 700) & (hgo[:, 3] <= 1300)) and np.all(hgo[:, 5] == 0.1)
