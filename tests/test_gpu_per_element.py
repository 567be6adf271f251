"""FemTech's per-element, per-Gauss-point routines on the GPU.

1. The device kernels behind them (ft_point_eval) against the reference
   build's known answers: StressUpdate with every material, including
   six-step Prony histories (kat_stress, from StressUpdate.cpp:5-29 compiled
   unmodified), the characteristic length of distorted hexes (kat_geom,
   CalculateCharacteristicLength_C3D8.cpp:3-30), and the shape-function
   gradients and Jacobians ShapeFunctions leaves in dshp and detJacobian
   (every golden run).
2. The routines of libFemTech.so themselves, called by a C++ driver one
   element and Gauss point at a time the way GetForce_3D.cpp:15-44 calls
   them, against the oracle and against the library's fused GetForce_3D,
   with the Prony history carried through the host arrays (and handed to the
   device when a fused pass follows).

Bars: the FP64 reference values to 1e-12 relative (stresses below 1 %
strain: 1e-12 of the stress of a 1 % strain; 1e-13 for geometry and shape
functions); element and nodal forces to 1e-11, histories to 1e-10.
"""
import os
import subprocess

import numpy as np
import pytest

import femtech_amd as ft
from conftest import HERE, ROOT, golden, rel
from oracle import oracle as O
from test_compat import LIB_DIR, write_case

pytestmark = pytest.mark.gpu

GP = 0.577350269189626
CHI = np.array([[-GP, -GP, GP], [GP, -GP, GP], [GP, GP, GP], [-GP, GP, GP],
                [-GP, -GP, -GP], [GP, -GP, -GP], [GP, GP, -GP], [-GP, GP, -GP]])


def test_point_stress_matches_reference_known_answers(gpu_ready):
    z = golden("kat_stress")
    n, T = z["F"].shape[:2]
    hist = np.zeros((n, 34))
    for t in range(T):
        rec = np.concatenate([z["F"][:, t, :], z["material"][:, None].astype(float), z["props"],
                              z["dt"][:, None]], axis=1)
        out = ft.point_eval(ft.PT_STRESS, rec, hist)
        # relative to the stress, or below 1 % strain to the stress of a 1 %
        # strain: the device forms the laws from H = F - I and keeps the digits
        # the reference's F^T F - I loses at small strain
        stiff = z["props"][:, 1] + z["props"][:, 2] + z["props"][:, 3]
        scale = np.maximum(np.abs(z["ref_pk2"][:, t]).max(axis=1), 1e-2 * stiff)[:, None]
        assert (np.abs(out[:, :6] - z["ref_pk2"][:, t]) / scale).max() < 1e-12, t
        detF = ft.point_eval(ft.PT_DETF, z["F"][:, t, :])[:, 0]
        np.testing.assert_allclose(detF, z["ref_detF"][:, t], rtol=1e-13)
        hist = out
    visco = z["material"] == 5
    assert visco.any()
    h = hist[visco, 6:33]
    ref = z["ref_hist"][visco]
    assert rel(h, ref) < 1e-10


def test_point_geometry_matches_reference_known_answers(gpu_ready):
    z = golden("kat_geom")
    le = ft.point_eval(ft.PT_CHARLEN, z["xe"])[:, 0]
    np.testing.assert_allclose(le, z["ref_char_length"], rtol=1e-13)


@pytest.mark.parametrize("name", ["ex9", "bench10", "c1", "impact", "materials"])
def test_point_shape_functions_match_reference(gpu_ready, name):
    z = golden("run_" + name)
    X = z["ref_coords"].reshape(-1, 3)
    conn = z["ref_conn"].reshape(-1, 8)
    xe = X[conn].reshape(-1, 24)
    rec = np.concatenate([np.tile(CHI, (len(xe), 1)), np.repeat(xe, 8, axis=0)], axis=1)
    out = ft.point_eval(ft.PT_SHAPE, rec)
    assert rel(out[:, 8:32].reshape(-1), z["ref_dshp"]) < 1e-13
    assert rel(out[:, 32], z["ref_detJ"]) < 1e-13
    np.testing.assert_allclose(out[:, :8].sum(axis=1), 1.0, rtol=1e-15)


# ---- the FemTech surface, through a C++ driver --------------------------------

def driver(tmp_path_factory):
    d = tmp_path_factory.mktemp("per_element")
    exe = str(d / "per_element_driver")
    res = subprocess.run(["g++", "-O1", "-std=c++17", "-w", "-I", os.path.join(ROOT, "include"),
                          os.path.join(HERE, "native", "per_element_driver.cpp"), "-L", LIB_DIR, "-lFemTech",
                          "-lfemtech_amd", "-Wl,-rpath," + LIB_DIR, "-o", exe], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-2000:]
    return exe


@pytest.fixture(scope="module")
def per_element_exe(tmp_path_factory):
    return driver(tmp_path_factory)


def field(z, scale, seed):
    rng = np.random.default_rng(seed)
    X = z["X"].reshape(-1, 3)
    return (scale * 1e-3 * (np.sin(X / 3e-3) + 0.3 * rng.uniform(-1, 1, X.shape))).reshape(-1)


def run(exe, tmp_path, z, mode, dt, us):
    os.makedirs(str(tmp_path), exist_ok=True)
    write_case(str(tmp_path), z)
    paths = []
    for k, u in enumerate(us):
        p = str(tmp_path / ("u%d.bin" % k))
        np.asarray(u, np.float64).tofile(p)
        paths.append(p)
    print("per-element driver:", mode, flush=True)
    res = subprocess.run([exe, "mesh.inp", mode, repr(dt)] + paths, cwd=str(tmp_path), capture_output=True,
                         text=True, env=dict(os.environ, FEMTECH_UID="t"), timeout=300)
    assert res.returncode == 0, (res.returncode, res.stdout[-2000:], res.stderr[-2000:])
    out = {}
    for f in os.listdir(tmp_path):
        if f.endswith(".f64"):
            out[f[:-4]] = np.fromfile(str(tmp_path / f))
    return out


def oracle_of(z):
    return O.Model(z["X"], z["conn"], z["pid"], z["material_id"], z["props"])


def permuted(z, seed):
    """The fixture's mesh with every element's local nodes renumbered by a
    random cube rotation (the fused pass then takes the node-slot layout)."""
    from cases import permute_elements
    d = {k: z[k] for k in z.files}
    d["conn"] = permute_elements(z["conn"], seed)
    return d


@pytest.mark.parametrize("name", ["impact", "materials", "impact_permuted", "materials_permuted"])
def test_per_point_loop_matches_oracle_and_fused_pass(gpu_ready, per_element_exe, tmp_path, name):
    base = name.replace("_permuted", "")
    z = golden("run_" + base)
    if name != base:
        z = permuted(z, 7)
    else:
        z = {k: z[k] for k in z.files}
    dt = 1e-5
    u = field(z, 2e-2, 3)
    man = run(per_element_exe, tmp_path / "m", z, "manual", dt, [u])
    fus = run(per_element_exe, tmp_path / "f", z, "fused", dt, [u])
    o = oracle_of(z)
    fo = o.get_force(u, dt)
    assert rel(man["fi"], fo) < 1e-11
    assert rel(fus["fi"], fo) < 1e-11
    if name == base:
        # ShapeFunctions' arrays equal the reference's
        assert rel(man["dshp"], z["ref_dshp"]) < 1e-13
        assert rel(man["detJ"], z["ref_detJ"]) < 1e-13
    np.testing.assert_array_equal(man["weights"], 1.0)
    # the per-point fields equal the fused pass's
    assert rel(man["F"], fus["F"]) < 1e-12
    assert rel(man["detF"], fus["detF"]) < 1e-12
    assert rel(man["pk2"], fus["pk2"]) < 1e-12
    # element time steps and characteristic lengths against the oracle's geometry
    X = z["X"].reshape(-1, 3)
    x = (X + u.reshape(-1, 3))[z["conn"].reshape(-1, 8)]
    le = np.array([O.char_length(xe) for xe in x])
    np.testing.assert_allclose(man["le"], le, rtol=1e-13)
    np.testing.assert_array_equal(man["le8"], man["le"])
    props = np.asarray(z["props"]).reshape(-1, 9)[z["pid"]]
    rho, mu, lam = props[:, 0], props[:, 1], props[:, 2]
    with np.errstate(invalid="ignore", divide="ignore"):  # rigid part: 0 / 0, as in the reference
        nu = 0.5 * lam / (lam + mu)
        ce = np.sqrt(lam * (1.0 / nu - 1.0) / rho)
    np.testing.assert_allclose(man["dte"], le / ce, rtol=1e-13)
    # InverseF and inverse3x3Matrix of Gauss point 0 against numpy
    F0 = man["F"].reshape(-1, 8, 9)[:, 0, :]
    A = F0.reshape(-1, 3, 3)  # rows = consecutive triples (InverseF.cpp)
    np.testing.assert_allclose(man["invF0"].reshape(-1, 3, 3), np.linalg.inv(A), rtol=1e-12, atol=1e-14)
    M = F0.reshape(-1, 3, 3).transpose(0, 2, 1)  # column-major (math.cpp:9)
    inv0 = man["inv0"].reshape(-1, 10)
    np.testing.assert_allclose(inv0[:, :9].reshape(-1, 3, 3).transpose(0, 2, 1), np.linalg.inv(M), rtol=1e-12,
                               atol=1e-14)
    np.testing.assert_allclose(inv0[:, 9], np.linalg.det(M), rtol=1e-13)
    # B of node 0 at Gauss point 0, contracted with a virtual displacement,
    # gives the rate of Green strain: B du = sym(F^T grad du) (Voigt,
    # engineering shears)
    dn = man["dshp"].reshape(-1, 8, 8, 3)[:, 0, 0, :]
    B0 = man["B0"].reshape(-1, 3, 6).transpose(0, 2, 1)  # [e][row][component]
    du = np.array([0.3, -0.7, 0.2])
    G = np.einsum("i,ej->eij", du, dn)  # grad du for a field du N_0
    Fm = F0.reshape(-1, 3, 3).transpose(0, 2, 1)
    E = 0.5 * (np.einsum("eki,ekj->eij", Fm, G) + np.einsum("eki,ekj->eij", G, Fm))
    voigt = np.stack([E[:, 0, 0], E[:, 1, 1], E[:, 2, 2], 2 * E[:, 1, 2], 2 * E[:, 0, 2], 2 * E[:, 0, 1]], 1)
    np.testing.assert_allclose(B0 @ du, voigt, rtol=1e-12, atol=1e-12 * np.abs(voigt).max())


@pytest.mark.parametrize("mode", ["mixed", "twice"])
def test_prony_history_carried_through_host_arrays(gpu_ready, per_element_exe, tmp_path, mode):
    """Material 5 through StressUpdate on the host arrays, then either a fused
    GetForce_3D (the history is handed to the device first) or a second host
    pass: both equal the oracle run through the same two force evaluations."""
    z = golden("run_materials")
    dt = 2e-5
    u1, u2 = field(z, 1e-2, 5), field(z, 3e-2, 6)
    out = run(per_element_exe, tmp_path, z, mode, dt, [u1, u2])
    o = oracle_of(z)
    o.element_forces(u1, dt, update_history=True)
    fo = o.get_force(u2, dt)
    assert rel(out["fi"], fo) < 1e-11
    ho = o.history().reshape(-1, 8, 3, 9)
    h = out["hist"].reshape(3, -1, 8, 9).transpose(1, 2, 0, 3)
    mask = z["material_id"][z["pid"]] == 5
    assert rel(h[mask], ho[mask]) < 1e-10
