"""The oracle against the reference build's golden vectors (tests/golden)."""
import numpy as np
import pytest

import cases
from conftest import golden, rel
from oracle import oracle as O

RUNS = ["ex9", "bench10", "c1", "impact", "materials"]


@pytest.mark.parametrize("name", RUNS)
def test_setup_matches_reference(name):
    z = golden("run_" + name)
    m = O.Model(z["X"], z["conn"], z["pid"], z["material_id"], z["props"])
    dN, dJ = m.shape_functions()
    # ShapeFunctions.cpp:207-228 and AssembleLumpedMass (Mass3D.cpp:127-157)
    assert np.array_equal(z["X"], z["ref_coords"])
    assert np.array_equal(z["conn"], z["ref_conn"])
    assert rel(dN, z["ref_dshp"]) < 1e-14
    assert rel(dJ, z["ref_detJ"]) < 1e-14
    assert rel(m.lumped_mass(), z["ref_mass"]) < 1e-14


@pytest.mark.parametrize("name", RUNS)
def test_run_matches_reference(name):
    z = golden("run_" + name)
    m = O.Model(z["X"], z["conn"], z["pid"], z["material_id"], z["props"])
    m.set_bc(z["dof_motion"], z["motions"])
    m.set_initial(None, z["v0"] if "v0" in z else None)
    r = m.run(int(z["max_steps"]), float(z["tmax"]), float(z["dt_factor"]), float(z["failure_dt"]),
              list(z["snaps"]))
    assert r["steps"] == int(z["ref_steps"])
    # FP64 restatement vs FP64 reference: round-off only
    assert rel(r["u"], z["ref_u"]) < 1e-12
    assert rel(r["snap_u"], z["ref_snap_u"]) < 1e-12
    assert rel(r["v"], z["ref_v"]) < 1e-10
    assert rel(r["a"], z["ref_a"]) < 1e-9
    assert rel(r["fi"], z["ref_fi"]) < 1e-10
    assert rel(r["time"], z["ref_time"]) < 1e-14
    assert rel(r["dt"], z["ref_dt"]) < 1e-13
    assert rel(r["energy"], z["ref_energy"]) < 1e-12
    # CheckEnergy's own rows (printed with 7 significant digits)
    ef = z["ref_energy_file"]
    E = r["energy"][1:]
    for col, k in ((1, 0), (2, 1), (3, 2)):
        scale = max(np.abs(ef[:, col]).max(), 1e-300)
        assert np.abs(E[:, k] - ef[:, col]).max() <= 1e-6 * scale + 1e-30


def test_initial_force_matches_reference():
    for name in RUNS:
        z = golden("run_" + name)
        m = O.Model(z["X"], z["conn"], z["pid"], z["material_id"], z["props"])
        m.set_bc(z["dof_motion"], z["motions"])
        m.set_initial(None, z["v0"] if "v0" in z else None)
        r = m.run(0, float(z["tmax"]), float(z["dt_factor"]), float(z["failure_dt"]))
        assert rel(r["fi"], z["ref_fi0"]) < 1e-12
        assert rel(r["a"], z["ref_a0"]) < 1e-12


def test_stress_kats():
    z = golden("kat_stress")
    for i in range(len(z["material"])):
        hist = np.zeros(27)
        for t in range(z["F"].shape[1]):
            S, J = O.stress(int(z["material"][i]), z["props"][i], z["F"][i, t], float(z["dt"][i]), hist)
            ref = z["ref_pk2"][i, t]
            assert abs(J - z["ref_detF"][i, t]) <= 1e-15 * abs(J)
            scale = max(np.abs(ref).max(), 1e-12 * float(z["props"][i][1]))
            assert np.abs(S - ref).max() <= 1e-12 * scale, (i, t, S, ref)
        assert np.abs(hist - z["ref_hist"][i]).max() <= 1e-12 * max(np.abs(z["ref_hist"][i]).max(), 1.0)


def test_geometry_kats():
    z = golden("kat_geom")
    for i, xe in enumerate(z["xe"]):
        assert O.hex_volume(xe) == pytest.approx(z["ref_volume"][i], rel=1e-14)
        for f in range(6):
            assert O.face_area(xe, f) == pytest.approx(z["ref_area"][i, f], rel=1e-14)
        assert O.char_length(xe) == pytest.approx(z["ref_char_length"][i], rel=1e-14)


def test_geometry_kats_cover_shortcut_branch():
    """Some faces take the parallelogram shortcut while not being
    parallelograms (negative twist, Geometry.cpp:46 has no fabs)."""
    z = golden("kat_geom")
    twisted = 0
    for xe in z["xe"]:
        x = xe.reshape(8, 3)
        for f in range(6):
            p = x[O.FACES[4 * f:4 * f + 4]]
            cD = 0.25 * (p[0] - p[1] + p[2] - p[3])
            if np.all(cD < 1e-6) and np.abs(cD).max() > 1e-6:
                twisted += 1
    assert twisted > 20


def test_ex9_against_abaqus_report():
    """The reference CI check (examples/ex9/compareResults.py:14-28): final
    time within 1 %, displacement of the loaded corner within 5 %."""
    z = golden("run_ex9")
    rpt = z["abaqus_rpt"]
    m = O.Model(z["X"], z["conn"], z["pid"], z["material_id"], z["props"])
    m.set_bc(z["dof_motion"], z["motions"])
    r = m.run(int(z["max_steps"]), float(z["tmax"]), float(z["dt_factor"]))
    X = z["X"].reshape(-1, 3)
    node = int(np.nonzero(np.all(np.abs(X - 0.005) < 1e-5, axis=1))[0][0])
    u = r["u"][3 * node:3 * node + 3]
    t_ab, u_ab = rpt[-1, 0], rpt[-1, 9:12]  # columns X, A(4), RF(4), U mag, U1..U3
    assert abs(r["time"][-1] - t_ab) / t_ab < 0.01
    assert np.max(np.abs(u - rpt[-1, 10:13]) / np.abs(rpt[-1, 10:13])) < 0.05


def test_ex7_free_dofs_follow_the_ex9_step():
    """ex7's loop (ex7.cpp:56-125) differs from ex9's only on constrained
    DOFs (velocities stay zero, the previous-step force arrays are never
    saved): the displacement history of the reference's ex7 run equals the
    oracle's ex9-style run of the same loading."""
    z = golden("run_ex7")
    c = cases.case_ex9_like(z["X"], z["conn"], z["pid"], z["material_id"], z["props"], L=float(z["length"]),
                            dmax=float(z["dmax"]), tmax=float(z["tmax"]))
    m = O.Model(z["X"], z["conn"], z["pid"], z["material_id"], z["props"])
    m.set_bc(c["dof_motion"], c["motions"])
    pe = int(z["ref_plot_every"])
    snaps = list(range(pe, int(z["ref_steps"]) + 1, pe))
    r = m.run(100000, float(z["tmax"]), 0.8, 1e-11, snaps)
    assert r["steps"] == int(z["ref_steps"])
    assert rel(r["u"], z["ref_u"]) < 1e-12
    assert rel(r["time"], z["ref_time"]) < 1e-14 and rel(r["dt"], z["ref_dt"]) < 1e-13
    X = z["X"].reshape(-1, 3)
    node = int(np.nonzero(np.all(np.abs(X - 0.005) < 1e-5, axis=1))[0][0])
    assert rel(r["snap_u"][:, 3 * node:3 * node + 3], z["ref_corner"][1:, 1:]) < 1e-12
    free = c["dof_motion"] < 0
    assert rel(r["v"][free], z["ref_v"][free]) < 1e-10
    # ex7 never sets a constrained velocity: what they hold comes from the
    # round-off accelerations of the force at u = 0, taken before any
    # boundary condition existed
    assert np.abs(z["ref_v"][~free]).max() < 1e-10 * np.abs(z["ref_v"][free]).max()


@pytest.mark.parametrize("name", RUNS)
def test_strain_matches_reference(name):
    """Eavg and the principal summary at the final state of each golden run
    (CalculateStrain.cpp:8-97 run by the reference build)."""
    z = golden("run_" + name)
    o = O.Model(z["X"], z["conn"], z["pid"], z["material_id"], z["props"])
    E, P = o.element_strain(z["ref_u"])
    assert np.abs(E - z["ref_eavg"]).max() <= 1e-14 * max(1.0, np.abs(z["ref_eavg"]).max())
    ref = z["ref_princ"]
    np.testing.assert_array_equal(np.isnan(P), np.isnan(ref))
    ok = ~np.isnan(ref)
    assert np.abs(P[ok] - ref[ok]).max(initial=0.0) <= 1e-12 * max(1.0, np.abs(ref[ok]).max(initial=0.0))


def test_strain_principal_against_eigenvalues():
    """The trigonometric cubic solution against a symmetric eigen-solver on
    the reference's own Eavg outputs (finite entries)."""
    for name in RUNS:
        z = golden("run_" + name)
        E = z["ref_eavg"].reshape(-1, 3, 3).transpose(0, 2, 1)
        P = z["ref_princ"].reshape(-1, 3)
        w = np.linalg.eigvalsh(0.5 * (E + E.transpose(0, 2, 1)))
        ok = ~np.isnan(P[:, 2])
        scale = max(1e-300, np.abs(w).max())
        assert np.abs(P[ok, 0] - np.maximum(w[ok, 2], 0)).max(initial=0) <= 1e-7 * scale
        assert np.abs(P[ok, 1] - np.minimum(w[ok, 0], 0)).max(initial=0) <= 1e-7 * scale
        assert np.abs(P[ok, 2] - 0.5 * (w[ok, 2] - w[ok, 0])).max(initial=0) <= 1e-7 * scale


def test_strain_closed_forms():
    """Homogeneous deformations of a distorted hex: rigid rotation gives
    E = 0; u = (F - I) X gives E = (F^T F - I) / 2 exactly at every Gauss
    point, and the principal values are its eigenvalues."""
    rng = np.random.default_rng(5)
    X = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                  [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float) * 1e-3
    X = X + rng.uniform(-1e-4, 1e-4, X.shape)
    conn = np.arange(8, dtype=np.int32)
    o = O.Model(X.reshape(-1), conn, np.zeros(1, np.int32), np.array([1], np.int32),
                np.array([1000.0, 1e6, 1e7, 0, 0, 0, 0, 0, 0]))
    a = 0.7
    R = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
    E, P = o.element_strain(((X @ R.T) - X).reshape(-1))
    assert np.abs(E).max() < 1e-14 and np.abs(P[:2]).max() < 1e-14
    F = np.eye(3) + np.array([[0.10, 0.03, -0.02], [0.01, -0.05, 0.04], [0.02, 0.00, 0.07]])
    E, P = o.element_strain(((X @ F.T) - X).reshape(-1))
    Ex = 0.5 * (F.T @ F - np.eye(3))
    assert np.abs(E.reshape(3, 3).T - Ex).max() < 1e-14
    w = np.linalg.eigvalsh(Ex)
    np.testing.assert_allclose(P, [max(w[2], 0), min(w[0], 0), 0.5 * (w[2] - w[0])], rtol=0, atol=1e-12)
