"""HIP engine against the oracle and the reference's golden vectors.

All calls go through the C ABI (include/femtech_amd.h) via femtech_amd.
Tolerances: FP64 engine <= 1e-10 relative nodal displacement after the full
run (and round-off-level bounds on the per-routine outputs); FP32 element
arithmetic with FP64 nodal state <= 1e-5 relative L2 displacement and
<= 1e-4 energy-balance error (relative to the largest energy term).
"""
import numpy as np
import pytest

import femtech_amd as ft
from conftest import golden, rel
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RUNS = ["ex9", "bench10", "c1", "impact", "materials"]


def engine_for(z, precision):
    e = ft.Engine(z["X"], z["conn"], z["pid"], z["material_id"], z["props"], precision)
    e.set_boundary(z["dof_motion"], z["motions"])
    return e


def random_u(z, scale, seed):
    rng = np.random.default_rng(seed)
    X = z["X"].reshape(-1, 3)
    h = 1e-3
    # smooth field plus noise: large enough to engage every material branch
    u = scale * h * (np.sin(X / (3 * h)) + 0.3 * rng.uniform(-1, 1, X.shape))
    return u.reshape(-1)


@pytest.mark.parametrize("name", RUNS)
def test_mass_and_setup(gpu_ready, name):
    z = golden("run_" + name)
    with engine_for(z, ft.FT_FP64) as e:
        m = e.lumped_mass()
        assert rel(m, z["ref_mass"]) < 1e-13


@pytest.mark.parametrize("name", RUNS)
def test_element_forces_fp64(gpu_ready, name):
    z = golden("run_" + name)
    o = O.Model(z["X"], z["conn"], z["pid"], z["material_id"], z["props"])
    for seed, scale in ((1, 1e-3), (2, 5e-2)):
        u = random_u(z, scale, seed)
        with engine_for(z, ft.FT_FP64) as e:
            fe = e.element_forces(u, dt=1e-5)
        fo = o.element_forces(u, dt=1e-5)
        assert rel(fe, fo) < 1e-11, (name, scale)


@pytest.mark.parametrize("name", RUNS)
def test_get_force_and_dt_fp64(gpu_ready, name):
    z = golden("run_" + name)
    o = O.Model(z["X"], z["conn"], z["pid"], z["material_id"], z["props"])
    boundary = (z["dof_motion"] >= 0).astype(np.int32)
    u = random_u(z, 2e-2, 3)
    with engine_for(z, ft.FT_FP64) as e:
        fi = e.get_force(u, dt=1e-5)
        dt = e.stable_time_step(u, boundary)
        dt_flags = e.stable_time_step(u)  # flags from set_boundary
    assert rel(fi, o.get_force(u, 1e-5)) < 1e-11
    ref = o.stable_dt(u, boundary)
    assert dt == pytest.approx(ref, rel=1e-13)
    assert dt_flags == pytest.approx(ref, rel=1e-13)


@pytest.mark.parametrize("name", RUNS)
def test_run_fp64_matches_reference(gpu_ready, name):
    z = golden("run_" + name)
    nsteps = int(z["ref_steps"])
    with engine_for(z, ft.FT_FP64) as e:
        e.set_state(None, z["v0"] if "v0" in z else None)
        e.lumped_mass()
        e.set_history(nsteps + 4)
        e.run_init(float(z["tmax"]), float(z["dt_factor"]), float(z["failure_dt"]))
        u0, v0, a0, fi0 = e.state()
        # an unloaded start has round-off-level reference forces: measure
        # against the force scale of the run
        fscale = max(np.linalg.norm(z["ref_fi0"]), np.linalg.norm(z["ref_fi"]))
        ascale = max(np.linalg.norm(z["ref_a0"]), np.linalg.norm(z["ref_a"]))
        assert np.linalg.norm(fi0 - z["ref_fi0"]) <= 1e-11 * fscale
        assert np.linalg.norm(a0 - z["ref_a0"]) <= 1e-11 * ascale
        done = e.run_steps(int(z["max_steps"]))
        assert done == nsteps
        u, v, a, fi = e.state()
        t, dt_next, k = e.time()
        th, dth, en = e.history()
        E = e.energy()
    assert k == nsteps
    assert rel(u, z["ref_u"]) < 1e-10
    assert rel(v, z["ref_v"]) < 1e-9
    assert rel(fi, z["ref_fi"]) < 1e-9
    assert rel(a, z["ref_a"]) < 1e-8
    assert t == pytest.approx(z["ref_time"][-1], rel=1e-13)
    assert dt_next == pytest.approx(z["ref_dt"][-1], rel=1e-11)
    assert len(th) == nsteps
    assert rel(th, z["ref_time"][1:]) < 1e-13
    assert rel(dth, z["ref_dt"][1:]) < 1e-11
    assert rel(en[:, :3], z["ref_energy"][1:, :3]) < 1e-9
    assert rel(E[:3], z["ref_energy"][-1, :3]) < 1e-9


@pytest.mark.parametrize("name", ["c1", "impact", "bench10", "ex9"])
def test_run_fp32_split_within_tolerance(gpu_ready, name):
    """FP32 element arithmetic, FP64 nodal state and energy."""
    z = golden("run_" + name)
    nsteps = int(z["ref_steps"])
    with engine_for(z, ft.FT_FP32) as e:
        e.set_state(None, z["v0"] if "v0" in z else None)
        e.lumped_mass()
        e.set_history(nsteps + 4)
        e.run_init(float(z["tmax"]), float(z["dt_factor"]), float(z["failure_dt"]))
        e.run_steps(nsteps)
        u, v, a, fi = e.state()
        th, dth, en = e.history()
    assert rel(u, z["ref_u"]) < 1e-5
    ref_e = z["ref_energy"][1:len(en) + 1]
    scale = np.abs(ref_e[:, :3]).max()
    bal = en[:, 2] + en[:, 0] - en[:, 1]
    ref_bal = ref_e[:, 2] + ref_e[:, 0] - ref_e[:, 1]
    assert np.abs(bal - ref_bal).max() / scale < 1e-4


def test_run_fp32_materials(gpu_ready):
    z = golden("run_materials")
    nsteps = int(z["ref_steps"])
    with engine_for(z, ft.FT_FP32) as e:
        e.set_state(None, z["v0"])
        e.lumped_mass()
        e.run_init(float(z["tmax"]), float(z["dt_factor"]), float(z["failure_dt"]))
        e.run_steps(nsteps)
        u, v, a, fi = e.state()
    assert rel(u, z["ref_u"]) < 1e-5


def test_visco_history_matches_oracle(gpu_ready):
    z = golden("run_materials")
    o = O.Model(z["X"], z["conn"], z["pid"], z["material_id"], z["props"])
    with engine_for(z, ft.FT_FP64) as e:
        for k in range(4):
            u = random_u(z, 1e-2 * (k + 1), 10 + k)
            fg = e.element_forces(u, dt=2e-5, update_history=True)
            fo = o.element_forces(u, dt=2e-5, update_history=True)
            assert rel(fg, fo) < 1e-11
        hg = e.visco_history().reshape(-1, 8, 3, 6)
    ho = o.history().reshape(-1, 8, 3, 9)  # H1, H2, S0 column-major 3x3
    vo = ho[..., [0, 4, 8, 7, 6, 3]]
    mask = z["material_id"][z["pid"]] == 5
    assert rel(hg[mask], vo[mask]) < 1e-10


def test_runs_are_bitwise_deterministic(gpu_ready):
    z = golden("run_impact")
    outs = []
    for _ in range(2):
        with engine_for(z, ft.FT_FP32) as e:
            e.set_state(None, z["v0"])
            e.lumped_mass()
            e.run_init(float(z["tmax"]), float(z["dt_factor"]), float(z["failure_dt"]))
            e.run_steps(100)
            outs.append(e.state()[0])
    assert np.array_equal(outs[0], outs[1])


def test_run_in_chunks_equals_one_call(gpu_ready):
    z = golden("run_c1")
    res = []
    for chunks in ((300,), (100, 1, 199)):
        with engine_for(z, ft.FT_FP64) as e:
            e.lumped_mass()
            e.run_init(float(z["tmax"]), float(z["dt_factor"]), float(z["failure_dt"]))
            for c in chunks:
                e.run_steps(c)
            res.append(e.state()[0])
    assert np.array_equal(res[0], res[1])


@pytest.mark.parametrize("name", RUNS)
def test_element_strain_matches_reference(gpu_ready, name):
    """Eavg and principal strains (CalculateStrain.cpp:8-97) at the reference's
    final displacement.  Eavg to 1e-13 of its scale (the device forms
    E = (H + H^T + H^T H) / 2 from H = grad u, the reference F^T F - I; the
    two differ by round-off only).  Principal values to 1e-7 of the strain
    scale (the trigonometric cubic loses up to half the digits at a repeated
    eigenvalue); where the reference yields NaN from a degenerate triple the
    device value is checked against a symmetric eigen-solver instead."""
    z = golden("run_" + name)
    with engine_for(z, ft.FT_FP64) as e:
        E, P = e.element_strain(z["ref_u"])
    ref_E, ref_P = z["ref_eavg"], z["ref_princ"]
    scale = max(1e-300, np.abs(ref_E).max())
    assert np.abs(E - ref_E).max() <= 1e-13 * max(1.0, scale)
    assert not np.isnan(P).any()
    Pm = P.reshape(-1, 3)
    ok = ~np.isnan(ref_P.reshape(-1, 3)).any(axis=1)
    assert np.abs(Pm[ok] - ref_P.reshape(-1, 3)[ok]).max(initial=0) <= 1e-7 * scale
    Es = E.reshape(-1, 3, 3).transpose(0, 2, 1)
    w = np.linalg.eigvalsh(0.5 * (Es + Es.transpose(0, 2, 1)))
    want = np.stack([np.maximum(w[:, 2], 0), np.minimum(w[:, 0], 0), 0.5 * (w[:, 2] - w[:, 0])], 1)
    assert np.abs(Pm - want).max() <= 1e-7 * scale


def test_element_strain_device_state(gpu_ready):
    """u = NULL reads the engine's own displacement (after a short run)."""
    z = golden("run_impact")
    o = O.Model(z["X"], z["conn"], z["pid"], z["material_id"], z["props"])
    with engine_for(z, ft.FT_FP64) as e:
        e.set_state(None, z["v0"])
        e.lumped_mass()
        e.run_init(1e9, float(z["dt_factor"]), float(z["failure_dt"]))
        assert e.run_steps(40) == 40
        u, _, _, _ = e.state()
        E1, P1 = e.element_strain()
        E2, P2 = e.element_strain(u)
    np.testing.assert_array_equal(E1, E2)
    np.testing.assert_array_equal(P1, P2)
    Eo, _ = o.element_strain(u)
    assert np.abs(E1 - Eo).max() <= 1e-13 * max(1.0, np.abs(Eo).max())


def rotation_run(precision, steps):
    from cases import case_rotation
    c = case_rotation(steps=steps)
    e = ft.Engine(c["X"], c["conn"], c["pid"], c["material_id"], c["props"], precision)
    e.set_boundary(c["dof_motion"], c["motions"])
    e.lumped_mass()
    e.set_history(steps + 4)
    e.run_init(c["tmax"], c["dt_factor"], c["failure_dt"])
    assert e.run_steps(steps) == steps
    out = e.state(), e.time(), e.history(), e.energy()
    e.close()
    return c, out


def test_rotational_halfsine_loading_matches_oracle(gpu_ready):
    """Outer-surface nodes driven by the rigid half-sine rotation motion,
    through and beyond the pulse (600 steps ~ 16 ms); FP64 engine against
    the oracle's loop with the same motion table."""
    steps = 600
    c, ((u, v, a, fi), (t, dt, k), (th, dth, en), E) = rotation_run(ft.FT_FP64, steps)
    o = O.Model(c["X"], c["conn"], c["pid"], c["material_id"], c["props"])
    o.set_bc(c["dof_motion"], c["motions"])
    r = o.run(steps, c["tmax"], c["dt_factor"], c["failure_dt"])
    assert r["steps"] == steps and k == steps
    assert t > 0.0105  # past the end of the pulse
    assert rel(u, r["u"]) < 1e-10
    assert rel(v, r["v"]) < 1e-9
    assert rel(th, r["time"][1:]) < 1e-13
    assert rel(en[:, :3], r["energy"][1:, :3]) < 1e-9
    # the surface follows the closed-form rotation exactly
    X = c["X"].reshape(-1, 3)
    surf = c["dof_motion"].reshape(-1, 3)[:, 0] == 0
    for i in np.nonzero(surf)[0][:20]:
        ue, _, _ = O.motion_eval(c["motions"][0], X[i], t)
        np.testing.assert_allclose(u.reshape(-1, 3)[i], ue, rtol=1e-12, atol=1e-15)


def test_rotational_halfsine_loading_fp32_split(gpu_ready):
    steps = 600
    c, ((u, v, a, fi), _, (th, dth, en), E) = rotation_run(ft.FT_FP32, steps)
    o = O.Model(c["X"], c["conn"], c["pid"], c["material_id"], c["props"])
    o.set_bc(c["dof_motion"], c["motions"])
    r = o.run(steps, c["tmax"], c["dt_factor"], c["failure_dt"])
    assert rel(u, r["u"]) < 1e-5
    ref = r["energy"][1:]
    scale = np.abs(ref[:, :3]).max()
    assert np.abs((en[:, 2] + en[:, 0] - en[:, 1]) - (ref[:, 2] + ref[:, 0] - ref[:, 1])).max() / scale < 1e-4


def compare_injury(out, fl, ro, rfl):
    """Running extrema and percentiles (values, element ids, times) and the
    per-element flags of the injury criteria, FP64 engine vs oracle."""
    for k in (0, 3, 6, 9, 12, 14):
        assert out[k] == pytest.approx(ro[k], rel=1e-9, abs=1e-300), k
    for k in (1, 4, 7, 10):
        assert out[k] == ro[k], k
    for k in (2, 5, 8, 11, 13, 15):
        assert out[k] == pytest.approx(ro[k], rel=1e-13), k
    np.testing.assert_array_equal(fl, rfl)
    np.testing.assert_allclose(out[16:21], ro[16:21], rtol=1e-12)


def test_injury_criteria_in_resident_loop_match_oracle(gpu_ready):
    """CalculateInjuryCriterions (ex5.cpp:1311-1430) after every step of a
    large-strain rotational run, on the device, against the oracle's update
    after every step of the same run (FP64 engine)."""
    from cases import case_injury
    c = case_injury()
    steps = c["max_steps"]
    with ft.Engine(c["X"], c["conn"], c["pid"], c["material_id"], c["props"], ft.FT_FP64) as e:
        e.set_boundary(c["dof_motion"], c["motions"])
        e.lumped_mass()
        e.run_init(c["tmax"], c["dt_factor"], c["failure_dt"])
        e.injury_enable()
        assert e.run_steps(steps) == steps
        out, fl = e.injury_results()
        u = e.state()[0]
    o = O.Model(c["X"], c["conn"], c["pid"], c["material_id"], c["props"])
    o.set_bc(c["dof_motion"], c["motions"])
    o.injury_init()
    r = o.run(steps, c["tmax"], c["dt_factor"], c["failure_dt"])
    ro, rfl = o.injury_results()
    assert rel(u, r["u"]) < 1e-10
    # every threshold is exercised by this case
    assert ro[16] > 0.5 and ro[17] > 0.3 and ro[18] > 0 and ro[19] > 0
    compare_injury(out, fl, ro, rfl)


def test_injury_update_on_host_displacements_matches_oracle(gpu_ready):
    """ft_injury_update with caller-supplied displacement, time and step."""
    from cases import case_injury
    c = case_injury()
    rng = np.random.default_rng(12)
    Xr = c["X"].reshape(-1, 3)
    o = O.Model(c["X"], c["conn"], c["pid"], c["material_id"], c["props"])
    o.injury_init(np.array([1], np.int32))
    with ft.Engine(c["X"], c["conn"], c["pid"], c["material_id"], c["props"], ft.FT_FP64) as e:
        e.injury_enable(np.array([1], np.int32))
        for k in range(4):
            F = np.eye(3) + (0.1 * (k + 1)) * rng.uniform(-1, 1, (3, 3))
            u = (Xr @ F.T - Xr).reshape(-1) + 1e-4 * rng.uniform(-1, 1, Xr.size)
            e.injury_update(u, 0.01 * (k + 1), 1e-3)
            o.injury_update(u, 0.01 * (k + 1), 1e-3)
        out, fl = e.injury_results()
    ro, rfl = o.injury_results()
    compare_injury(out, fl, ro, rfl)


def ellipsoid_run(precision, steps=300):
    from femtech_amd import workloads as W
    w = W.ellipsoid_head(n=12, peak=5e5, duration=2e-3)
    with ft.Engine(w["X"], w["conn"], w["pid"], w["material_id"], w["props"], precision) as e:
        e.set_boundary(w["dof_motion"], w["motions"])
        e.set_fibres(w["fibres"])
        e.lumped_mass()
        e.set_history(steps + 4)
        e.run_init(1e9, 0.9, 1e-11)
        assert e.run_steps(steps) == steps
        out = e.state(), e.history(), e.visco_history()
    o = O.Model(w["X"], w["conn"], w["pid"], w["material_id"], w["props"])
    o.set_fibres(w["fibres"])
    o.set_bc(w["dof_motion"], w["motions"])
    r = o.run(steps, 1e9, 0.9, 1e-11)
    return w, out, r, o


def test_ellipsoid_three_regions_fp64_matches_oracle(gpu_ready):
    """The ellipsoid workload in miniature (1,728 elements): linear-elastic
    shell, Ogden + Prony middle (strains up to ~23 %), anisotropic HGO core
    with seeded fibres, rotational surface loading; FP64 engine against the
    oracle."""
    w, ((u, v, a, fi), (th, dth, en), hv), r, o = ellipsoid_run(ft.FT_FP64)
    assert rel(u, r["u"]) < 1e-10
    assert rel(v, r["v"]) < 1e-9
    assert rel(th, r["time"][1:]) < 1e-13
    assert rel(en[:, :3], r["energy"][1:, :3]) < 1e-9
    ho = o.history().reshape(-1, 8, 3, 9)[..., [0, 4, 8, 7, 6, 3]]
    mask = w["material_id"][w["pid"]] == 7
    assert rel(hv.reshape(-1, 8, 3, 6)[mask], ho[mask]) < 1e-9


def test_ellipsoid_three_regions_fp32_split(gpu_ready):
    w, ((u, v, a, fi), (th, dth, en), hv), r, o = ellipsoid_run(ft.FT_FP32)
    assert rel(u, r["u"]) < 1e-5
    ref = r["energy"][1:]
    scale = np.abs(ref[:, :3]).max()
    assert np.abs((en[:, 2] + en[:, 0] - en[:, 1]) - (ref[:, 2] + ref[:, 0] - ref[:, 1])).max() / scale < 1e-4


def test_ellipsoid_fp32_through_vanishing_strain(gpu_ready):
    """The start of the rotational pulse (strains rising from zero through
    the FP32 denormal range) stays finite in FP32 and follows the FP64 run:
    the Ogden eigen-solver's reciprocals never meet a vanishing off-diagonal
    term."""
    from femtech_amd import workloads as W
    w = W.ellipsoid_head(n=20)
    outs = []
    for prec in (ft.FT_FP32, ft.FT_FP64):
        with ft.Engine(w["X"], w["conn"], w["pid"], w["material_id"], w["props"], prec) as e:
            e.set_boundary(w["dof_motion"], w["motions"])
            e.set_fibres(w["fibres"])
            e.lumped_mass()
            e.run_init(1e9, 0.9, 1e-11)
            assert e.run_steps(200) == 200
            outs.append((e.state()[0], e.status()))
    (u32, st32), (u64, _) = outs
    assert np.all(np.isfinite(u32)) and not st32 & (ft.STATUS_NONFINITE | ft.STATUS_DT_COLLAPSE)
    assert rel(u32, u64) < 1e-5


def test_new_laws_element_forces_fp64(gpu_ready):
    """Materials 6-8 in the device kernels against the oracle."""
    from test_cpu_kernel_math import new_law_block, field
    X, conn, pid, mid, props, fib = new_law_block()
    o = O.Model(X, conn, pid, mid, props)
    o.set_fibres(fib)
    with ft.Engine(X, conn, pid, mid, props, ft.FT_FP64) as e:
        e.set_fibres(fib)
        for scale, seed in ((1e-3, 3), (8e-2, 4)):
            u = field(X.reshape(-1, 3), scale, seed)
            assert rel(e.element_forces(u, dt=1e-5), o.element_forces(u, 1e-5)) < 1e-11


def batch_instances():
    """Independent instances of different sizes, laws, motions and step
    sizes: jittered impact blocks, a rotating ellipsoid and the
    rotational-loading block."""
    from femtech_amd import workloads as W
    from cases import case_rotation
    a = W.impact_block(n=4, slab=1, seed=11)
    b = W.impact_block(n=6, slab=2, seed=12)
    b["props"] = b["props"].copy()
    b["props"][0, 1] *= 1.3  # stiffer: smaller dt
    c = W.ellipsoid_head(n=8, peak=4e5, duration=2e-3)
    r = case_rotation(n=4, steps=0)
    r = {k: r[k] for k in ("X", "conn", "pid", "material_id", "props", "dof_motion", "motions")}
    r["v0"] = None
    return [a, b, c, r]


def test_batched_instances_match_separate_runs(gpu_ready):
    """Instances batched in one engine, each with its own stable step, time,
    motions and energy, against the oracle run of each instance alone (FP64
    engine; end times differ per instance)."""
    from femtech_amd import workloads as W
    insts = batch_instances()
    m, no, eo = W.concatenate(insts)
    tmax = np.array([4e-4, 3e-4, 2e-4, 5e-4])
    steps = 400
    with ft.Engine(m["X"], m["conn"], m["pid"], m["material_id"], m["props"], ft.FT_FP64) as e:
        e.set_instances(no, eo)
        e.set_fibres(m["fibres"])
        e.set_boundary(m["dof_motion"], m["motions"])
        e.set_state(None, m["v0"])
        e.lumped_mass()
        e.set_instance_end_times(tmax)
        e.run_init(1e9, 0.9, 1e-11)
        done = e.run_steps(100)
        done += e.run_steps(steps)
        u = e.state()[0]
        info = [e.instance(i) for i in range(len(insts))]
    assert done > 0
    for i, w in enumerate(insts):
        o = O.Model(w["X"], w["conn"], w["pid"], w["material_id"], w["props"])
        if w.get("fibres") is not None:
            o.set_fibres(w["fibres"])
        o.set_bc(w["dof_motion"], w["motions"])
        o.set_initial(None, w.get("v0"))
        r = o.run(100000, tmax[i], 0.9, 1e-11)
        assert info[i][3] == 1.0 and int(info[i][2]) == r["steps"], (i, info[i], r["steps"])
        assert info[i][0] == pytest.approx(r["time"][-1], rel=1e-13)
        ui = u[3 * no[i]:3 * no[i + 1]]
        assert rel(ui, r["u"]) < 1e-10, i
        assert rel(info[i][4:7], r["energy"][-1, :3]) < 1e-9, i


def test_single_instance_declared_equals_default(gpu_ready):
    """ft_set_instances with one instance changes nothing."""
    z = golden("run_impact")
    outs = []
    for declare in (False, True):
        with engine_for(z, ft.FT_FP32) as e:
            if declare:
                e.set_instances([0, z["X"].size // 3], [0, z["conn"].size // 8])
                e.set_boundary(z["dof_motion"], z["motions"])
            e.set_state(None, z["v0"])
            e.lumped_mass()
            e.run_init(1e9, float(z["dt_factor"]), float(z["failure_dt"]))
            e.run_steps(50)
            outs.append(e.state()[0])
    assert np.array_equal(outs[0], outs[1])


def run_batch_or_single(insts, tmax, precision, steps):
    from femtech_amd import workloads as W
    m, no, eo = W.concatenate(insts)
    with ft.Engine(m["X"], m["conn"], m["pid"], m["material_id"], m["props"], precision) as e:
        if len(insts) > 1:
            e.set_instances(no, eo)
        if m.get("fibres") is not None:
            e.set_fibres(m["fibres"])
        e.set_boundary(m["dof_motion"], m["motions"])
        e.set_state(None, m["v0"])
        e.lumped_mass()
        if len(insts) > 1:
            e.set_instance_end_times(np.asarray(tmax))
        e.run_init(1e9 if len(insts) > 1 else float(tmax[0]), 0.9, 1e-11)
        e.run_steps(steps)
        u = e.state()[0]
        info = np.array([e.instance(i) for i in range(len(insts))]) if len(insts) > 1 else None
        t = e.time()
    return u, no, info, t


def test_batched_fp32_equals_separate_fp32_runs(gpu_ready):
    """FP32 engine: each instance of a batch follows exactly (bitwise) the
    run of the same instance alone: its tiles, step slot and refinement of
    the step's minimiser are the same."""
    insts = batch_instances()
    tmax = [4e-4, 3e-4, 2e-4, 5e-4]
    u, no, info, _ = run_batch_or_single(insts, tmax, ft.FT_FP32, 600)
    for i, w in enumerate(insts):
        ui, _, _, ti = run_batch_or_single([w], [tmax[i]], ft.FT_FP32, 600)
        assert np.array_equal(u[3 * no[i]:3 * no[i + 1]], ui), i
        assert info[i][0] == ti[0] and int(info[i][2]) == ti[2], (i, info[i], ti)


def test_set_instances_guards_single_instance_features(gpu_ready):
    """ft_set_instances refuses several instances while the step history or
    the injury criteria (single-instance features) are on, and forgets end
    times declared for the previous instances: ft_run_init's end time then
    applies to every instance."""
    from femtech_amd import workloads as W
    insts = [W.impact_block(n=3, slab=1, seed=21), W.impact_block(n=4, slab=1, seed=22)]
    m, no, eo = W.concatenate(insts)
    with ft.Engine(m["X"], m["conn"], m["pid"], m["material_id"], m["props"], ft.FT_FP64) as e:
        e.set_history(10)
        with pytest.raises(ft.FemTechError) as ei:
            e.set_instances(no, eo)
        assert ei.value.code == ft.FT_ERR_STATE
        e.set_history(0)
        e.injury_enable()
        with pytest.raises(ft.FemTechError) as ei:
            e.set_instances(no, eo)
        assert ei.value.code == ft.FT_ERR_STATE
        e.injury_disable()
        e.set_instances(no, eo)
        e.set_instance_end_times([1e-4, 2e-4])
        e.set_instances(no, eo)  # new instances: the end times above no longer apply
        e.set_boundary(m["dof_motion"], m["motions"])
        e.set_state(None, m["v0"])
        e.lumped_mass()
        e.run_init(3e-4, 0.9, 1e-11)
        e.run_steps(5000)
        info = [e.instance(i) for i in range(2)]
    for t, dt, step, state in (x[:4] for x in info):
        assert state == 1.0 and 3e-4 <= t < 3e-4 + 2 * dt, info


@pytest.mark.parametrize("cause", ["inverted", "nonfinite"])
@pytest.mark.parametrize("precision", [ft.FT_FP64, ft.FT_FP32])
def test_batch_instance_failure_is_isolated(gpu_ready, cause, precision):
    """An instance whose elements invert, or whose every element force turns
    non-finite (NaN initial velocities), stops (state 2) at a finite time
    while the others run on to their end times, bitwise as when run alone;
    the batch call returns normally.  (No finite element step at all must
    stop an instance, not send it past its end time.)"""
    from femtech_amd import workloads as W
    if cause == "inverted":
        bad = W.ellipsoid_head(n=10, peak=5e5, duration=2e-3)
    else:
        bad = W.impact_block(n=3, slab=1, seed=13)
        bad["v0"] = np.full_like(bad["v0"], np.nan)
    good = W.impact_block(n=4, slab=1, seed=11)
    tmax = [1e-2, 3e-4]
    u, no, info, _ = run_batch_or_single([bad, good], tmax, precision, 2000)
    assert info[0][3] == 2.0 and info[1][3] == 1.0, info
    assert 0.0 <= info[0][0] < tmax[0], info[0]
    ug, _, _, tg = run_batch_or_single([good], [tmax[1]], precision, 2000)
    assert np.array_equal(u[3 * no[1]:3 * no[2]], ug)
    assert int(info[1][2]) == tg[2]


@pytest.mark.parametrize("name", ["impact", "materials"])
def test_fp32_step_sizes_come_from_fp64_geometry(gpu_ready, name):
    """FP32 engine, one step per call: each next step size equals dt_factor x
    the FP64 stable step (the oracle's StableTimeStep) at the engine's own
    displacement, including the steps where the minimum rises by more than
    0.1 % (half of the impact run's steps): the element kernel's last block
    refines the best unrefined candidate in FP64."""
    z = golden("run_" + name)
    o = O.Model(z["X"], z["conn"], z["pid"], z["material_id"], z["props"])
    bnd = (z["dof_motion"] >= 0).astype(np.int32)
    factor = float(z["dt_factor"])
    worst, rises, prev = 0.0, 0, None
    with engine_for(z, ft.FT_FP32) as e:
        e.set_state(None, z["v0"] if "v0" in z else None)
        e.lumped_mass()
        e.run_init(float(z["tmax"]), factor, float(z["failure_dt"]))
        for k in range(120):
            assert e.run_steps(1) == 1
            u = e.state()[0]
            _, dt_next, _ = e.time()
            want = factor * o.stable_dt(u, bnd)
            worst = max(worst, abs(dt_next - want) / want)
            if prev is not None and dt_next > prev * 1.001:
                rises += 1
            prev = dt_next
        assert e.status() == 0
    assert rises > 0 or name == "materials"
    assert worst < 1e-12, worst
