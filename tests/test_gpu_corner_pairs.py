"""Corner pairs (FT_LAYOUT_CORNER_PAIRS): in the run loop, element i and
element i - 1 of one material segment that share a face along xi (i's local
nodes 0, 3, 4, 7 are i - 1's nodes 1, 2, 5, 6) add that face's corner
forces in registers (each wavefront's lane 0 recomputes the element before
the wavefront's own 63); element i writes the four sums into its own corner
slots and element i - 1 writes only its other four.  An element without a
right partner writes each xi+ corner to the matching xi- slot of its node
when no element owns that slot (a row end of a block), else to an extras
entry that a small kernel adds to that slot after the element pass (the
ellipsoid, whose material segments interleave along xi), so the node sum
reads only the four xi- slots.

The pair sum is one more floating-point addition before the node sum (the
reference's scatter-add, GetForce_3D.cpp:39-44, adds the eight element
contributions one by one), so runs differ from the plain corner slots by
round-off only: FP64 <= 1e-12 relative displacement and both within the
oracle bars (<= 1e-10), FP32 <= 1e-6 of each other.  Forces asked for
outside the run loop (ft_get_force, ft_element_forces) are written without
pairs and equal the corner layout's bit for bit.
"""
import numpy as np
import pytest

import femtech_amd as ft
from conftest import golden, rel
from test_general_mesh import engine_for, run_engine

pytestmark = pytest.mark.gpu


def golden_case(name):
    z = golden("run_" + name)
    c = {k: z[k] for k in ("X", "conn", "pid", "material_id", "props", "dof_motion", "motions")}
    c["v0"] = z["v0"] if "v0" in z else None
    c.update(tmax=float(z["tmax"]), dt_factor=float(z["dt_factor"]), failure_dt=float(z["failure_dt"]),
             steps=int(z["ref_steps"]), fibres=None)
    return c, z


def block_case(n=24, steps=300, reverse=False):
    from femtech_amd import workloads as W
    w = W.impact_block(n=n, slab=4, seed=9)
    conn = w["conn"].reshape(-1, 8)
    pid = np.asarray(w["pid"])
    if reverse:  # the same mesh with its elements in reverse order: no element follows its -xi neighbour
        conn, pid = conn[::-1], pid[::-1]
    c = {k: w[k] for k in ("X", "material_id", "props", "dof_motion", "motions", "v0")}
    c.update(conn=np.ascontiguousarray(conn).reshape(-1), pid=np.ascontiguousarray(pid), fibres=None,
             tmax=1e9, dt_factor=0.9, failure_dt=1e-11, steps=steps)
    return c


def test_structured_meshes_take_corner_pairs(gpu_ready):
    """Creation picks corner pairs for a structured block; the plain corner
    slots stay selectable; with no element after its -xi neighbour (the
    same block in reverse element order) the layout reports corner slots."""
    c = block_case()
    with engine_for(c, ft.FT_FP32) as e:
        assert e.layout()[0] == ft.LAYOUT_CORNER_PAIRS
        pairs = e.traffic_model()
        e.set_layout(ft.LAYOUT_CORNER)
        assert e.layout()[0] == ft.LAYOUT_CORNER
        plain = e.traffic_model()
        # about half of the corner slots are written, and the node sum reads four
        assert pairs[0] < 0.85 * plain[0] and pairs[1] < 0.8 * plain[1]
        e.set_layout(ft.LAYOUT_CORNER_PAIRS)
        assert e.layout()[0] == ft.LAYOUT_CORNER_PAIRS
    with engine_for(block_case(reverse=True), ft.FT_FP32) as e:
        assert e.layout()[0] == ft.LAYOUT_CORNER
    # partner-less elements across material segments: extras, and still a
    # four-slot node sum
    from femtech_amd import workloads as W
    w = W.ellipsoid_head(n=12, peak=5e5, duration=2e-3)
    w["fibres"] = w.get("fibres")
    with engine_for(w, ft.FT_FP32) as e:
        assert e.layout()[0] == ft.LAYOUT_CORNER_PAIRS
        pairs = e.traffic_model()
        e.set_layout(ft.LAYOUT_CORNER)
        plain = e.traffic_model()
        assert pairs[0] < plain[0] and pairs[1] < 0.8 * plain[1]
    # a mesh off the corner layout cannot take pairs
    from test_general_mesh import general_case
    with engine_for(general_case("impact"), ft.FT_FP64) as e:
        with pytest.raises(ft.FemTechError) as ei:
            e.set_layout(ft.LAYOUT_CORNER_PAIRS)
        assert ei.value.code == ft.FT_ERR_ARG


@pytest.mark.parametrize("name", ["impact", "materials", "c1"])
def test_pairs_fp64_match_the_reference_run(gpu_ready, name):
    """The reference-built golden runs: FP64 engine with corner pairs against
    the reference's trajectory (<= 1e-10) and against the plain corner slots
    (<= 1e-12: round-off of the pair sums)."""
    c, z = golden_case(name)
    a = run_engine(c, ft.FT_FP64, history=False, layout=ft.LAYOUT_CORNER_PAIRS)
    b = run_engine(c, ft.FT_FP64, history=False, layout=ft.LAYOUT_CORNER)
    assert a["done"] == b["done"] == c["steps"]
    assert rel(a["state"][0], z["ref_u"]) < 1e-10
    assert rel(a["state"][0], b["state"][0]) < 1e-12
    assert rel(a["state"][1], b["state"][1]) < 1e-11
    assert rel(np.array(a["energy"][:3]), np.array(b["energy"][:3])) < 1e-11


@pytest.mark.parametrize("tile", [256, 128, 64])
def test_pairs_fp32_block_every_tile_size(gpu_ready, tile):
    """The 24^3 block in FP32 with 64-, 128- and 256-element tiles (the
    wavefront boundaries, where pairs stop, move with them): within 1e-6 of
    the plain corner slots, same step count and time."""
    c = block_case()
    outs = []
    for layout in (ft.LAYOUT_CORNER_PAIRS, ft.LAYOUT_CORNER):
        with engine_for(c, ft.FT_FP32, layout) as e:
            e.set_state(None, c["v0"])
            e.lumped_mass()
            e.run_init(c["tmax"], c["dt_factor"], c["failure_dt"])
            e.set_segment_shape(0, 0, tile)
            assert e.layout()[0] == layout
            e.run_steps(c["steps"])
            u, v, _, _ = e.state()
            outs.append((u, v, e.time(), e.energy(), e.status()))
    (u1, v1, t1, e1, s1), (u2, v2, t2, e2, s2) = outs
    assert s1 == s2 == 0
    assert t1[2] == t2[2] == c["steps"]
    assert t1[0] == pytest.approx(t2[0], rel=1e-12)
    assert rel(u1, u2) < 1e-6
    assert rel(v1, v2) < 1e-5
    assert abs(e1[3] - e2[3]) <= 1e-6 * max(abs(x) for x in e2[:3])


def test_forces_outside_the_loop_are_unpaired(gpu_ready):
    """ft_get_force and ft_element_forces after a paired run: the same bits
    as the plain corner layout's (those passes write every corner)."""
    c = block_case(n=12, steps=40)
    rng = np.random.default_rng(3)
    u = 1e-6 * rng.standard_normal(c["X"].size)
    res = []
    for layout in (ft.LAYOUT_CORNER_PAIRS, ft.LAYOUT_CORNER):
        with engine_for(c, ft.FT_FP32, layout) as e:
            e.set_state(None, c["v0"])
            e.lumped_mass()
            e.run_init(c["tmax"], c["dt_factor"], c["failure_dt"])
            e.run_steps(c["steps"])
            res.append((e.get_force(u, 1e-7), e.element_forces(u, 1e-7)))
    assert np.array_equal(res[0][0], res[1][0])
    assert np.array_equal(res[0][1], res[1][1])


@pytest.mark.parametrize("precision", [ft.FT_FP32, ft.FT_FP64])
@pytest.mark.parametrize("kind", ["block", "ellipsoid"])
def test_pairs_graph_and_persistent(gpu_ready, precision, kind):
    """Captured chunks equal the per-step launches bit for bit with corner
    pairs, without (block) or with extras (ellipsoid); the persistent loop
    agrees to round-off."""
    if kind == "block":
        c = block_case(n=16, steps=120)
    else:
        from femtech_amd import workloads as W
        w = W.ellipsoid_head(n=12, peak=5e5, duration=2e-3)
        c = dict(w, fibres=w.get("fibres"), tmax=1e9, dt_factor=0.9, failure_dt=1e-11, steps=120)
    chunks = [1, 16, 40, 63]
    plain = run_engine(c, precision, history=False, chunks=chunks)
    graph = run_engine(c, precision, history=False, graph=16, chunks=chunks)
    pers = run_engine(c, precision, history=False, persistent=True, chunks=chunks)
    for x, y in zip(plain["state"], graph["state"]):
        assert np.array_equal(x, y)
    assert rel(pers["state"][0], plain["state"][0]) < (1e-6 if precision == ft.FT_FP32 else 1e-12)


@pytest.mark.parametrize("precision", [ft.FT_FP64, ft.FT_FP32])
def test_pairs_incomplete_ellipsoid(gpu_ready, precision):
    """The three-material ellipsoid (pairs only inside each material's runs
    along xi, the rest through extras): corner pairs against plain corner
    slots, FP64 to 1e-12 and FP32 to 1e-6, the Prony history included."""
    from femtech_amd import workloads as W
    w = W.ellipsoid_head(n=12, peak=5e5, duration=2e-3)
    c = dict(w, fibres=w.get("fibres"), tmax=1e9, dt_factor=0.9, failure_dt=1e-11, steps=200)
    a = run_engine(c, precision, history=False, layout=ft.LAYOUT_CORNER_PAIRS)
    b = run_engine(c, precision, history=False, layout=ft.LAYOUT_CORNER)
    assert a["done"] == b["done"] == c["steps"]
    bar = 1e-12 if precision == ft.FT_FP64 else 1e-6
    assert rel(a["state"][0], b["state"][0]) < bar


def test_pairs_repeat_bitwise_with_prony_history(gpu_ready):
    """Two runs of the ellipsoid (Ogden + Prony core, anisotropic HGO) with
    corner pairs give the same bits: no lane reads an element's Prony
    history while another rewrites it (the Prony segments pair only inside
    a wavefront, without shadow lanes)."""
    from femtech_amd import workloads as W
    w = W.ellipsoid_head(n=12, peak=5e5, duration=2e-3)
    c = dict(w, fibres=w.get("fibres"), tmax=1e9, dt_factor=0.9, failure_dt=1e-11, steps=150)
    a = run_engine(c, ft.FT_FP32, history=False, chunks=[1, 2, 47, 100])
    b = run_engine(c, ft.FT_FP32, history=False, chunks=[150])
    for x, y in zip(a["state"], b["state"]):
        assert np.array_equal(x, y)
    assert a["time"] == b["time"]
