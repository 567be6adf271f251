"""Launch shapes and captured steps: HIP-graph chunks of the run loop, the
element-kernel autotuner and per-segment shapes give the same run as the
default launches (bitwise where only the launch changes, within round-off
where the tile size moves which elements refine the step size in FP64)."""
import numpy as np
import pytest

import femtech_amd as ft
from conftest import rel
from test_gpu_parity import batch_instances

pytestmark = pytest.mark.gpu


def batched_engine(precision):
    from femtech_amd import workloads as W
    insts = batch_instances()
    m, no, eo = W.concatenate(insts)
    e = ft.Engine(m["X"], m["conn"], m["pid"], m["material_id"], m["props"], precision)
    e.set_instances(no, eo)
    e.set_fibres(m["fibres"])
    e.set_boundary(m["dof_motion"], m["motions"])
    e.set_state(None, m["v0"])
    e.lumped_mass()
    e.set_instance_end_times(np.array([4e-4, 3e-4, 2e-4, 5e-4]))
    e.run_init(1e9, 0.9, 1e-11)
    return e, len(insts)


def snapshot(e, n_inst):
    u, v, a, fi = e.state()
    return u, v, np.array([e.instance(i) for i in range(n_inst)])


@pytest.mark.parametrize("graph_steps", [16, 5])
def test_graph_chunks_equal_separate_launches(gpu_ready, graph_steps):
    """Every replay of the captured chunks (even and odd chunk lengths, so
    both starting parities) matches the separately launched run bitwise,
    with host reads and synchronisations between the replays; instances
    finish at different steps inside the chunks."""
    chunks = [16, 16, 7, 40, 3, 120, 200]
    ref, got = [], []
    for g in (0, graph_steps):
        e, n = batched_engine(ft.FT_FP32)
        with e:
            e.set_graph(g)
            out = []
            for c in chunks:
                e.run_steps(c)
                out.append(snapshot(e, n))
        (ref if g == 0 else got).extend(out)
    for k, (r, s) in enumerate(zip(ref, got)):
        for x, y in zip(r, s):
            assert np.array_equal(x, y), k


def test_graph_with_history_and_injury(gpu_ready):
    """The per-step history row and the injury update run inside captured
    chunks exactly as in separate launches (single-instance ellipsoid with
    Prony history)."""
    from femtech_amd import workloads as W
    w = W.ellipsoid_head(n=12, peak=5e5, duration=2e-3)
    outs = []
    for g in (0, 8):
        with ft.Engine(w["X"], w["conn"], w["pid"], w["material_id"], w["props"], ft.FT_FP32) as e:
            e.set_boundary(w["dof_motion"], w["motions"])
            e.set_fibres(w["fibres"])
            e.lumped_mass()
            e.set_history(200)
            e.injury_enable()
            e.run_init(1e9, 0.9, 1e-11)
            e.set_graph(g)
            e.run_steps(64)
            e.run_steps(37)
            u = e.state()[0]
            th, dth, en = e.history()
            res, fl = e.injury_results()
            outs.append((u, th, dth, en, res, fl, e.visco_history()))
    for x, y in zip(*outs):
        assert np.array_equal(x, y)


def test_graph_dropped_on_reconfiguration(gpu_ready):
    """A new run (other factor) after captured chunks uses its own arguments."""
    from femtech_amd import workloads as W
    w = W.impact_block(n=6, slab=2, seed=3)
    outs = []
    for g in (0, 4):
        with ft.Engine(w["X"], w["conn"], w["pid"], w["material_id"], w["props"], ft.FT_FP32) as e:
            e.set_boundary(w["dof_motion"], w["motions"])
            e.set_state(None, w["v0"])
            e.lumped_mass()
            e.set_graph(g)
            e.run_init(1e9, 0.9, 1e-11)
            e.run_steps(20)
            e.set_state(None, w["v0"])
            e.run_init(1e9, 0.5, 1e-11)
            e.run_steps(20)
            outs.append((e.state()[0], e.time()))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert outs[0][1] == outs[1][1]


def test_autotune_keeps_the_run(gpu_ready):
    """Tuning between run calls leaves the state untouched, and the tuned
    shapes continue the run within round-off of the default shapes."""
    e0, n = batched_engine(ft.FT_FP32)
    e1, _ = batched_engine(ft.FT_FP32)
    with e0, e1:
        e0.run_steps(30)
        e1.run_steps(30)
        before = snapshot(e1, n)
        rep = e1.autotune(2)
        assert rep.shape == (len(e1.segments()), 6)
        assert np.all(np.isin(rep[:, 3], [64, 128, 256])) and np.all((rep[:, 2] >= 0) & (rep[:, 2] <= 5))
        assert np.all(rep[:, 4] > 0) and np.all(rep[:, 4] <= rep[:, 5])
        assert np.all((rep[:, 2] > 0) | (rep[:, 3] == 256))  # waves 0 only at the default tile
        seg = e1.segments()
        assert np.array_equal(seg[:, 3], rep[:, 2].astype(int)) and np.array_equal(seg[:, 4], rep[:, 3].astype(int))
        for x, y in zip(before, snapshot(e1, n)):
            assert np.array_equal(x, y)
        e0.run_steps(200)
        e1.run_steps(200)
        u0, v0, i0 = snapshot(e0, n)
        u1, v1, i1 = snapshot(e1, n)
    assert rel(u1, u0) < 1e-6
    assert np.array_equal(i0[:, 2], i1[:, 2])  # same step counts


def test_occupancy_variants_are_bitwise_equal(gpu_ready):
    """The occupancy variants differ only in register allocation."""
    outs = []
    for waves in (0, 1, 3, 5):
        e, n = batched_engine(ft.FT_FP32)
        with e:
            for s in range(len(e.segments())):
                e.set_segment_shape(s, waves, 256)
            e.run_steps(60)
            outs.append(snapshot(e, n))
    for o in outs[1:]:
        for x, y in zip(outs[0], o):
            assert np.array_equal(x, y)


def test_segment_shape_errors(gpu_ready):
    e, n = batched_engine(ft.FT_FP64)
    with e:
        with pytest.raises(ft.FemTechError):
            e.set_segment_shape(0, 0, 96)
        with pytest.raises(ft.FemTechError):
            e.set_segment_shape(0, 2, 256)  # occupancy variants are FP32-only
        with pytest.raises(ft.FemTechError):
            e.set_segment_shape(99, 0, 256)
        with pytest.raises(ft.FemTechError):
            e.set_graph(-1)
        e.set_segment_shape(0, 0, 64)
        e.run_steps(10)


@pytest.mark.parametrize("precision", [ft.FT_FP64, ft.FT_FP32])
def test_inverted_element_collapses_the_step(gpu_ready, precision):
    """A coarse ellipsoid under an extreme pulse inverts elements within 100
    steps; the reference stops with code 19 (a negative element time step is
    below FailureTimeStep), and so do the oracle and the engine, in both
    precisions, instead of running on with non-finite values."""
    from femtech_amd import workloads as W
    from oracle import oracle as O
    w = W.ellipsoid_head(n=10, peak=5e5, duration=2e-3)
    o = O.Model(w["X"], w["conn"], w["pid"], w["material_id"], w["props"])
    o.set_fibres(w["fibres"])
    o.set_bc(w["dof_motion"], w["motions"])
    with pytest.raises(RuntimeError, match="code 19"):
        o.run(101, 1e9, 0.9, 1e-11)
    with ft.Engine(w["X"], w["conn"], w["pid"], w["material_id"], w["props"], precision) as e:
        e.set_boundary(w["dof_motion"], w["motions"])
        e.set_fibres(w["fibres"])
        e.lumped_mass()
        e.run_init(1e9, 0.9, 1e-11)
        with pytest.raises(ft.FemTechError) as err:
            e.run_steps(101)
        assert err.value.code == ft.FT_ERR_DT
        assert e.status() & ft.STATUS_DT_COLLAPSE
        assert e.status() & ft.STATUS_INVERTED


@pytest.mark.parametrize("graph", [0, 8])
def test_snapshot_overlaps_later_steps_and_equals_a_stopped_run(gpu_ready, graph):
    """ft_snapshot_begin after 48 steps, 40 more steps, then ft_snapshot_end:
    the snapshot is the displacement, time and step of step 48, bit for bit
    as a run stopped there (with and without captured chunks); a second
    begin while one is pending, or an end without a begin, is a call-order
    error."""
    from femtech_amd import workloads as W
    w = W.impact_block(n=12, slab=2, seed=5)

    def engine():
        e = ft.Engine(w["X"], w["conn"], w["pid"], w["material_id"], w["props"], ft.FT_FP32)
        e.set_boundary(w["dof_motion"], w["motions"])
        e.set_state(None, w["v0"])
        e.lumped_mass()
        e.run_init(1e9, 0.9, 1e-11)
        e.set_graph(graph)
        return e

    with engine() as e:
        with pytest.raises(ft.FemTechError) as ei:
            e.snapshot_end()
        assert ei.value.code == ft.FT_ERR_STATE
        e.run_steps(48)
        e.snapshot_begin()
        with pytest.raises(ft.FemTechError) as ei:
            e.snapshot_begin()
        assert ei.value.code == ft.FT_ERR_STATE
        e.run_steps(40)
        us, ts, ks = e.snapshot_end()
        np.testing.assert_array_equal(e.snapshot_view(), us)
        later = e.state()[0]
    with engine() as f:
        f.run_steps(48)
        u48 = f.state()[0]
        t48, _, k48 = f.time()
    assert ks == k48 == 48 and ts == t48
    assert np.array_equal(us, u48)
    assert not np.array_equal(later, u48)


@pytest.mark.parametrize("precision", [ft.FT_FP32, ft.FT_FP64])
def test_boundary_replaced_after_creation(gpu_ready, precision):
    """Only the node tiles that hold a constrained node stage the motion
    states.  A boundary replaced after creation (first every DOF free, then
    the ellipsoid's rotating surface; or set once) marks the tiles again: the
    runs equal each other bit for bit."""
    from femtech_amd import workloads as W
    w = W.ellipsoid_head(n=12, peak=5e5, duration=2e-3)
    outs = []
    for first in (None, "free"):
        with ft.Engine(w["X"], w["conn"], w["pid"], w["material_id"], w["props"], precision) as e:
            if first == "free":
                e.set_boundary(-np.ones_like(w["dof_motion"]), w["motions"])
            e.set_boundary(w["dof_motion"], w["motions"])
            e.set_fibres(w["fibres"])
            e.lumped_mass()
            e.run_init(1e9, 0.9, 1e-11)
            e.run_steps(60)
            assert e.status() == 0
            outs.append((e.state()[0], e.time()))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert outs[0][1] == outs[1][1]
    # the surface motion moved the nodes
    assert np.abs(outs[0][0]).max() > 0
