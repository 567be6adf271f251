"""Persistent run loop (ft_set_persistent): one cooperative launch runs all
the steps of a run_steps call, its phases separated by grid-wide barriers.
The element and node bodies are the per-step kernels' own source, the tiles
and the order of the reductions are the same; compiled into one kernel, the
compiler may still fuse a different set of multiply-adds, so the runs must
agree with the launch-per-step loop to floating-point round-off (they agree
bitwise on the FP32 engines measured so far): relative 1e-9 for the FP64
nodal state of the FP32 engine, 1e-12 for the FP64 engine, step counts and
stopping states exactly."""
import time

import numpy as np
import pytest

import femtech_amd as ft
from test_gpu_launch import batched_engine, snapshot

pytestmark = pytest.mark.gpu

CHUNKS = [16, 16, 7, 40, 3, 120, 200]


def close(a, b, tol):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    scale = max(float(np.max(np.abs(b))), 1e-300) if b.size else 1.0
    return a.shape == b.shape and float(np.max(np.abs(a - b), initial=0.0)) <= tol * scale


def tol_of(precision):
    return 1e-9 if precision == ft.FT_FP32 else 1e-12


@pytest.mark.parametrize("precision", [ft.FT_FP32, ft.FT_FP64])
def test_persistent_batch_equals_launches(gpu_ready, precision):
    """Four instances (neo-Hookean blocks, three-region ellipsoids with Ogden +
    Prony, shells and fibre-reinforced cores), each with its own step and end
    time; the chunks stop instances inside a launch and the last ones run
    past every end (the launch ends early)."""
    outs = []
    for persistent in (False, True):
        e, n = batched_engine(precision)
        with e:
            e.set_persistent(persistent)
            out = []
            for c in CHUNKS:
                e.run_steps(c)
                out.append(snapshot(e, n) + (e.energy(),))
            outs.append(out)
    tol = tol_of(precision)
    for k, (r, s) in enumerate(zip(*outs)):
        u0, v0, i0, e0 = r
        u1, v1, i1, e1 = s
        assert close(u1, u0, tol) and close(v1, v0, tol), k
        assert np.array_equal(i1[:, 2:4], i0[:, 2:4]), k  # steps and stopping states
        assert close(i1[:, :2], i0[:, :2], 1e-13), k      # time and step size
        assert close(i1[:, 4:], i0[:, 4:], tol), k        # energy terms
        assert close(e1, e0, tol), k


def test_persistent_neo_hookean_block(gpu_ready):
    """An all-neo-Hookean block takes the neo-Hookean-only persistent kernel;
    the time history rows and the energy balance come out the same."""
    from femtech_amd import workloads as W
    w = W.impact_block(n=16, slab=3, seed=11)
    outs = []
    for persistent in (False, True):
        with ft.Engine(w["X"], w["conn"], w["pid"], w["material_id"], w["props"], ft.FT_FP32) as e:
            e.set_boundary(w["dof_motion"], w["motions"])
            e.set_state(None, w["v0"])
            e.lumped_mass()
            e.set_history(400)
            e.run_init(1e9, 0.9, 1e-11)
            e.set_persistent(persistent)
            e.run_steps(150)
            e.run_steps(51)
            outs.append(e.state()[:2] + e.history() + (e.energy(), e.time()))
    for x, y in zip(*outs):
        assert close(x, y, 1e-9)
    assert outs[0][-1][2] == outs[1][-1][2] == 201


def test_persistent_rotational_pulse_with_prony(gpu_ready):
    """Timed boundary motions (evaluated inside the node tiles of the
    persistent launch) and Prony history on the three-region ellipsoid."""
    from femtech_amd import workloads as W
    w = W.ellipsoid_head(n=12, peak=5e5, duration=2e-3)
    outs = []
    for persistent in (False, True):
        with ft.Engine(w["X"], w["conn"], w["pid"], w["material_id"], w["props"], ft.FT_FP32) as e:
            e.set_boundary(w["dof_motion"], w["motions"])
            e.set_fibres(w["fibres"])
            e.lumped_mass()
            e.set_history(300)
            e.run_init(1e9, 0.9, 1e-11)
            e.set_persistent(persistent)
            e.run_steps(90)
            e.run_steps(33)
            outs.append(e.state()[:2] + e.history() + (e.visco_history(), e.time()))
    for x, y in zip(*outs):
        assert close(x, y, 1e-9)
    assert outs[0][-1][2] == outs[1][-1][2] == 123


def test_persistent_ends_with_the_run(gpu_ready):
    """A call asking for far more steps than the run has left ends when the
    run does (end time reached): the same step count and state as the
    per-step loop, and no time spent on the steps it did not need."""
    from femtech_amd import workloads as W
    w = W.impact_block(n=6, slab=2, seed=3)
    outs = []
    for persistent in (False, True):
        with ft.Engine(w["X"], w["conn"], w["pid"], w["material_id"], w["props"], ft.FT_FP32) as e:
            e.set_boundary(w["dof_motion"], w["motions"])
            e.set_state(None, w["v0"])
            e.lumped_mass()
            e.run_init(2e-5, 0.9, 1e-11)
            e.set_persistent(persistent)
            steps = 500 if not persistent else 1_000_000
            t0 = time.perf_counter()
            done = e.run_steps(steps)
            dt = time.perf_counter() - t0
            outs.append((done, e.time(), e.state()[0]))
            if persistent:
                assert dt < 5.0
                assert e.run_steps(10) == 0  # stopped: nothing left to run
    assert outs[0][0] == outs[1][0] and outs[0][0] < 500
    assert outs[0][1][2] == outs[1][1][2] and close(outs[1][1][:2], outs[0][1][:2], 1e-13)
    assert close(outs[1][2], outs[0][2], 1e-9)


def test_persistent_falls_back_with_injury_criteria(gpu_ready):
    """Runs with the per-step injury update keep the per-step launches: the
    persistent switch changes nothing there, results included."""
    from femtech_amd import workloads as W
    w = W.ellipsoid_head(n=12, peak=5e5, duration=2e-3)
    outs = []
    for persistent in (False, True):
        with ft.Engine(w["X"], w["conn"], w["pid"], w["material_id"], w["props"], ft.FT_FP32) as e:
            e.set_boundary(w["dof_motion"], w["motions"])
            e.set_fibres(w["fibres"])
            e.lumped_mass()
            e.injury_enable()
            e.run_init(1e9, 0.9, 1e-11)
            e.set_persistent(persistent)
            e.run_steps(60)
            res, fl = e.injury_results()
            outs.append((e.state()[0], res, fl))
    for x, y in zip(*outs):
        assert np.array_equal(np.asarray(x), np.asarray(y))
