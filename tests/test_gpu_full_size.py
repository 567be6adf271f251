"""The FP32/FP64 split at full size: the 1M-hex jittered neo-Hookean impact
block run for 5,000 steps with FP32 element arithmetic and FP64 nodal state,
against the FP64 engine on the same mesh, loading and step count (the FP64
engine matches the FP64 oracle to 1e-10 relative displacement on the smaller
cases of test_gpu_parity.py; the oracle itself would need about 15 minutes of
host time here).  Bars: 1e-5 relative L2 nodal displacement, 1e-4
energy-balance error relative to the largest energy term; the peak maximum
principal strain is held to the displacement bar as well."""
import numpy as np
import pytest

import femtech_amd as ft
from femtech_amd import workloads as W

pytestmark = pytest.mark.gpu

STEPS = 5000


def run(w, precision):
    with ft.Engine(w["X"], w["conn"], w["pid"], w["material_id"], w["props"], precision) as e:
        e.set_boundary(w["dof_motion"], w["motions"])
        e.set_state(None, w["v0"])
        e.lumped_mass()
        e.run_init(1e9, 0.9, 1e-11)
        assert e.run_steps(STEPS) == STEPS
        _, P = e.element_strain()
        return e.state()[0], e.energy(), P.reshape(-1, 3)[:, 0], e.time()


def test_million_hex_fp32_split_within_bars(gpu_ready):
    w = W.impact_block(n=100)
    u32, e32, p32, t32 = run(w, ft.FT_FP32)
    u64, e64, p64, t64 = run(w, ft.FT_FP64)
    assert np.linalg.norm(u32 - u64) / np.linalg.norm(u64) < 1e-5
    assert abs(e32[3] - e64[3]) / np.max(np.abs(e64[:3])) < 1e-4
    assert abs(p32.max() - p64.max()) / abs(p64.max()) < 1e-5
    assert abs(t32[0] - t64[0]) / t64[0] < 1e-5  # the same time grid
    assert np.max(np.abs(e64[:3])) > 0.1  # the impact has loaded the block
