"""Training of the constitutive surrogate (materials 7 and 8) from plain
runs of the engine.

Training samples are the Gauss-point states of plain FP32 runs of the
three-region ellipsoid under seeded variants of the rotational pulse: every
`every` steps, the inputs and the analytic isochoric stress of every
material-7/8 Gauss point are appended on the device (ft_surrogate_collect).
The network is then trained on the device by its own BF16 matrix-core
kernels (ft_surrogate_train), and its parameters and normalisation block
are returned for any engine holding the same laws (ft_surrogate_set).
"""
import numpy as np

from . import FT_FP32, Engine
from .workloads import ellipsoid_head


def training_pulses(count=3, seed=7, peak=(4000.0, 8000.0), duration=(0.006, 0.015)):
    """Seeded pulse variants (peak rad/s^2, duration s)."""
    rng = np.random.default_rng(seed)
    return [(float(rng.uniform(*peak)), float(rng.uniform(*duration))) for _ in range(count)]


def collect_and_train(n=40, pulses=None, t_end=0.004, snapshots=40, capacity=None, steps=6000, batch=8192,
                      lr=2e-3, seed=1, semi_axes=(0.090, 0.075, 0.065), log=None):
    """Plain runs of an n^3-element ellipsoid, one per pulse, to t_end with
    `snapshots` collections each; then device training.  Returns (params,
    norm, report) with report = dict(loss, samples, rel_err, steps).  `log`
    (optional) receives a line after each run and after training."""
    pulses = pulses if pulses is not None else training_pulses()
    w = ellipsoid_head(n=n, semi_axes=semi_axes, peak=pulses[0][0], duration=pulses[0][1])
    ne = w["conn"].size // 8
    n_surr = int(np.sum(np.isin(w["material_id"][w["pid"]], [7, 8])))
    cap = capacity or 8 * n_surr * snapshots * len(pulses)
    with Engine(w["X"], w["conn"], w["pid"], w["material_id"], w["props"], FT_FP32) as e:
        e.set_fibres(w["fibres"])
        e.lumped_mass()
        e.surrogate_reserve(cap)
        held = 0
        for peak, duration in pulses:
            mo = np.array(w["motions"], dtype=np.float64).copy()
            mo[0, 7], mo[0, 8] = peak, duration
            e.set_boundary(w["dof_motion"], mo)
            e.set_state(None, None)
            e.reset_history()
            e.run_init(t_end, 0.9, 1e-11)
            t_snap = [t_end * (k + 1) / snapshots for k in range(snapshots)]
            for ts in t_snap:
                t, dt, _ = e.time()
                while t < ts:
                    k = int(min(max(np.ceil((ts - t) / max(dt, 1e-300)), 1), 1000))
                    if e.run_steps(k) == 0:
                        break
                    t, dt, _ = e.time()
                held = e.surrogate_collect()
            if log:
                log("surrogate training data: pulse %.0f rad/s^2 over %.1f ms, %d steps, %d samples held"
                    % (peak, 1e3 * duration, e.time()[2], held))
        rep = e.surrogate_train(steps, batch=batch, lr=lr, seed=seed)
        if log:
            log("surrogate trained: %d steps, loss %.3e, rel. L2 error of S_iso %.3e" % (steps, rep[0], rep[2]))
        params, norm = e.surrogate_get()
    report = dict(loss=float(rep[0]), samples=int(rep[1]), rel_err=float(rep[2]), steps=int(rep[3]),
                  elements=ne, held=held)
    return params, norm, report
