"""Synthetic workloads of the benchmark configurations, as engine inputs.

Each builder returns a dict with X, conn, pid, material_id, props (9 per
part), dof_motion, motions, v0 and, where the laws need them, fibres.
"""
import numpy as np

from . import (MOTION_FIXED, MOTION_LINEAR, MOTION_ROT_HALFSINE, mesh_block, mesh_ellipsoid, motion)

IMPACT_SEED = 20261202


def props_row(rho, mu, lam, *rest):
    p = np.zeros(9)
    p[:3] = [rho, mu, lam]
    p[3:3 + len(rest)] = rest
    return p


def impact_block(n=100, slab=10, h=1e-3, jitter=0.15, seed=IMPACT_SEED):
    """n^3 jittered hex block (h = 1 mm), compressible neo-Hookean G = 1 MPa,
    K = 10 MPa, rho = 1000; base face fixed; the top `slab` element layers
    start at v_z = -5 m/s."""
    X, conn = mesh_block(n, n, n, h, jitter=jitter, seed=seed)
    nn, ne = X.size // 3, conn.size // 8
    G, K = 1.0e6, 10.0e6
    Xr = X.reshape(-1, 3)
    dm = -np.ones(3 * nn, np.int32)
    base = np.nonzero(Xr[:, 2] < 1e-9 * h)[0]
    for d in range(3):
        dm[3 * base + d] = 0
    v0 = np.zeros(3 * nn)
    v0[2::3][Xr[:, 2] >= (n - slab) * h - 0.2 * h] = -5.0
    return dict(X=X, conn=conn, pid=np.zeros(ne, np.int32), material_id=np.array([1], np.int32),
                props=props_row(1000.0, G, K - 2.0 * G / 3.0)[None, :], dof_motion=dm,
                motions=motion(MOTION_FIXED)[None, :], v0=v0, fibres=None)


def benchmark_cube(n=27, side=0.005, dmax=0.007, tmax=0.1):
    """The reference's published benchmark problem (examples/Benchmarking-
    Parallel: Benchmarking-Parallel.cpp:68-160 with 27elements.inp): an n^3
    cube of side 5 mm, neo-Hookean rho 1040, mu = lambda = 100 Pa, symmetry
    planes x = y = z = 0 and the face y = side pulled at dmax / tmax to
    tmax = 0.1 s, dt = 0.8 x stable.  The mesh is generated (regular grid),
    not read from the deck, whose coordinates agree with it to their nine
    printed digits."""
    X, conn = mesh_block(n, n, n, side / n)
    nn, ne = X.size // 3, conn.size // 8
    Xr = X.reshape(-1, 3)
    tol = 1e-5
    dm = -np.ones(3 * nn, np.int32)
    for d in range(3):
        dm[3 * np.nonzero(np.abs(Xr[:, d]) < tol)[0] + d] = 0
    dm[3 * np.nonzero(np.abs(Xr[:, 1] - side) < tol)[0] + 1] = 1
    motions = np.stack([motion(MOTION_FIXED), motion(MOTION_LINEAR, dmax / tmax)])
    return dict(X=X, conn=conn, pid=np.zeros(ne, np.int32), material_id=np.array([1], np.int32),
                props=props_row(1040.0, 100.0, 100.0)[None, :], dof_motion=dm, motions=motions, v0=None,
                fibres=None, tmax=tmax, dt_factor=0.8, failure_dt=1e-11)


def ellipsoid_head(n=126, semi_axes=(0.090, 0.075, 0.065), peak=6000.0, duration=0.010, seed=7):
    """Mapped-cube ellipsoid (n^3 elements; n = 126 gives 2.0M) in three
    regions by normalised radius r of the element centroid:
      r > 0.93          linear elastic shell, E = 6 GPa, nu = 0.22, rho = 2000
      0.3 < r <= 0.93   1-term Ogden mu = 530 Pa, alpha = -4.4, K = 2.1 MPa,
                        rho = 1040, with Prony g = 0.5 / 0.3, tau = 2 / 20 ms
      r <= 0.3          anisotropic HGO k1 = 1 kPa, k2 = 10, kappa = 0.1
                        (mu = 530 Pa, K = 2.1 MPa, rho = 1040), seeded unit
                        fibre directions
    Outer-surface nodes follow a rigid rotation about the x axis through the
    centre with angular acceleration peak sin(pi t / duration) for t <=
    duration."""
    a, b, c = semi_axes
    X, conn, r = mesh_ellipsoid(n, a, b, c)
    nn, ne = X.size // 3, conn.size // 8
    pid = np.where(r > 0.93, 0, np.where(r > 0.3, 1, 2)).astype(np.int32)
    E, nu = 6.0e9, 0.22
    mu_s, lam_s = E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))
    mu, K = 530.0, 2.1e6
    props = np.stack([props_row(2000.0, mu_s, lam_s),
                      props_row(1040.0, mu, K - 2 * mu / 3, -4.4, 0.0, 0.5, 2e-3, 0.3, 2e-2),
                      props_row(1040.0, mu, K - 2 * mu / 3, 1.0e3, 10.0, 0.1)])
    # surface nodes: the cube's faces (node grid indices 0 or n)
    s = n + 1
    idx = np.arange(nn)
    i, j, k = idx % s, (idx // s) % s, idx // (s * s)
    surf = (i == 0) | (i == n) | (j == 0) | (j == n) | (k == 0) | (k == n)
    dm = -np.ones(3 * nn, np.int32)
    for d in range(3):
        dm[3 * np.nonzero(surf)[0] + d] = 0
    motions = motion(MOTION_ROT_HALFSINE, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, peak, duration)[None, :]
    rng = np.random.default_rng(seed)
    fib = rng.normal(size=(ne, 3))
    fib /= np.linalg.norm(fib, axis=1)[:, None]
    return dict(X=X, conn=conn, pid=pid, material_id=np.array([3, 7, 8], np.int32), props=props,
                dof_motion=dm, motions=motions, v0=None, fibres=fib.reshape(-1), radius=r)


def concatenate(instances):
    """One engine input from independent instances: nodes, elements, parts
    and motions renumbered; returns (merged dict, node_offsets, elem_offsets)."""
    X, conn, pid, mid, props, dm, motions, v0, fib = [], [], [], [], [], [], [], [], []
    node_off, elem_off = [0], [0]
    n_parts = n_motions = 0
    any_fib = any(w.get("fibres") is not None for w in instances)
    for w in instances:
        nn, ne = w["X"].size // 3, w["conn"].size // 8
        X.append(w["X"])
        conn.append(w["conn"] + node_off[-1])
        pid.append(w["pid"] + n_parts)
        mid.append(w["material_id"])
        props.append(np.asarray(w["props"]).reshape(-1, 9))
        d = np.asarray(w["dof_motion"]).copy()
        d[d >= 0] += n_motions
        dm.append(d)
        motions.append(np.asarray(w["motions"]).reshape(-1, 16))
        v0.append(w["v0"] if w.get("v0") is not None else np.zeros(3 * nn))
        if any_fib:
            f = w.get("fibres")
            fib.append(f if f is not None else np.tile([1.0, 0.0, 0.0], ne))
        n_parts += len(w["material_id"])
        n_motions += motions[-1].shape[0]
        node_off.append(node_off[-1] + nn)
        elem_off.append(elem_off[-1] + ne)
    merged = dict(X=np.concatenate(X), conn=np.concatenate(conn).astype(np.int32),
                  pid=np.concatenate(pid).astype(np.int32), material_id=np.concatenate(mid).astype(np.int32),
                  props=np.concatenate(props), dof_motion=np.concatenate(dm).astype(np.int32),
                  motions=np.concatenate(motions), v0=np.concatenate(v0),
                  fibres=np.concatenate(fib) if any_fib else None)
    return merged, np.array(node_off, np.int32), np.array(elem_off, np.int32)


BATCH_SEED = 4242


def batch_instances(count=512, ne_range=(2000, 80000), t_end=0.020, seed=BATCH_SEED):
    """`count` independent impact simulations for one batched run.

    Each instance is, with equal odds, a jittered hex block (h = 1 mm,
    jitter 0.15, its own seed; compressible neo-Hookean) or a three-region
    ellipsoid (the regions and laws of ellipsoid_head).  Its element count is
    n^3 with n = round(Ne^(1/3)) and Ne drawn log-uniformly from ne_range.
    Every modulus (neo-Hookean G and K; shell E; Ogden / HGO mu and K; HGO
    k1) is scaled by its own uniform factor in [0.7, 1.3]; densities, the
    Ogden exponent, the Prony terms, k2 and kappa keep their nominal values.
    Loading: a half-sine angular acceleration about the x axis with peak
    drawn from [2, 8] krad/s^2 and duration from [5, 20] ms, applied to the
    ellipsoid's outer-surface nodes or to the block's base face (rotation
    about the base face's centre line).  Every instance runs to t_end.

    Returns (merged, node_offsets, elem_offsets, end_times, info) where info
    lists (kind, n, elements, peak, duration) per instance."""
    rng = np.random.default_rng(seed)
    inst, info = [], []
    lo, hi = np.log(ne_range[0]), np.log(ne_range[1])
    for i in range(count):
        ne_target = float(np.exp(rng.uniform(lo, hi)))
        n = max(2, int(round(ne_target ** (1.0 / 3.0))))
        peak = float(rng.uniform(2000.0, 8000.0))
        duration = float(rng.uniform(0.005, 0.020))
        f = rng.uniform(0.7, 1.3, size=4)
        if rng.uniform() < 0.5:
            w = impact_block(n=n, slab=0, seed=int(rng.integers(1, 2**62)))
            G, K = 1.0e6 * f[0], 10.0e6 * f[1]
            w["props"] = props_row(1000.0, G, K - 2.0 * G / 3.0)[None, :]
            Xr = w["X"].reshape(-1, 3)
            base = np.nonzero(Xr[:, 2] < 1e-9)[0]
            dm = -np.ones(Xr.shape[0] * 3, np.int32)
            for d in range(3):
                dm[3 * base + d] = 0
            c = 0.5 * n * 1e-3
            w["dof_motion"] = dm
            w["motions"] = motion(MOTION_ROT_HALFSINE, 1.0, 0.0, 0.0, 0.0, c, 0.0, peak, duration)[None, :]
            w["v0"] = None
            kind = "block"
        else:
            w = ellipsoid_head(n=n, peak=peak, duration=duration, seed=int(rng.integers(1, 2**62)))
            p = np.array(w["props"])
            E, nu = 6.0e9 * f[0], 0.22
            p[0, 1], p[0, 2] = E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))
            mu, K = 530.0 * f[1], 2.1e6 * f[2]
            for r in (1, 2):
                p[r, 1], p[r, 2] = mu, K - 2 * mu / 3
            p[2, 3] = 1.0e3 * f[3]
            w["props"] = p
            kind = "ellipsoid"
        info.append((kind, n, w["conn"].size // 8, peak, duration))
        inst.append(w)
    merged, node_off, elem_off = concatenate(inst)
    return merged, node_off, elem_off, np.full(count, t_end), info


# Reference-cube coordinates of the C3D8 local nodes (ShapeFunction_C3D8.cpp:22-29).
C3D8_SIGNS = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                       [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], np.int64)


def cube_rotations():
    """The 24 proper rotations of the reference cube as node renumberings:
    row r maps new local node a to old local node perm[r, a], the node that
    sits at R s_a (s_a: reference coordinates of node a).  The renumbered
    element x'(xi) = x(R xi) is the same hexahedron with det J of the same
    sign, and the 2x2x2 Gauss points map onto each other."""
    import itertools
    out = []
    for p in itertools.permutations(range(3)):
        for sg in itertools.product((-1, 1), repeat=3):
            R = np.zeros((3, 3), np.int64)
            for i in range(3):
                R[i, p[i]] = sg[i]
            if round(np.linalg.det(R)) != 1:
                continue
            img = C3D8_SIGNS @ R.T
            perm = [int(np.nonzero(np.all(C3D8_SIGNS == img[a], axis=1))[0][0]) for a in range(8)]
            out.append(perm)
    out = np.array(out, np.int64)
    assert out.shape == (24, 8) and len({tuple(r) for r in out}) == 24
    return out


def permute_elements(conn, seed):
    """Each element's local node order renumbered by a random proper cube
    rotation (the identity excluded), so that nodes are the same local corner
    of several elements and the engine's general force gather is used."""
    rots = cube_rotations()
    ident = np.nonzero(np.all(rots == np.arange(8), axis=1))[0][0]
    choices = np.delete(np.arange(24), ident)
    rng = np.random.default_rng(seed)
    c = np.asarray(conn, np.int64).reshape(-1, 8)
    pick = rots[rng.choice(choices, size=c.shape[0])]
    return np.take_along_axis(c, pick, axis=1).reshape(-1).astype(np.int32)
