"""Canonical test problems shared by the golden-vector generator and tests.

Each builder returns a dict of plain arrays (mesh, materials, boundary motions
as data, run parameters).  Materials follow ReadMaterials.cpp:43-122
(9 properties per part: rho, mu, lambda, k1, k2, g1, tau1, g2, tau2).
"""
import numpy as np

import femtech_amd as ft

TOL = 1e-5  # coordinate test tolerance of the reference drivers (ex9.cpp:188)


def props_row(rho, mu=0.0, lam=0.0, k1=0.0, k2=0.0, g1=0.0, t1=0.0, g2=0.0, t2=0.0):
    return np.array([rho, mu, lam, k1, k2, g1, t1, g2, t2], np.float64)


def plane_rules(X, rules, nn):
    """rules: (axis, value, dof, motion index) applied in order (later wins),
    like the coordinate tests of ex9.cpp:199-243."""
    X = X.reshape(-1, 3)
    dm = -np.ones(3 * nn, np.int32)
    for axis, value, dof, mi in rules:
        nodes = np.nonzero(np.abs(X[:, axis] - value) < TOL)[0]
        dm[3 * nodes + dof] = mi
    return dm


def case_ex9_like(X, conn, pid, material_id, props, L=0.005, dmax=0.007, tmax=1.0, steps=100000):
    """ex9's loading (ex9.cpp:187-248): symmetry planes x=y=z=0 and a
    prescribed y-displacement Time*dMax/tMax on the face y = L."""
    nn = X.size // 3
    motions = np.stack([ft.motion(ft.MOTION_FIXED), ft.motion(ft.MOTION_LINEAR, dmax / tmax)])
    dm = plane_rules(X, [(0, 0.0, 0, 0), (1, 0.0, 1, 0), (2, 0.0, 2, 0), (1, L, 1, 1)], nn)
    return dict(X=X, conn=conn, pid=pid, material_id=material_id, props=props, dof_motion=dm,
                motions=motions, u0=None, v0=None, max_steps=steps, tmax=tmax, dt_factor=0.8,
                failure_dt=1e-11, snaps=[0, 1, 10, 50])


def case_c1(steps=2000):
    """10x10x10 block, h = 1 mm; neo-Hookean G = 1 MPa, K = 10 MPa, rho = 1000;
    bottom face u_z = 0 with one corner pinned; top face u_z ramps to 10 %
    stretch over 1 ms; dt = 0.9 x stable, 2000 steps."""
    X, conn = ft.mesh_block(10, 10, 10, 1e-3)
    nn = X.size // 3
    G, K = 1.0e6, 10.0e6
    props = props_row(1000.0, G, K - 2.0 * G / 3.0)[None, :]
    motions = np.stack([ft.motion(ft.MOTION_FIXED), ft.motion(ft.MOTION_RAMP_HOLD, 1e-3, 1e-3)])
    Xr = X.reshape(-1, 3)
    dm = plane_rules(X, [(2, 0.0, 2, 0), (2, 0.01, 2, 1)], nn)
    corner = np.nonzero(np.all(np.abs(Xr) < TOL, axis=1))[0]
    dm[3 * corner + 0] = 0
    dm[3 * corner + 1] = 0
    dm[3 * corner + 2] = 0
    return dict(X=X, conn=conn, pid=np.zeros(conn.size // 8, np.int32),
                material_id=np.array([1], np.int32), props=props, dof_motion=dm, motions=motions,
                u0=None, v0=None, max_steps=steps, tmax=1e9, dt_factor=0.9, failure_dt=1e-11,
                snaps=[0, 1, 10, 100, 500, 1000, 2000])


def case_impact(nx=6, ny=5, nz=4, jitter=0.15, seed=20261202, steps=300, slab=1):
    """Jittered block (h = 1 mm) with the impact loading of the 1M-hex
    workload: base face fixed, the top `slab` element layers start at
    v_z = -5 m/s; neo-Hookean G = 1 MPa, K = 10 MPa."""
    X, conn = ft.mesh_block(nx, ny, nz, 1e-3, jitter=jitter, seed=seed)
    nn = X.size // 3
    G, K = 1.0e6, 10.0e6
    props = props_row(1000.0, G, K - 2.0 * G / 3.0)[None, :]
    motions = np.stack([ft.motion(ft.MOTION_FIXED)])
    dm = plane_rules(X, [(2, 0.0, 0, 0), (2, 0.0, 1, 0), (2, 0.0, 2, 0)], nn)
    v0 = np.zeros(3 * nn)
    Xr = X.reshape(-1, 3)
    top = np.nonzero(Xr[:, 2] >= (nz - slab) * 1e-3 - 0.2e-3)[0]
    v0[3 * top + 2] = -5.0
    return dict(X=X, conn=conn, pid=np.zeros(conn.size // 8, np.int32),
                material_id=np.array([1], np.int32), props=props, dof_motion=dm, motions=motions,
                u0=None, v0=v0, max_steps=steps, tmax=1e9, dt_factor=0.9, failure_dt=1e-11,
                snaps=[0, 1, 10, 100, steps])


def case_materials(steps=250):
    """4x4x6 jittered block in six parts (one per element layer): neo-Hookean,
    St Venant-Kirchhoff, linear elastic, HGO isotropic, HGO + Prony and rigid;
    stretched 10 % along z by a prescribed top motion with symmetry planes,
    plus a lateral initial velocity."""
    X, conn = ft.mesh_block(4, 4, 6, 1e-3, jitter=0.1, seed=7)
    nn, ne = X.size // 3, conn.size // 8
    # part = element layer (z index), parts first appear in increasing order
    pid = (np.arange(ne) // 16).astype(np.int32)
    mu, lam = 2.0e4, 2.0e5
    props = np.stack([
        props_row(1000.0, mu, lam),                                    # 1 NH
        props_row(1000.0, mu, lam),                                    # 2 SVK
        props_row(1000.0, mu, lam),                                    # 3 LE
        props_row(1000.0, mu, lam, 3.0e4, 8.0),                        # 4 HGO
        props_row(1000.0, mu, lam, 3.0e4, 8.0, 0.4, 2e-3, 0.3, 2e-2),  # 5 HGO + Prony
        props_row(1000.0),                                             # 0 rigid
    ])
    material_id = np.array([1, 2, 3, 4, 5, 0], np.int32)
    motions = np.stack([ft.motion(ft.MOTION_FIXED), ft.motion(ft.MOTION_RAMP_HOLD, 0.6e-3, 4e-3)])
    dm = plane_rules(X, [(0, 0.0, 0, 0), (1, 0.0, 1, 0), (2, 0.0, 2, 0), (2, 6e-3, 2, 1)], nn)
    v0 = np.zeros(3 * nn)
    Xr = X.reshape(-1, 3)
    v0[0::3] = 0.5 * Xr[:, 2] / 6e-3
    return dict(X=X, conn=conn, pid=pid, material_id=material_id, props=props.reshape(-1),
                dof_motion=dm, motions=motions, u0=None, v0=v0, max_steps=steps, tmax=1e9,
                dt_factor=0.8, failure_dt=1e-11, snaps=[0, 1, 10, 100, steps])


def random_F(rng, n, scale):
    F = np.eye(3)[None, :, :] + scale * rng.uniform(-1, 1, (n, 3, 3))
    # keep det F > 0
    for i in range(n):
        if np.linalg.det(F[i]) <= 0.2:
            F[i] = np.eye(3) + 0.1 * scale * rng.uniform(-1, 1, (3, 3))
    return F


def random_hexes(rng, n):
    """Distorted hexes around the unit-millimetre cube; a third get large
    negative twists so the parallelogram shortcut of Geometry.cpp:46 (no fabs)
    is taken on non-parallelogram faces."""
    base = 1e-3 * np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                            [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], np.float64)
    out = []
    for i in range(n):
        x = base + 1e-3 * rng.uniform(-0.2, 0.2, (8, 3)) * (i % 3 != 0)
        if i % 3 == 1:
            x[2] -= 1e-3 * np.array([0.3, 0.3, 0.3])
            x[6] -= 1e-3 * np.array([0.2, 0.25, 0.3])
        if i % 7 == 0:
            x = x @ np.array([[0.8, 0.6, 0.0], [-0.6, 0.8, 0.0], [0.0, 0.0, 1.0]])
        out.append(x.reshape(24))
    return np.array(out)


def case_rotation(n=5, h=1e-3, steps=600, peak=6000.0, T=0.01):
    """Jittered n^3 block whose outer-surface nodes follow a rigid rotation
    about the x axis through the block centre, with angular acceleration
    peak*sin(pi t/T) for t <= T and constant angular velocity afterwards
    (the half-sine pulse of the rotational loading); soft neo-Hookean
    interior (G = 10 kPa, K = 1 MPa, rho = 1040)."""
    X, conn = ft.mesh_block(n, n, n, h, jitter=0.15, seed=31)
    nn = X.size // 3
    Xr = X.reshape(-1, 3)
    L = n * h
    surf = np.any((np.abs(Xr) < TOL) | (np.abs(Xr - L) < TOL), axis=1)
    dm = -np.ones(3 * nn, np.int32)
    for d in range(3):
        dm[3 * np.nonzero(surf)[0] + d] = 0
    c = 0.5 * L
    motions = np.stack([ft.motion(ft.MOTION_ROT_HALFSINE, 1.0, 0.0, 0.0, c, c, c, peak, T)])
    G, K = 1.0e4, 1.0e6
    props = props_row(1040.0, G, K - 2.0 * G / 3.0)[None, :]
    return dict(X=X, conn=conn, pid=np.zeros(conn.size // 8, np.int32),
                material_id=np.array([1], np.int32), props=props, dof_motion=dm, motions=motions,
                u0=None, v0=None, max_steps=steps, tmax=1e9, dt_factor=0.9, failure_dt=1e-11,
                snaps=[steps])


def case_injury(steps=150):
    """Large-strain variant of the rotational loading (6^3 jittered block,
    h = 5 mm, G = 500 Pa, K = 20 kPa, peak 20 krad/s^2 over 5 ms): the
    strains cross the 15 % and 30 % thresholds and the strain rates the
    120 1/s and 28 1/s thresholds of the injury criteria."""
    c = case_rotation(n=6, h=5e-3, steps=steps, peak=20000.0, T=0.005)
    G, K = 500.0, 2.0e4
    c["props"] = props_row(1040.0, G, K - 2.0 * G / 3.0)[None, :]
    return c


# Proper cube rotations as local-node renumberings (the workload module's,
# shared with the bench's general-gather option).
from femtech_amd.workloads import C3D8_SIGNS, cube_rotations, permute_elements  # noqa: E402,F401


def corner_eligible(conn, nn):
    """True when every node is local corner k of at most one element for each
    k (the engine then writes forces to per-node corner slots)."""
    c = np.asarray(conn).reshape(-1, 8)
    return all(np.unique(c[:, k]).size == c.shape[0] for k in range(8))


def fan_mesh(nz=3, r0=1e-3, h=0.5e-3, jitter=0.08, seed=5):
    """Hexagonal prism meshed as an unstructured fan: per z layer, 6 kite
    quads around the axis (centre, edge midpoint, vertex, edge midpoint of a
    regular hexagon of radius r0) and a ring of 12 quads out to radius 2 r0,
    extruded over nz layers of height h.  Axis nodes of interior layers
    belong to 12 elements (valence above 8); interior nodes are jittered by
    +-jitter r0 in x and y.  Returns X (3 per node), conn, pid (0 inner fan,
    1 outer ring) and the per-layer node count."""
    ang = np.pi / 3.0 * np.arange(6)
    V = r0 * np.stack([np.cos(ang), np.sin(ang)], 1)
    M = 0.5 * (V + np.roll(V, -1, axis=0))
    W = 2.0 * V
    N = 0.5 * (W + np.roll(W, -1, axis=0))
    P = np.concatenate([[[0.0, 0.0]], V, M, W, N])  # 0 | 1-6 V | 7-12 M | 13-18 W | 19-24 N
    npl = P.shape[0]
    iv, im, iw, inn = 1, 7, 13, 19
    quads, parts = [], []
    for i in range(6):
        quads.append([0, im + (i - 1) % 6, iv + i, im + i])
        parts.append(0)
        quads.append([iv + i, iw + i, inn + i, im + i])
        parts.append(1)
        quads.append([im + i, inn + i, iw + (i + 1) % 6, iv + (i + 1) % 6])
        parts.append(1)
    fixed = []
    for q in quads:
        a = P[q]
        area = 0.5 * sum(a[k, 0] * a[(k + 1) % 4, 1] - a[(k + 1) % 4, 0] * a[k, 1] for k in range(4))
        fixed.append(q if area > 0 else q[::-1])
    rng = np.random.default_rng(seed)
    X = []
    for k in range(nz + 1):
        Pk = P.copy()
        if 0 < k < nz:
            Pk[:iw] += jitter * r0 * rng.uniform(-1, 1, (iw, 2))
        X.append(np.concatenate([Pk, np.full((npl, 1), k * h)], 1))
    X = np.concatenate(X).reshape(-1)
    conn, pid = [], []
    for k in range(nz):
        for q, p in zip(fixed, parts):
            conn.append([k * npl + n for n in q] + [(k + 1) * npl + n for n in q])
            pid.append(p)
    return X, np.array(conn, np.int32).reshape(-1), np.array(pid, np.int32), npl


def case_fan(steps=300):
    """fan_mesh under impact loading: base layer fixed, top layer starting at
    v_z = -5 m/s; neo-Hookean fan (G = 1 MPa, K = 10 MPa) inside an isotropic
    HGO + Prony ring (Prony time constants 2 and 20 ms, as case_materials)."""
    X, conn, pid, npl = fan_mesh()
    nn = X.size // 3
    props = np.stack([props_row(1000.0, 1.0e6, 1.0e7 - 2.0e6 / 3.0),
                      props_row(1000.0, 4.0e5, 4.0e6, 3.0e5, 6.0, 0.4, 2e-3, 0.3, 2e-2)])
    Xr = X.reshape(-1, 3)
    zmax = Xr[:, 2].max()
    dm = -np.ones(3 * nn, np.int32)
    base = np.nonzero(Xr[:, 2] < 1e-12)[0]
    for d in range(3):
        dm[3 * base + d] = 0
    v0 = np.zeros(3 * nn)
    v0[3 * np.nonzero(Xr[:, 2] > zmax - 1e-12)[0] + 2] = -5.0
    return dict(X=X, conn=conn, pid=pid, material_id=np.array([1, 5], np.int32), props=props.reshape(-1),
                dof_motion=dm, motions=ft.motion(ft.MOTION_FIXED)[None, :], u0=None, v0=v0,
                max_steps=steps, tmax=1e9, dt_factor=0.9, failure_dt=1e-11)


def case_mesh_file(X, conn, pid, kind):
    """Loading and materials for two of the reference's own unstructured
    meshes (the fixture holds the arrays the reader made of them):
    - "brain" (examples/ex5/brain.inp: 14,698 hexes in 10 parts, nodes of up
      to 12 elements): soft tissue, the lowest 15 % of the height fixed, the
      rest with an initial shearing velocity that grows with height;
    - "spheres" (examples/ex14/sphere-in-sphere-hex.inp: 9,317 hexes in 3
      parts, at most 8 elements per node): the nodes with x < -0.8 m fixed,
      the rest with an initial sideways velocity growing with x.
    The parts cycle through neo-Hookean, HGO, HGO + Prony, St Venant-
    Kirchhoff and linear elastic laws with moduli varied per part."""
    Xr = np.asarray(X).reshape(-1, 3)
    nn = Xr.shape[0]
    n_parts = int(np.max(pid)) + 1
    cycle = [1, 4, 5, 2, 3]
    mids, rows = [], []
    if kind == "brain":
        mu, lam, rho, k1, taus = 2.0e3, 2.0e4, 1040.0, 1.0e3, (2e-2, 2e-1)
        lo, hi = Xr[:, 2].min(), Xr[:, 2].max()
        s = (Xr[:, 2] - lo) / (hi - lo)
        fixed = s < 0.15
        v0 = np.zeros((nn, 3))
        v0[:, 0] = 0.4 * s
        v0[:, 2] = -0.3 * s
        steps = 200
    elif kind == "spheres":
        mu, lam, rho, k1, taus = 1.0e4, 1.0e5, 1000.0, 5.0e3, (5e-2, 5e-1)
        fixed = Xr[:, 0] < -0.8
        v0 = np.zeros((nn, 3))
        v0[:, 1] = 0.5 * (Xr[:, 0] + 1.0) / 2.0
        steps = 200
    else:
        raise KeyError(kind)
    for p in range(n_parts):
        m = cycle[p % len(cycle)]
        f = 1.0 + 0.1 * (p % 4)
        if m in (1, 2, 3):
            rows.append(props_row(rho, mu * f, lam * f))
        elif m == 4:
            rows.append(props_row(rho, mu * f, lam * f, k1, 5.0))
        else:
            rows.append(props_row(rho, mu * f, lam * f, k1, 5.0, 0.4, taus[0], 0.3, taus[1]))
        mids.append(m)
    dm = -np.ones(3 * nn, np.int32)
    for d in range(3):
        dm[3 * np.nonzero(fixed)[0] + d] = 0
    v0[fixed] = 0.0
    return dict(X=np.asarray(X, np.float64), conn=np.asarray(conn, np.int32), pid=np.asarray(pid, np.int32),
                material_id=np.array(mids, np.int32), props=np.array(rows), dof_motion=dm,
                motions=np.stack([ft.motion(ft.MOTION_FIXED)]), u0=None, v0=v0.reshape(-1), max_steps=steps,
                tmax=1e9, dt_factor=0.8, failure_dt=1e-11, snaps=[])
