"""Writes the golden-vector fixtures in tests/golden/ from the reference build.

Run in a container that has the FemTech source tree (default
/root/reference) and the image's MPICH/MKL:

    make -C oracle/ref REF=/root/reference      # builds oracle/_ref/ref_golden
    python tests/golden/make_golden.py [out_dir]   # default: tests/golden

Each fixture holds the inputs (mesh, materials, boundary motions as data, run
parameters) and the reference's outputs for them.  Outputs are produced by
FemTech's own library routines (compiled unmodified) driven by
oracle/ref/golden_driver.cpp.  The reference's Abaqus reports for ex9 and
the 10-element benchmark are parsed into the fixtures as data.
"""
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import femtech_amd as ft  # noqa: E402
import cases  # noqa: E402

REF = os.environ.get("FEMTECH_REF", "/root/reference")
DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_golden")


def run_driver(mode, d):
    env = dict(os.environ, MKL_NUM_THREADS="1", MKL_CBWR="COMPATIBLE", OMP_NUM_THREADS="1")
    res = subprocess.run([DRIVER, mode, d], capture_output=True, text=True, env=env, timeout=1800)
    if res.returncode != 0:
        raise RuntimeError("golden driver failed (%d): %s %s" % (res.returncode, res.stdout, res.stderr))
    return res.stdout


def write_materials(path, material_id, props):
    props = np.asarray(props).reshape(-1, 9)
    nprops = {0: 1, 1: 3, 2: 3, 3: 3, 4: 5, 5: 9}
    with open(path, "w") as f:
        for p, m in enumerate(material_id):
            vals = " ".join("%.17g" % v for v in props[p, :nprops[int(m)]])
            f.write("%d %d %s\n" % (p, int(m), vals))


def golden_run(name, c, mesh_file=None):
    d = tempfile.mkdtemp(prefix="golden_")
    try:
        nn = c["X"].size // 3
        if mesh_file:
            shutil.copy(mesh_file, os.path.join(d, "mesh" + os.path.splitext(mesh_file)[1]))
        else:
            ft.write_inp(os.path.join(d, "mesh.inp"), c["X"], c["conn"], c["pid"])
        write_materials(os.path.join(d, "materials.dat"), c["material_id"], c["props"])
        motions = np.asarray(c["motions"], np.float64).reshape(-1, 16)
        bc = np.concatenate([np.asarray(c["dof_motion"], np.int32), [motions.shape[0]]]).astype(np.int32)
        bc.tofile(os.path.join(d, "bc.bin"))
        motions.astype(np.float64).tofile(os.path.join(d, "motions.bin"))
        if c.get("v0") is not None:
            np.asarray(c["v0"], np.float64).tofile(os.path.join(d, "v0.bin"))
        if c.get("u0") is not None:
            np.asarray(c["u0"], np.float64).tofile(os.path.join(d, "u0.bin"))
        with open(os.path.join(d, "params.txt"), "w") as f:
            f.write("%d %.17g %.17g %.17g %d %s\n" % (c["max_steps"], c["tmax"], c["dt_factor"],
                                                      c["failure_dt"], len(c["snaps"]),
                                                      " ".join(str(s) for s in c["snaps"])))
        out = run_driver("run", d).split()
        steps = int(out[1])
        uid = out[out.index("uid") + 1]
        r = lambda f, t=np.float64: np.fromfile(os.path.join(d, f), dtype=t)  # noqa: E731
        res = dict(ref_coords=r("out_coords.f64"), ref_conn=r("out_conn.i32", np.int32),
                   ref_dshp=r("out_dshp.f64"), ref_detJ=r("out_detJ.f64"), ref_mass=r("out_mass.f64"),
                   ref_fi0=r("out_fi0.f64"), ref_a0=r("out_a0.f64"), ref_u=r("out_u.f64"),
                   ref_v=r("out_v.f64"), ref_a=r("out_a.f64"), ref_fi=r("out_fi.f64"),
                   ref_time=r("out_time.f64"), ref_dt=r("out_dt.f64"),
                   ref_energy=r("out_energy.f64").reshape(-1, 4),
                   ref_snap_u=r("out_snap_u.f64").reshape(-1, 3 * nn), ref_steps=np.int64(steps),
                   ref_eavg=r("out_eavg.f64"), ref_princ=r("out_princ.f64"))
        rows = []
        with open(os.path.join(d, "energy_%s.dat" % uid)) as f:
            for line in f:
                if not line.startswith("#") and line.strip():
                    rows.append([float(x) for x in line.split()])
        res["ref_energy_file"] = np.array(rows)
    finally:
        shutil.rmtree(d, ignore_errors=True)
    inputs = {k: np.asarray(v) for k, v in c.items() if v is not None and k != "snaps"}
    inputs["snaps"] = np.asarray(c["snaps"], np.int64)
    return inputs, res


def golden_ex7(mesh_file, materials_file):
    """ex7's own run (mesh, materials, loop and boundary routine of
    examples/ex7) through the golden driver's ex7 mode."""
    d = tempfile.mkdtemp(prefix="golden_ex7_")
    try:
        shutil.copy(mesh_file, os.path.join(d, "mesh.k"))
        shutil.copy(materials_file, os.path.join(d, "materials.dat"))
        out = run_driver("ex7", d).split()
        uid = out[out.index("uid") + 1]
        r = lambda f, t=np.float64: np.fromfile(os.path.join(d, f), dtype=t)  # noqa: E731
        res = dict(ref_coords=r("out_coords.f64"), ref_conn=r("out_conn.i32", np.int32), ref_u=r("out_u.f64"),
                   ref_v=r("out_v.f64"), ref_a=r("out_a.f64"), ref_fi=r("out_fi.f64"),
                   ref_time=r("out_time.f64"), ref_dt=r("out_dt.f64"),
                   ref_energy=r("out_energy.f64").reshape(-1, 4), ref_corner=r("out_plot.f64").reshape(-1, 4),
                   ref_steps=np.int64(int(out[1])), ref_plot_every=np.int64(int(out[out.index("plot_every") + 1])))
        rows = []
        with open(os.path.join(d, "energy_%s.dat" % uid)) as f:
            for line in f:
                if not line.startswith("#") and line.strip():
                    rows.append([float(x) for x in line.split()])
        res["ref_energy_file"] = np.array(rows)
    finally:
        shutil.rmtree(d, ignore_errors=True)
    return res


def golden_mesh_file(name, path, kind):
    """A golden run on one of the reference's own unstructured meshes, kept
    lean: inputs, mass, initial force, final state, dt and energy histories
    and CheckEnergy's rows (no shape-function arrays or snapshots).  The
    reader's arrays must equal the ones the reference read from the file
    (their SHA-256 digests are kept)."""
    import hashlib
    X, conn, pid, gid = ft.read_mesh(path)
    c = cases.case_mesh_file(X, conn, pid, kind)
    inp, res = golden_run(name, c, mesh_file=path)
    if not (np.array_equal(res["ref_coords"], X) and np.array_equal(res["ref_conn"], conn)):
        raise SystemExit("%s: the reader's arrays differ from the reference's" % name)
    keep = ("ref_mass", "ref_fi0", "ref_u", "ref_v", "ref_fi", "ref_time", "ref_dt", "ref_energy",
            "ref_energy_file", "ref_steps")
    out = {k: res[k] for k in keep}
    out["ref_coords_sha256"] = np.array(hashlib.sha256(np.ascontiguousarray(res["ref_coords"]).tobytes()).hexdigest())
    out["ref_conn_sha256"] = np.array(hashlib.sha256(np.ascontiguousarray(res["ref_conn"]).tobytes()).hexdigest())
    inp.pop("snaps", None)
    inp.pop("u0", None)
    save("run_" + name, **inp, **out)


def abaqus_rpt(path):
    rows = []
    with open(path) as f:
        for line in f:
            parts = line.split()
            if not parts:
                continue
            try:
                rows.append([float(p) for p in parts])
            except ValueError:
                continue
    return np.array(rows)


def stress_kats(rng):
    mats = []
    recs = []
    mu, lam = 2.0e4, 2.0e5
    plist = {
        1: cases.props_row(1000.0, mu, lam),
        2: cases.props_row(1000.0, mu, lam),
        3: cases.props_row(1000.0, mu, lam),
        4: cases.props_row(1000.0, mu, lam, 3.0e4, 8.0),
        5: cases.props_row(1000.0, mu, lam, 3.0e4, 8.0, 0.4, 2e-3, 0.3, 2e-2),
    }
    T = 6
    for m in (1, 2, 3, 4, 5):
        for scale in (1e-4, 1e-2, 0.15, 0.35):
            for _ in range(6):
                F0 = cases.random_F(rng, 1, scale)[0]
                seq = [np.eye(3) + (k + 1) / T * (F0 - np.eye(3)) for k in range(T)]
                dt = 1e-5 * (1 + rng.uniform())
                rec = [m] + list(plist[m]) + [dt, T]
                for F in seq:
                    rec += list(F.T.reshape(9))  # column-major
                recs.append(np.array(rec, np.float64))
                mats.append((m, plist[m], dt, np.array([F.T.reshape(9) for F in seq])))
    d = tempfile.mkdtemp(prefix="golden_")
    try:
        np.concatenate(recs).tofile(os.path.join(d, "stress_in.f64"))
        run_driver("stress", d)
        out = np.fromfile(os.path.join(d, "stress_out.f64")).reshape(len(recs), T, 7)
        hist = np.fromfile(os.path.join(d, "stress_hist.f64")).reshape(len(recs), 27)
    finally:
        shutil.rmtree(d, ignore_errors=True)
    return dict(material=np.array([m[0] for m in mats], np.int32),
                props=np.array([m[1] for m in mats]), dt=np.array([m[2] for m in mats]),
                F=np.array([m[3] for m in mats]), ref_pk2=out[:, :, :6], ref_detF=out[:, :, 6],
                ref_hist=hist)


def geom_kats(rng):
    xe = cases.random_hexes(rng, 240)
    d = tempfile.mkdtemp(prefix="golden_")
    try:
        xe.astype(np.float64).tofile(os.path.join(d, "geom_in.f64"))
        run_driver("geom", d)
        out = np.fromfile(os.path.join(d, "geom_out.f64")).reshape(len(xe), 8)
    finally:
        shutil.rmtree(d, ignore_errors=True)
    return dict(xe=xe, ref_volume=out[:, 0], ref_area=out[:, 1:7], ref_char_length=out[:, 7])


OUT_DIR = HERE


def save(name, **arrays):
    path = os.path.join(OUT_DIR, name + ".npz")
    np.savez_compressed(path, **arrays)
    print("wrote %s (%d bytes)" % (path, os.path.getsize(path)))


def main():
    global OUT_DIR
    if len(sys.argv) > 1:  # write the fixtures elsewhere (e.g. to compare them)
        OUT_DIR = sys.argv[1]
        os.makedirs(OUT_DIR, exist_ok=True)
    if not os.path.exists(DRIVER):
        raise SystemExit("build the reference first: make -C oracle/ref REF=%s" % REF)
    rng = np.random.default_rng(20261202)
    save("kat_stress", **stress_kats(rng))
    save("kat_geom", **geom_kats(rng))

    # ex9: the reference's own one-element cube and its Abaqus report
    kfile = os.path.join(REF, "examples", "ex9", "1-elt-cube.k")
    X, conn, pid, gid = ft.read_mesh(kfile)
    c = cases.case_ex9_like(X, conn, pid, np.array([1], np.int32),
                            cases.props_row(1040.0, 100.0, 100.0))
    inp, res = golden_run("ex9", c, mesh_file=kfile)
    res["abaqus_rpt"] = abaqus_rpt(os.path.join(REF, "examples", "ex9", "abaqus", "abaqus.rpt"))
    save("run_ex9", **inp, **res)

    # 10-element benchmark mesh (Benchmarking-Parallel) and its Abaqus report
    ifile = os.path.join(REF, "examples", "Benchmarking-Parallel", "10elements.inp")
    X, conn, pid, gid = ft.read_mesh(ifile)
    c = cases.case_ex9_like(X, conn, pid, np.array([1], np.int32),
                            cases.props_row(1040.0, 100.0, 100.0), tmax=0.1)
    c["snaps"] = [0, 1, 10, 100]
    inp, res = golden_run("bench10", c, mesh_file=ifile)
    res["abaqus_rpt"] = abaqus_rpt(os.path.join(REF, "examples", "Benchmarking-Parallel", "abaqus",
                                                "10elements.rpt"))
    res["global_node_id"] = gid
    save("run_bench10", **inp, **res)

    # the 27^3 benchmark mesh itself (the reference's published benchmark),
    # kept lean, with its Abaqus report
    ifile = os.path.join(REF, "examples", "Benchmarking-Parallel", "27elements.inp")
    X, conn, pid, gid = ft.read_mesh(ifile)
    c = cases.case_ex9_like(X, conn, pid, np.array([1], np.int32),
                            cases.props_row(1040.0, 100.0, 100.0), tmax=0.1)
    c["snaps"] = []
    inp, res = golden_run("bench27", c, mesh_file=ifile)
    if not (np.array_equal(res["ref_coords"], X) and np.array_equal(res["ref_conn"], conn)):
        raise SystemExit("bench27: the reader's arrays differ from the reference's")
    keep = ("ref_u", "ref_time", "ref_dt", "ref_energy", "ref_energy_file", "ref_steps")
    out = {k: res[k] for k in keep}
    out["abaqus_rpt"] = abaqus_rpt(os.path.join(REF, "examples", "Benchmarking-Parallel", "abaqus",
                                                "27elements.rpt"))
    inp.pop("snaps", None)
    inp.pop("u0", None)
    inp.pop("v0", None)
    save("run_bench27", **inp, **out)

    # ex7: the reference's two-element column under its own loop, with the
    # Abaqus corner-node history (TensionNoDamping.dat: time, U1, U2, U3)
    ex7 = os.path.join(REF, "examples", "ex7")
    X, conn, pid, gid = ft.read_mesh(os.path.join(ex7, "2elts-hex.k"))
    res = golden_ex7(os.path.join(ex7, "2elts-hex.k"), os.path.join(ex7, "materials.dat"))
    mat, props = ft.read_materials(os.path.join(ex7, "materials.dat"), 1)
    res["abaqus_corner"] = np.loadtxt(os.path.join(ex7, "TensionNoDamping.dat"))
    res["abaqus_rpt"] = abaqus_rpt(os.path.join(ex7, "abaqus", "abaqus.rpt"))
    save("run_ex7", X=X, conn=conn, pid=pid, material_id=mat, props=props, tmax=np.float64(1.0),
         dmax=np.float64(0.001), length=np.float64(0.005), plots=np.int64(50), **res)

    # two of the reference's own unstructured meshes
    golden_mesh_file("brain", os.path.join(REF, "examples", "ex5", "brain.inp"), "brain")
    golden_mesh_file("spheres", os.path.join(REF, "examples", "ex14", "sphere-in-sphere-hex.inp"), "spheres")

    inp, res = golden_run("c1", cases.case_c1())
    save("run_c1", **inp, **res)
    inp, res = golden_run("impact", cases.case_impact())
    save("run_impact", **inp, **res)
    inp, res = golden_run("materials", cases.case_materials())
    save("run_materials", **inp, **res)


if __name__ == "__main__":
    main()
