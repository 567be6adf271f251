"""CPU-baseline calibration: the oracle (the FP64 C++ restatement timed as
the CPU baseline on the GPU host) against the reference build itself, on the
same cases, in a container that has the FemTech source tree.

    make -C oracle/ref REF=/root/reference
    python tests/golden/calibrate_cpu.py [report.txt]   # default: tests/golden/cpu_calibration.txt

Cases: the 27^3 Benchmarking-Parallel cube (examples/Benchmarking-Parallel/
27elements.inp, rho 1040, mu = lambda = 100 Pa, stretched 7 mm over
tMax = 0.1 s: 445 steps) and the 10x10x10 neo-Hookean block of 2,000 steps.
The reference is the golden driver's `time` mode (the ex9 loop, ex9.cpp:
107-175, over FemTech's own routines, MKL on one thread) under MPICH at 1 and
8 ranks; the oracle runs the same loop with 1 and 8 OpenMP threads.  Only the
step loop is timed on both sides; each figure is the best of `REPEATS` runs.
"""
import os
import platform
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import femtech_amd as ft  # noqa: E402
import cases  # noqa: E402
from oracle import oracle as O  # noqa: E402
from make_golden import DRIVER, REF, write_materials  # noqa: E402

MPIEXEC = os.environ.get("MPIEXEC", "/opt/conda/bin/mpiexec")
REPEATS = 3


def case_dir(c, mesh_file=None):
    d = tempfile.mkdtemp(prefix="calib_")
    if mesh_file:
        shutil.copy(mesh_file, os.path.join(d, "mesh.inp"))
    else:
        ft.write_inp(os.path.join(d, "mesh.inp"), c["X"], c["conn"], c["pid"])
    write_materials(os.path.join(d, "materials.dat"), c["material_id"], c["props"])
    motions = np.asarray(c["motions"], np.float64).reshape(-1, 16)
    np.concatenate([np.asarray(c["dof_motion"], np.int32), [motions.shape[0]]]).astype(np.int32).tofile(
        os.path.join(d, "bc.bin"))
    motions.tofile(os.path.join(d, "motions.bin"))
    np.asarray(c["X"], np.float64).tofile(os.path.join(d, "coords.bin"))
    with open(os.path.join(d, "params.txt"), "w") as f:
        f.write("%d %.17g %.17g %.17g 0\n" % (c["max_steps"], c["tmax"], c["dt_factor"], c["failure_dt"]))
    return d


def reference_loop(d, ranks):
    env = dict(os.environ, MKL_NUM_THREADS="1", OMP_NUM_THREADS="1")
    best = None
    for _ in range(REPEATS):
        res = subprocess.run([MPIEXEC, "-n", str(ranks), DRIVER, "time", d], capture_output=True, text=True,
                             env=env, timeout=1800)
        if res.returncode != 0:
            raise RuntimeError("reference run failed (%d): %s %s" % (res.returncode, res.stdout[-2000:],
                                                                     res.stderr[-2000:]))
        line = [ln for ln in res.stdout.splitlines() if ln.startswith("steps ")][-1].split()
        steps, secs = int(line[1]), float(line[5])
        best = secs if best is None else min(best, secs)
    return steps, best


def oracle_loop(c, threads):
    best = None
    for _ in range(REPEATS):
        o = O.Model(c["X"], c["conn"], c["pid"], c["material_id"], c["props"], threads=threads)
        o.set_bc(c["dof_motion"], c["motions"])
        o.set_initial(None, None)
        t0 = time.perf_counter()
        o.run(0, c["tmax"], c["dt_factor"], c["failure_dt"])
        t1 = time.perf_counter()
        steps = o.run(c["max_steps"], c["tmax"], c["dt_factor"], c["failure_dt"])["steps"]
        t2 = time.perf_counter()
        o.close()
        secs = (t2 - t1) - (t1 - t0)
        best = secs if best is None else min(best, secs)
    return steps, best


def cpu_model():
    try:
        for line in subprocess.run(["lscpu"], capture_output=True, text=True).stdout.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def main():
    out = sys.argv[1] if len(sys.argv) > 1 else os.path.join(HERE, "cpu_calibration.txt")
    if not os.path.exists(DRIVER):
        raise SystemExit("build the reference first: make -C oracle/ref REF=%s" % REF)
    mesh27 = os.path.join(REF, "examples", "Benchmarking-Parallel", "27elements.inp")
    X, conn, pid, _ = ft.read_mesh(mesh27)
    c27 = cases.case_ex9_like(X, conn, pid, np.array([1], np.int32), cases.props_row(1040.0, 100.0, 100.0),
                              tmax=0.1)
    rows = []
    for name, c, mesh in (("27^3 benchmark cube", c27, mesh27), ("10x10x10 neo-Hookean block", cases.case_c1(),
                                                                 None)):
        ne = c["conn"].size // 8
        d = case_dir(c, mesh)
        try:
            for n in (1, 8):
                rs, rt = reference_loop(d, n)
                os_, ot = oracle_loop(c, n)
                rows.append((name, ne, n, rs, rt, os_, ot))
                print(rows[-1], flush=True)
        finally:
            shutil.rmtree(d, ignore_errors=True)
    with open(out, "w") as f:
        f.write("# CPU-baseline calibration (tests/golden/calibrate_cpu.py)\n")
        f.write("# host: %s, nproc %d; reference: golden driver `time` mode under MPICH (MKL 1 thread); "
                "oracle: OpenMP threads; step loop only, best of %d\n" % (cpu_model(), os.cpu_count(), REPEATS))
        f.write("%-28s %9s %5s %9s %10s %12s %9s %10s %12s %8s\n" % (
            "case", "elements", "n", "ref_steps", "ref_s", "ref_el_st/s", "ora_steps", "ora_s", "ora_el_st/s",
            "ora/ref"))
        for name, ne, n, rs, rt, os_, ot in rows:
            rv, ov = ne * rs / rt, ne * os_ / ot
            f.write("%-28s %9d %5d %9d %10.3f %12.4g %9d %10.3f %12.4g %8.2f\n" % (
                name, ne, n, rs, rt, rv, os_, ot, ov, ov / rv))
    print(open(out).read())


if __name__ == "__main__":
    main()
