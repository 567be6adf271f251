"""FemTech's C++ call surface (include/FemTech.h, libFemTech.so), CPU side.

The library must export the reference's C++ names (include/FemTech.h:22-129
of the reference) with the same mangling, so that FemTech drivers link
unchanged; the reference's own example drivers are compiled and linked here
when the FemTech source tree is present.  Without a GPU the library must stop
with an error instead of computing anything on the host.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import femtech_amd as ft
from conftest import ROOT, golden

LIB_DIR = os.path.join(ROOT, "femtech_amd", "_lib")
COMPAT = os.path.join(LIB_DIR, "libFemTech.so")
DRIVER = os.path.join(LIB_DIR, "explicit_tension")
REF = os.environ.get("FEMTECH_REF", "/root/reference")

# Functions of the explicit path that drivers call (FemTech.h:22-129) and the
# logging functions behind the FILE_LOG macros (parallel_log.h:20-45).
PATH_FUNCTIONS = [
    "InitFemTechWoInput(int, char**)", "FinalizeFemTech()", "TerminateFemTech(int)",
    "ReadInputFile(char const*)", "ReadMaterials()", "PartitionMesh()", "AllocateArrays()",
    "ShapeFunctions()", "GaussQuadrature3D(int, int, double*, double*)", "AssembleLumpedMass()",
    "StableTimeStep()", "GetForce()", "GetForce_3D()", "CalculateAccelerations()",
    "CheckEnergy(double, int)", "CalculateStrain()",
    "CalculateMaximumPrincipalStrain(int, double*, double*, double*)",
    "WriteVTU(char const*, int, int**, int, char const**)", "WritePVD(char const*, int)", "FreeArrays()",
    "fileLog(logLevel, char const*, ...)", "fileLogMaster(logLevel, char const*, ...)",
    "fileLogSingle(logLevel, char const*, ...)",
    "fileLogMatrixRM(logLevel, double const*, int, int, char const*, ...)",
    "initLog(char const*)", "finaliseLog()",
    # per-element, per-Gauss-point routines (FemTech.h:29-104)
    "ShapeFunction_C3D8(int, int, double*, double*)", "CalculateTimeStep(int)",
    "CalculateDeformationGradient(int, int)", "StrainDisplacementMatrix(int, int, int, double*)",
    "StressUpdate(int, int)", "DeterminateF(int, int)", "InverseF(int, int, double*)",
    "InternalForceUpdate(int, int, double*)", "StVenantKirchhoff(int, int)",
    "CompressibleNeoHookean(int, int)", "LinearElastic(int, int)", "HGOIsotropic(int, int)",
    "HGOIsotropicViscoelastic(int, int)", "inverse3x3Matrix(double*, double*, double*)",
    "CalculateCharacteristicLength(int)", "CalculateCharacteristicLength_C3D8(int)",
]
GLOBALS = ["coordinates", "connectivity", "displacements", "velocities", "velocities_half",
           "shp", "dshp", "detJacobian", "gaussWeights", "F", "detF", "invF", "pk2", "gptr", "dsptr",
           "gpPtr", "fptr", "pk2ptr", "detFptr", "B", "fintGQ", "Hn_1", "Hn_2", "S0n",
           "accelerations", "boundary", "mass", "fi", "fe", "f_net", "fi_prev", "fe_prev",
           "displacements_prev", "accelerations_prev", "Eavg", "stepTime", "nNodes", "nelements",
           "ndim", "nDOF", "nPIDglobal", "materialID", "properties", "pid", "eptr", "energyFile",
           "world_rank", "world_size", "uid"]
DRIVER_DEFINED = ["Time", "dt", "ExplicitTimeStepReduction", "FailureTimeStep", "ImplicitStatic",
                  "ImplicitDynamic", "ExplicitDynamic"]


def nm(*args):
    return subprocess.run(["nm", "-D", "-C"] + list(args) + [COMPAT], capture_output=True, text=True,
                          check=True).stdout


def test_exports_the_reference_names():
    defined = nm("--defined-only")
    text = {line.split(None, 2)[2] for line in defined.splitlines() if len(line.split(None, 2)) == 3}
    missing = [f for f in PATH_FUNCTIONS if f not in text]
    assert not missing, missing
    missing = [g for g in GLOBALS if not any(t == g or t.startswith(g + "[") for t in text)]
    assert not missing, missing
    # the driver owns these (ex9.cpp:9-17): the library must leave them undefined
    undefined = nm("--undefined-only")
    for g in DRIVER_DEFINED:
        assert g not in text, g
    for g in ("dt", "FailureTimeStep", "ImplicitDynamic", "ExplicitDynamic"):  # read by the library
        assert re.search(r"U %s$" % g, undefined, re.M), g


def test_every_compute_routine_goes_through_the_engine():
    """The compat layer calls the C ABI for every arithmetic routine of the
    path; it has no host implementation of forces, time step, mass, strain,
    accelerations or energy sums."""
    src = open(os.path.join(ROOT, "femtech_amd", "csrc", "ft_femtech.cpp")).read()
    for routine, call in [("ShapeFunctions", "ft_engine_create"), ("AssembleLumpedMass", "ft_assemble_lumped_mass"),
                          ("StableTimeStep", "ft_stable_time_step"), ("GetForce_3D", "ft_get_force"),
                          ("CalculateAccelerations", "ft_accelerations"), ("CheckEnergy", "ft_energy_terms"),
                          ("CalculateStrain", "ft_element_strain"),
                          ("ShapeFunction_C3D8", "point(FT_PT_SHAPE"),
                          ("CalculateDeformationGradient", "point(FT_PT_DEFGRAD"),
                          ("DeterminateF", "point(FT_PT_DETF"), ("InverseF", "point(FT_PT_INVF"),
                          ("StrainDisplacementMatrix", "point(FT_PT_BMAT"),
                          ("InternalForceUpdate", "point(FT_PT_FINT"),
                          ("CalculateCharacteristicLength_C3D8", "point(FT_PT_CHARLEN"),
                          ("CalculateTimeStep", "point(FT_PT_TIMESTEP")]:
        m = re.search(r"^(?:void|double) %s\([^)]*\) \{(.*?)^\}" % routine, src, re.M | re.S)
        assert m, routine
        assert call in m.group(1), routine


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "examples")), reason="FemTech sources not present")
@pytest.mark.parametrize("driver", ["ex9/ex9.cpp", "ex7/ex7.cpp",
                                    "Benchmarking-Parallel/Benchmarking-Parallel.cpp"])
def test_reference_drivers_compile_and_link_unmodified(tmp_path, driver):
    exe = str(tmp_path / "drv")
    res = subprocess.run(["g++", "-O1", "-std=c++17", "-w", "-I", os.path.join(ROOT, "include"),
                          os.path.join(REF, "examples", driver), "-L", LIB_DIR, "-lFemTech", "-lfemtech_amd",
                          "-Wl,-rpath," + LIB_DIR, "-o", exe], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-2000:]
    assert os.path.exists(exe)


def write_case(d, z):
    ft.write_inp(os.path.join(d, "mesh.inp"), z["X"], z["conn"], z["pid"])
    nprops = {0: 1, 1: 3, 2: 3, 3: 3, 4: 5, 5: 9}
    props = np.asarray(z["props"]).reshape(-1, 9)
    with open(os.path.join(d, "materials.dat"), "w") as f:
        for p, m in enumerate(np.asarray(z["material_id"])):
            f.write("%d %d %s\n" % (p, int(m), " ".join("%.17g" % v for v in props[p, :nprops[int(m)]])))


@pytest.mark.skipif(ft.device_count() > 0, reason="checks the behaviour without a GPU")
@pytest.mark.parametrize("mode", [[], ["--resident"]])
def test_driver_without_gpu_fails_loudly(tmp_path, mode):
    write_case(str(tmp_path), golden("run_ex9"))
    res = subprocess.run([DRIVER, "mesh.inp", "--max-steps", "3"] + mode, cwd=str(tmp_path),
                         capture_output=True, text=True, env=dict(os.environ, FEMTECH_UID="t"), timeout=60)
    assert res.returncode == ft.FT_ERR_HIP, (res.returncode, res.stdout, res.stderr)
    assert "no HIP device" in res.stderr
    assert not os.path.exists(tmp_path / "tension.u.f64")


def test_input_errors_terminate_with_the_reference_codes(tmp_path):
    """ReadInputFile / ReadMaterials failures end the process with FemTech's
    code 3 (ReadInputFile.cpp:31-76, ReadMaterials.cpp:11-13)."""
    env = dict(os.environ, FEMTECH_UID="t")
    res = subprocess.run([DRIVER, "mesh.unknown"], cwd=str(tmp_path), capture_output=True, text=True, env=env,
                         timeout=60)
    assert res.returncode == 3 and "Input file type is unknown" in res.stderr
    write_case(str(tmp_path), golden("run_ex9"))
    os.remove(tmp_path / "materials.dat")
    res = subprocess.run([DRIVER, "mesh.inp"], cwd=str(tmp_path), capture_output=True, text=True, env=env,
                         timeout=60)
    assert res.returncode == 3, (res.returncode, res.stderr)
