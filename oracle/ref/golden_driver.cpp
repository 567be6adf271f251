/*
 * golden_driver.cpp -- TEST INFRASTRUCTURE ONLY.
 *
 * A driver of our own, linked against FemTech's library translation units
 * compiled unmodified from the reference source tree (oracle/ref/Makefile).
 * It writes golden vectors that pin the oracle.  It never ships and runs only
 * where the reference source tree is present.
 *
 * Modes
 *   run <dir>     ex9-style explicit run (ex9.cpp:19-175) on <dir>/mesh.inp and
 *                 <dir>/materials.dat with boundary motions given as data
 *                 (<dir>/bc.bin, <dir>/params.txt, optional v0.bin / u0.bin).
 *   stress <dir>  StressUpdate (StressUpdate.cpp:5-29) on given F sequences.
 *   geom <dir>    volumeHexahedron / areaHexahedronFace /
 *                 CalculateCharacteristicLength_C3D8 on given element corners.
 *   ex7 <dir>     ex7's run (ex7.cpp:19-170): its own loop and boundary routine.
 *   time <dir>    the run's loop timed alone, no output, on 1..N MPI ranks
 *                 (CPU-baseline calibration; <dir>/coords.bin added).
 *
 * Every FemTech routine it calls is the reference's own, compiled unmodified
 * (InitFemTechWoInput, ReadInputFile, ReadMaterials, PartitionMesh with the
 * vendored ParMETIS, AllocateArrays, ShapeFunctions, ..., FinalizeFemTech);
 * the driver defines only what ex9.cpp defines itself (ex9.cpp:9-17) and
 * its own boundary-condition and output code.
 */
#include "FemTech.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include <array>
#include <map>
#include <string>
#include <vector>

/* Driver-owned globals (ex9.cpp:9-17). */
double Time, dt;
double ExplicitTimeStepReduction = 0.8;
double FailureTimeStep = 1e-11;
bool ImplicitStatic = false;
bool ImplicitDynamic = false;
bool ExplicitDynamic = true;

namespace {

template <typename T>
std::vector<T> read_bin(const std::string &path, bool required = true) {
  std::vector<T> v;
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) {
    if (required) {
      fprintf(stderr, "cannot open %s\n", path.c_str());
      exit(2);
    }
    return v;
  }
  fseek(f, 0, SEEK_END);
  long n = ftell(f);
  fseek(f, 0, SEEK_SET);
  v.resize(n / sizeof(T));
  if (fread(v.data(), sizeof(T), v.size(), f) != v.size()) exit(2);
  fclose(f);
  return v;
}

template <typename T>
void write_bin(const std::string &path, const T *p, size_t n) {
  FILE *f = fopen(path.c_str(), "wb");
  fwrite(p, sizeof(T), n, f);
  fclose(f);
}

/* Boundary motions as data: same table layout as the oracle. */
struct Motions {
  std::vector<int> dof_motion;
  std::vector<double> table;
};

void motion_eval(const double *mo, const double *X, double t, double *u, double *v,
                 double *a) {
  int type = (int)mo[0];
  for (int d = 0; d < 3; ++d) u[d] = v[d] = a[d] = 0.0;
  if (type == 1) {
    for (int d = 0; d < 3; ++d) {
      u[d] = t * mo[1];
      v[d] = mo[1];
    }
  } else if (type == 2) {
    double amp = mo[1], T = mo[2];
    double s = t < T ? t / T : 1.0;
    for (int d = 0; d < 3; ++d) {
      u[d] = amp * s;
      v[d] = t < T ? amp / T : 0.0;
    }
  } else if (type != 0) {
    fprintf(stderr, "motion type %d not supported by the golden driver\n", type);
    exit(2);
  }
}

void apply_bc(const Motions &m) {
  for (int n = 0; n < nNodes; ++n)
    for (int d = 0; d < 3; ++d) {
      int k = m.dof_motion[3 * n + d];
      if (k < 0) continue;
      double u[3], v[3], a[3];
      motion_eval(&m.table[16 * k], &coordinates[3 * n], Time, u, v, a);
      boundary[3 * n + d] = 1;
      displacements[3 * n + d] = u[d];
      velocities[3 * n + d] = v[d];
      accelerations[3 * n + d] = a[d];
    }
}

// ex9.cpp:19-27 then the explicit branch, in the current directory.
int mode_run() {
  ReadInputFile(access("mesh.k", F_OK) == 0 ? "mesh.k" : "mesh.inp");
  ReadMaterials();
  PartitionMesh();
  AllocateArrays();
  Time = 0.0;

  Motions mot;
  {
    std::vector<int> bc = read_bin<int>("bc.bin");
    int nd = 3 * nNodes;
    mot.dof_motion.assign(bc.begin(), bc.begin() + nd);
    int nm = bc[nd];
    std::vector<double> tab = read_bin<double>("motions.bin");
    if ((int)tab.size() != 16 * nm) return 3;
    mot.table = tab;
  }
  int max_steps = 0, nsnap = 0;
  double tmax = 0, factor = 0;
  std::vector<int> snaps;
  {
    FILE *f = fopen("params.txt", "r");
    if (fscanf(f, "%d %lf %lf %lf %d", &max_steps, &tmax, &factor, &FailureTimeStep,
               &nsnap) != 5)
      return 4;
    snaps.resize(nsnap);
    for (int i = 0; i < nsnap; ++i)
      if (fscanf(f, "%d", &snaps[i]) != 1) return 4;
    fclose(f);
  }
  ExplicitTimeStepReduction = factor;
  std::vector<double> u0 = read_bin<double>("u0.bin", false);
  std::vector<double> v0 = read_bin<double>("v0.bin", false);
  if (!u0.empty()) memcpy(displacements, u0.data(), nDOF * sizeof(double));
  if (!v0.empty()) memcpy(velocities, v0.data(), nDOF * sizeof(double));

  write_bin("out_coords.f64", coordinates, nDOF);
  write_bin("out_conn.i32", connectivity, eptr[nelements]);

  ShapeFunctions();
  write_bin("out_dshp.f64", dshp, 192 * (size_t)nelements);
  write_bin("out_detJ.f64", detJacobian, 8 * (size_t)nelements);
  AssembleLumpedMass();
  write_bin("out_mass.f64", mass, nDOF);

  apply_bc(mot);
  dt = ExplicitTimeStepReduction * StableTimeStep();
  GetForce();
  CalculateAccelerations();
  write_bin("out_fi0.f64", fi, nDOF);
  write_bin("out_a0.f64", accelerations, nDOF);

  std::vector<double> hist_time(1, Time), hist_dt(1, dt), hist_energy(4, 0.0);
  std::vector<double> snap_u;
  auto snap = [&](int step) {
    for (int s = 0; s < nsnap; ++s)
      if (snaps[s] == step) snap_u.insert(snap_u.end(), displacements, displacements + nDOF);
  };
  snap(0);
  double Wint = 0.0, Wext = 0.0;
  int step = 0;
  // ex9.cpp:107-175
  while (Time < tmax && step < max_steps) {
    double t_n = Time;
    double t_np1 = Time + dt;
    Time = t_np1;
    double dt_nphalf = dt;
    double t_nphalf = 0.5 * (t_np1 + t_n);
    for (int i = 0; i < nDOF; i++) {
      if (boundary[i])
        velocities_half[i] = velocities[i];
      else
        velocities_half[i] = velocities[i] + (t_nphalf - t_n) * accelerations[i];
    }
    memcpy(displacements_prev, displacements, nDOF * sizeof(double));
    memcpy(accelerations_prev, accelerations, nDOF * sizeof(double));
    memcpy(fi_prev, fi, nDOF * sizeof(double));
    memcpy(fe_prev, fe, nDOF * sizeof(double));
    for (int i = 0; i < nDOF; i++)
      if (!boundary[i]) displacements[i] = displacements[i] + dt_nphalf * velocities_half[i];
    apply_bc(mot);
    GetForce();
    CalculateAccelerations();
    for (int i = 0; i < nDOF; i++)
      if (!boundary[i])
        velocities[i] = velocities_half[i] + (t_np1 - t_nphalf) * accelerations[i];
    CheckEnergy(Time, 0);  // writes its own rows (%12.6e) to energy_<uid>.dat
    // Full-precision copy of the same sums (CheckEnergy.cpp:15-66, one rank),
    // computed from the library's arrays.
    double sWint = 0.0, sWext = 0.0, WKE = 0.0;
    for (int i = 0; i < nDOF; ++i) {
      double dd = displacements[i] - displacements_prev[i];
      WKE += mass[i] * velocities[i] * velocities[i];
      if (boundary[i]) {
        double reaction = fi_prev[i] + fi[i] + mass[i] * (accelerations[i] + accelerations_prev[i]);
        sWext += dd * reaction;
      }
      sWint += dd * (fi_prev[i] + fi[i]);
      sWext += dd * (fe_prev[i] + fe[i]);
    }
    WKE *= 0.5;
    sWint *= 0.5;
    sWext *= 0.5;
    Wint += sWint;
    Wext += sWext;
    ++step;
    dt = ExplicitTimeStepReduction * StableTimeStep();
    hist_time.push_back(Time);
    hist_dt.push_back(dt);
    hist_energy.push_back(Wint);
    hist_energy.push_back(Wext);
    hist_energy.push_back(WKE);
    hist_energy.push_back(fabs(WKE + Wint - Wext));
    snap(step);
  }
  // CalculateStrain (CalculateStrain.cpp:77-97) and the principal-strain
  // summary (CalculateStrain.cpp:8-75) at the final state (F from the last
  // GetForce).
  CalculateStrain();
  write_bin("out_eavg.f64", Eavg, 9 * (size_t)nelements);
  std::vector<double> princ;
  for (int e = 0; e < nelements; ++e) {
    double mx, mn, sh;
    CalculateMaximumPrincipalStrain(e, &mx, &mn, &sh);
    princ.push_back(mx);
    princ.push_back(mn);
    princ.push_back(sh);
  }
  write_bin("out_princ.f64", princ.data(), princ.size());
  write_bin("out_u.f64", displacements, nDOF);
  write_bin("out_v.f64", velocities, nDOF);
  write_bin("out_a.f64", accelerations, nDOF);
  write_bin("out_fi.f64", fi, nDOF);
  write_bin("out_time.f64", hist_time.data(), hist_time.size());
  write_bin("out_dt.f64", hist_dt.data(), hist_dt.size());
  write_bin("out_energy.f64", hist_energy.data(), hist_energy.size());
  write_bin("out_snap_u.f64", snap_u.data(), snap_u.size());
  printf("steps %d time %.17g uid %s\n", step, Time, uid.c_str());
  return 0;
}

/* ex7's run (ex7.cpp:19-125) on <dir>/mesh.k and <dir>/materials.dat: the
 * initial force and accelerations before any boundary condition, the
 * half-step velocity and the displacement update on every DOF, only
 * displacements_prev saved each step (fi_prev, fe_prev and
 * accelerations_prev keep their zero allocation), and its boundary routine
 * (ex7.cpp:127-170), which sets displacements and flags only.  Constants as
 * in ex7.cpp: tMax 1.0 s, dMax 1 mm, loaded face y = 5 mm, 50 plot steps.
 * Outputs: final u/v/a/fi, time and dt per step, the energies per step
 * (full-precision copy of CheckEnergy's sums), and the corner node
 * (5, 5, 5) mm at every plot step as ex7's CustomPlot records it
 * (ex7.cpp:172-205) in full precision. */
const double kEx7Tol = 1e-5, kEx7L = 0.005, kEx7TMax = 1.0, kEx7DMax = 0.001;
const int kEx7Plots = 50;

void ex7_bc() {
  const double applied = Time * (kEx7DMax / kEx7TMax);
  for (int i = 0; i < nNodes; i++) {
    for (int d = 0; d < 3; ++d)
      if (fabs(coordinates[ndim * i + d] - 0.0) < kEx7Tol) {
        boundary[ndim * i + d] = 1;
        displacements[ndim * i + d] = 0.0;
      }
    if (fabs(coordinates[ndim * i + 1] - kEx7L) < kEx7Tol) {
      boundary[ndim * i + 1] = 1;
      displacements[ndim * i + 1] = applied;
    }
  }
}

int mode_ex7() {
  ReadInputFile(access("mesh.k", F_OK) == 0 ? "mesh.k" : "mesh.inp");
  ReadMaterials();
  PartitionMesh();
  AllocateArrays();
  Time = 0.0;
  write_bin("out_coords.f64", coordinates, nDOF);
  write_bin("out_conn.i32", connectivity, eptr[nelements]);
  int corner = -1;
  for (int i = 0; i < nNodes; i++)
    if (fabs(coordinates[ndim * i] - kEx7L) < kEx7Tol && fabs(coordinates[ndim * i + 1] - kEx7L) < kEx7Tol &&
        fabs(coordinates[ndim * i + 2] - kEx7L) < kEx7Tol)
      corner = i;
  if (corner < 0) return 6;
  std::vector<double> plot = {0.0, 0.0, 0.0, 0.0};
  ShapeFunctions();
  AssembleLumpedMass();
  dt = ExplicitTimeStepReduction * StableTimeStep();
  GetForce();
  CalculateAccelerations();
  const int nSteps = (int)(kEx7TMax / dt);
  const int nsteps_plot = (int)(nSteps / kEx7Plots);
  memcpy(displacements_prev, displacements, nDOF * sizeof(double));
  std::vector<double> hist_time(1, Time), hist_dt(1, dt), hist_energy(4, 0.0);
  double Wint = 0.0, Wext = 0.0;
  int counter = 1, step = 0;
  while (Time < kEx7TMax) {
    double t_n = Time;
    double t_np1 = Time + dt;
    Time = t_np1;
    double dt_nphalf = dt;
    double t_nphalf = 0.5 * (t_np1 + t_n);
    for (int i = 0; i < nDOF; i++) velocities_half[i] = velocities[i] + (t_nphalf - t_n) * accelerations[i];
    memcpy(displacements_prev, displacements, nDOF * sizeof(double));
    for (int i = 0; i < nDOF; i++) displacements[i] = displacements[i] + dt_nphalf * velocities_half[i];
    ex7_bc();
    GetForce();
    CalculateAccelerations();
    for (int i = 0; i < nDOF; i++) velocities[i] = velocities_half[i] + (t_np1 - t_nphalf) * accelerations[i];
    const int writeFlag = counter % nsteps_plot;
    CheckEnergy(Time, writeFlag);
    double sWint = 0.0, sWext = 0.0, WKE = 0.0;
    for (int i = 0; i < nDOF; ++i) {
      double dd = displacements[i] - displacements_prev[i];
      WKE += mass[i] * velocities[i] * velocities[i];
      if (boundary[i]) {
        double reaction = fi_prev[i] + fi[i] + mass[i] * (accelerations[i] + accelerations_prev[i]);
        sWext += dd * reaction;
      }
      sWint += dd * (fi_prev[i] + fi[i]);
      sWext += dd * (fe_prev[i] + fe[i]);
    }
    Wint += 0.5 * sWint;
    Wext += 0.5 * sWext;
    WKE *= 0.5;
    if (writeFlag == 0) {
      plot.push_back(Time);
      for (int d = 0; d < 3; ++d) plot.push_back(displacements[3 * corner + d]);
    }
    ++counter;
    ++step;
    dt = ExplicitTimeStepReduction * StableTimeStep();
    hist_time.push_back(Time);
    hist_dt.push_back(dt);
    hist_energy.push_back(Wint);
    hist_energy.push_back(Wext);
    hist_energy.push_back(WKE);
    hist_energy.push_back(fabs(WKE + Wint - Wext));
  }
  write_bin("out_u.f64", displacements, nDOF);
  write_bin("out_v.f64", velocities, nDOF);
  write_bin("out_a.f64", accelerations, nDOF);
  write_bin("out_fi.f64", fi, nDOF);
  write_bin("out_time.f64", hist_time.data(), hist_time.size());
  write_bin("out_dt.f64", hist_dt.data(), hist_dt.size());
  write_bin("out_energy.f64", hist_energy.data(), hist_energy.size());
  write_bin("out_plot.f64", plot.data(), plot.size());
  printf("steps %d time %.17g uid %s corner %d plot_every %d\n", step, Time, uid.c_str(), corner, nsteps_plot);
  return 0;
}

/* Wall time of the ex9 loop alone (ex9.cpp:107-175), with no fixture output,
 * on any number of MPI ranks: the boundary motions of bc.bin are assigned to
 * the rank's nodes by their coordinates (coords.bin: the global mesh's nodes
 * in bc.bin's order, as read from the same mesh file).  Rank 0 prints the
 * step count and the seconds between barriers around the loop. */
int mode_time() {
  ReadInputFile(access("mesh.k", F_OK) == 0 ? "mesh.k" : "mesh.inp");
  ReadMaterials();
  PartitionMesh();
  AllocateArrays();
  Time = 0.0;
  std::vector<int> bc = read_bin<int>("bc.bin");
  std::vector<double> gx = read_bin<double>("coords.bin");
  const size_t ng = gx.size() / 3;
  Motions mot;
  mot.table = read_bin<double>("motions.bin");
  if ((int)mot.table.size() != 16 * bc[3 * ng]) return 3;
  std::map<std::array<double, 3>, size_t> where;
  for (size_t n = 0; n < ng; ++n) where[{gx[3 * n], gx[3 * n + 1], gx[3 * n + 2]}] = n;
  mot.dof_motion.assign(3 * (size_t)nNodes, -1);
  for (int n = 0; n < nNodes; ++n) {
    auto it = where.find({coordinates[3 * n], coordinates[3 * n + 1], coordinates[3 * n + 2]});
    if (it == where.end()) return 5;
    for (int d = 0; d < 3; ++d) mot.dof_motion[3 * n + d] = bc[3 * it->second + d];
  }
  int max_steps = 0, nsnap = 0;
  double tmax = 0, factor = 0;
  {
    FILE *f = fopen("params.txt", "r");
    if (fscanf(f, "%d %lf %lf %lf %d", &max_steps, &tmax, &factor, &FailureTimeStep,
               &nsnap) != 5)
      return 4;
    fclose(f);
  }
  ExplicitTimeStepReduction = factor;
  ShapeFunctions();
  AssembleLumpedMass();
  apply_bc(mot);
  dt = ExplicitTimeStepReduction * StableTimeStep();
  GetForce();
  CalculateAccelerations();
  MPI_Barrier(MPI_COMM_WORLD);
  const double t0 = MPI_Wtime();
  int step = 0;
  while (Time < tmax && step < max_steps) {
    double t_n = Time;
    double t_np1 = Time + dt;
    Time = t_np1;
    double dt_nphalf = dt;
    double t_nphalf = 0.5 * (t_np1 + t_n);
    for (int i = 0; i < nDOF; i++) {
      if (boundary[i])
        velocities_half[i] = velocities[i];
      else
        velocities_half[i] = velocities[i] + (t_nphalf - t_n) * accelerations[i];
    }
    memcpy(displacements_prev, displacements, nDOF * sizeof(double));
    memcpy(accelerations_prev, accelerations, nDOF * sizeof(double));
    memcpy(fi_prev, fi, nDOF * sizeof(double));
    memcpy(fe_prev, fe, nDOF * sizeof(double));
    for (int i = 0; i < nDOF; i++)
      if (!boundary[i]) displacements[i] = displacements[i] + dt_nphalf * velocities_half[i];
    apply_bc(mot);
    GetForce();
    CalculateAccelerations();
    for (int i = 0; i < nDOF; i++)
      if (!boundary[i])
        velocities[i] = velocities_half[i] + (t_np1 - t_nphalf) * accelerations[i];
    CheckEnergy(Time, 0);
    ++step;
    dt = ExplicitTimeStepReduction * StableTimeStep();
  }
  MPI_Barrier(MPI_COMM_WORLD);
  const double t1 = MPI_Wtime();
  if (world_rank == 0)
    printf("steps %d time %.17g loop_seconds %.6f ranks %d elements_rank0 %d\n", step, Time,
           t1 - t0, world_size, nelements);
  return 0;
}

/* Sets up n one-Gauss-point "elements" so the reference's material routines
 * can be called on given deformation gradients. */
int mode_stress() {
  std::vector<double> in = read_bin<double>("stress_in.f64");
  // record: mat, props[9], dt, T, F[T][9]  (T fixed for the file: first record)
  size_t pos = 0;
  std::vector<double> out_pk2, out_hist;
  ndim = 3;
  nelements = 1;
  nPIDglobal = 1;
  fptr = (int *)calloc(2, sizeof(int));
  detFptr = (int *)calloc(2, sizeof(int));
  gpPtr = (int *)calloc(2, sizeof(int));
  pk2ptr = (int *)calloc(2, sizeof(int));
  pid = (int *)calloc(1, sizeof(int));
  materialID = (int *)calloc(1, sizeof(int));
  properties = (double *)calloc(9, sizeof(double));
  F = (double *)calloc(9, sizeof(double));
  detF = (double *)calloc(1, sizeof(double));
  pk2 = (double *)calloc(6, sizeof(double));
  Hn_1 = (double *)calloc(9, sizeof(double));
  Hn_2 = (double *)calloc(9, sizeof(double));
  S0n = (double *)calloc(9, sizeof(double));
  mat1 = (double *)calloc(9, sizeof(double));
  mat2 = (double *)calloc(9, sizeof(double));
  mat3 = (double *)calloc(9, sizeof(double));
  mat4 = (double *)calloc(9, sizeof(double));
  while (pos < in.size()) {
    int mat = (int)in[pos++];
    materialID[0] = mat;
    for (int k = 0; k < 9; ++k) properties[k] = in[pos++];
    dt = in[pos++];
    int T = (int)in[pos++];
    memset(Hn_1, 0, 9 * sizeof(double));
    memset(Hn_2, 0, 9 * sizeof(double));
    memset(S0n, 0, 9 * sizeof(double));
    for (int t = 0; t < T; ++t) {
      for (int k = 0; k < 9; ++k) F[k] = in[pos++];
      memset(pk2, 0, 6 * sizeof(double));
      DeterminateF(0, 0);
      StressUpdate(0, 0);
      out_pk2.insert(out_pk2.end(), pk2, pk2 + 6);
      out_pk2.push_back(detF[0]);
    }
    out_hist.insert(out_hist.end(), Hn_1, Hn_1 + 9);
    out_hist.insert(out_hist.end(), Hn_2, Hn_2 + 9);
    out_hist.insert(out_hist.end(), S0n, S0n + 9);
  }
  write_bin("stress_out.f64", out_pk2.data(), out_pk2.size());
  write_bin("stress_hist.f64", out_hist.data(), out_hist.size());
  return 0;
}

int mode_geom() {
  std::vector<double> in = read_bin<double>("geom_in.f64");
  int n = (int)(in.size() / 24);
  ndim = 3;
  nelements = n;
  nNodes = 8 * n;
  coordinates = in.data();
  displacements = (double *)calloc(24 * (size_t)n, sizeof(double));
  connectivity = (int *)malloc(8 * (size_t)n * sizeof(int));
  eptr = (int *)malloc((n + 1) * sizeof(int));
  for (int i = 0; i < 8 * n; ++i) connectivity[i] = i;
  for (int e = 0; e <= n; ++e) eptr[e] = 8 * e;
  static const int idx[24] = {0, 1, 2, 3, 4, 5, 6, 7, 0, 3, 7, 4,
                              1, 2, 6, 5, 0, 1, 5, 4, 3, 2, 6, 7};
  std::vector<double> out;
  for (int e = 0; e < n; ++e) {
    double *xe = &in[24 * (size_t)e];
    out.push_back(volumeHexahedron(xe));
    for (int f = 0; f < 6; ++f) out.push_back(areaHexahedronFace(xe, &idx[4 * f]));
    out.push_back(CalculateCharacteristicLength_C3D8(e));
  }
  write_bin("geom_out.f64", out.data(), out.size());
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: %s run|ex7|time|stress|geom <dir>\n", argv[0]);
    return 2;
  }
  const std::string mode = argv[1];
  // The log and energy files are written to the working directory.
  if (chdir(argv[2]) != 0) return 2;
  InitFemTechWoInput(argc, argv);
  int rc = 2;
  if (mode == "run" || mode == "time" || mode == "ex7") {
    rc = mode == "run" ? mode_run() : mode == "time" ? mode_time() : mode_ex7();
    FinalizeFemTech();  // frees the arrays and closes energy_<uid>.dat
    return rc;
  }
  if (mode == "stress") rc = mode_stress();
  if (mode == "geom") rc = mode_geom();
  // The unit modes set up no mesh and no energy file (FinalizeFemTech
  // would close it): the log and MPI are closed directly.
  finaliseLog();
  MPI_Finalize();
  return rc;
}
