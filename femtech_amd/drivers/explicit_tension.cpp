// Explicit uniaxial-tension driver over FemTech's C++ call surface.
//
// The scenario of FemTech's examples/ex9 and Benchmarking-Parallel drivers: a
// C3D8 block with symmetry planes x = 0, y = 0, z = 0 and the face y = L
// pulled in y with u = Time * dMax / tMax; central-difference steps with
// dt = ExplicitTimeStepReduction * StableTimeStep() recomputed every step and
// the energy balance checked every step.  The call sequence is that of
// ex9.cpp:19-175, so this file exercises the same library surface on the GPU
// box, where the reference's own drivers are not available.
//
//   explicit_tension MESH [--tmax T] [--dmax D] [--length L] [--plots N]
//                         [--max-steps K] [--resident] [--loop ex9|ex7]
//                         [--out PREFIX]
//
// Default mode (compatibility): host arrays are authoritative; boundary
// conditions and the velocity/displacement updates run here, every library
// call goes to the GPU.  --resident: the same problem with the boundary
// conditions handed over as data and the loop run on the device through
// ft_run_init / ft_run_steps; its plot.dat rows come from asynchronous
// displacement snapshots (no VTU files).
//
// --loop ex7 (compatibility mode): the loop of ex7.cpp:56-125 instead of
// ex9's: the initial force and accelerations come before any boundary
// condition, the half-step velocity and the displacement update cover every
// DOF, only displacements_prev is saved each step (fi_prev, fe_prev and
// accelerations_prev keep their zero allocation, so CheckEnergy sums with
// them as ex7's run does), and the boundary routine (ex7.cpp:127-170) sets
// displacements and flags only.  PREFIX.corner.f64 then holds, as ex7's
// CustomPlot records (ex7.cpp:172-205), time and displacement of the node at
// (L, L, L) at step 0 and every plot step.
//
// Output: PREFIX.u.f64 (final displacement, 3 per node, FP64),
// PREFIX.summary.txt (steps, time, dt, energies), the energy file
// energy_<uid>.dat and, in compatibility mode, VTU/PVD files and plot.dat
// rows every nSteps / plots steps.
#include "FemTech.h"
#include "femtech_amd.h"

#include <algorithm>
#include <string>
#include <vector>

double Time, dt;
double ExplicitTimeStepReduction = 0.8;
double FailureTimeStep = 1e-11;
bool ImplicitStatic = false;
bool ImplicitDynamic = false;
bool ExplicitDynamic = true;

namespace {

struct Options {
  const char *mesh = nullptr;
  double tmax = 1.0, dmax = 0.007, length = 0.005;
  int plots = 50;
  long max_steps = -1;
  bool resident = false;
  bool ex7 = false;
  std::string out = "tension";
};

const double kTol = 1e-5;  // plane detection tolerance (ex9.cpp:188)

// Symmetry planes fixed, the loaded face moving at dMax / tMax.
void apply_tension_bc(const Options &o) {
  const double applied = Time * (o.dmax / o.tmax);
  for (int n = 0; n < nNodes; ++n) {
    for (int d = 0; d < 3; ++d) {
      const int i = 3 * n + d;
      if (fabs(coordinates[i]) < kTol) {
        boundary[i] = 1;
        displacements[i] = velocities[i] = accelerations[i] = 0.0;
      }
    }
    const int iy = 3 * n + 1;
    if (fabs(coordinates[iy] - o.length) < kTol) {
      boundary[iy] = 1;
      displacements[iy] = applied;
      velocities[iy] = o.dmax / o.tmax;
      accelerations[iy] = 0.0;
    }
  }
}

// ex7's boundary routine: planes fixed and the loaded face displaced;
// velocities and accelerations are left as they are.
void apply_ex7_bc(const Options &o) {
  const double applied = Time * (o.dmax / o.tmax);
  for (int n = 0; n < nNodes; ++n) {
    for (int d = 0; d < 3; ++d)
      if (fabs(coordinates[3 * n + d]) < kTol) {
        boundary[3 * n + d] = 1;
        displacements[3 * n + d] = 0.0;
      }
    if (fabs(coordinates[3 * n + 1] - o.length) < kTol) {
      boundary[3 * n + 1] = 1;
      displacements[3 * n + 1] = applied;
    }
  }
}

int corner_node(const Options &o) {
  for (int n = 0; n < nNodes; ++n)
    if (fabs(coordinates[3 * n] - o.length) < kTol && fabs(coordinates[3 * n + 1] - o.length) < kTol &&
        fabs(coordinates[3 * n + 2] - o.length) < kTol)
      return n;
  return -1;
}

// plot.dat as the examples' CustomPlot writes it (ex9.cpp:250-279,
// ex7.cpp:172-205): a header and zeros at Time = 0, then Time and the
// displacement of the node at (L, L, L) at each plot step; ex9 prints
// %11.5e, ex7 %11.3e.
void custom_plot(const Options &o, bool first) {
  const char *fmt = o.ex7 ? "%11.3e %11.3e  %11.3e  %11.3e\n" : "%11.5e %11.5e  %11.5e  %11.5e\n";
  FILE *f = fopen("plot.dat", first ? "w" : "a");
  if (!f) return;
  if (first) {
    fprintf(f, "# Results for Node ?\n");
    fprintf(f, "# Time  DispX    DispY   DispZ\n");
    fprintf(f, fmt, 0.0, 0.0, 0.0, 0.0);
  } else {
    const int n = corner_node(o);
    if (n >= 0) fprintf(f, fmt, Time, displacements[3 * n], displacements[3 * n + 1], displacements[3 * n + 2]);
  }
  fclose(f);
}

void write_outputs(const Options &o, long steps, double energy[4]) {
  FILE *f = fopen((o.out + ".u.f64").c_str(), "wb");
  if (f) {
    fwrite(displacements, sizeof(double), nDOF, f);
    fclose(f);
  }
  f = fopen((o.out + ".summary.txt").c_str(), "w");
  if (f) {
    fprintf(f, "steps %ld\ntime %.17g\ndt %.17g\nW_int %.17g\nW_ext %.17g\nKE %.17g\nbalance %.17g\n", steps,
            Time, dt, energy[0], energy[1], energy[2], energy[3]);
    fclose(f);
  }
}

long run_ex7(const Options &o, const std::string &base) {
  const int corner = corner_node(o);
  std::vector<double> plot = {0.0, 0.0, 0.0, 0.0};
  dt = ExplicitTimeStepReduction * StableTimeStep();
  GetForce();
  CalculateAccelerations();
  const long n_steps = (long)(o.tmax / dt);
  const long plot_every = n_steps / o.plots > 0 ? n_steps / o.plots : 1;
  FILE_LOG_MASTER(INFO, "initial dt = %3.3e, nSteps = %ld, plot every %ld", dt, n_steps, plot_every);
  memcpy(displacements_prev, displacements, nDOF * sizeof(double));
  int nplot = 0;
  long step = 1, taken = 0;
  while (Time < o.tmax && (o.max_steps < 0 || taken < o.max_steps)) {
    const double t_n = Time, t_np1 = Time + dt, t_half = 0.5 * (t_n + t_np1);
    Time = t_np1;
    for (int i = 0; i < nDOF; ++i) velocities_half[i] = velocities[i] + (t_half - t_n) * accelerations[i];
    memcpy(displacements_prev, displacements, nDOF * sizeof(double));
    for (int i = 0; i < nDOF; ++i) displacements[i] += dt * velocities_half[i];
    apply_ex7_bc(o);
    GetForce();
    CalculateAccelerations();
    for (int i = 0; i < nDOF; ++i) velocities[i] = velocities_half[i] + (t_np1 - t_half) * accelerations[i];
    const int write_flag = (int)(step % plot_every);
    CheckEnergy(Time, write_flag);
    if (write_flag == 0) {
      ++nplot;
      WriteVTU(base.c_str(), nplot);
      if (nplot < MAXPLOTSTEPS) {
        stepTime[nplot] = Time;
        WritePVD(base.c_str(), nplot);
      }
      if (corner >= 0) {
        plot.push_back(Time);
        plot.insert(plot.end(), displacements + 3 * corner, displacements + 3 * corner + 3);
      }
      custom_plot(o, false);
    }
    ++step;
    ++taken;
    dt = ExplicitTimeStepReduction * StableTimeStep();
    MPI_Barrier(MPI_COMM_WORLD);
  }
  FILE *f = fopen((o.out + ".corner.f64").c_str(), "wb");
  if (f) {
    fwrite(plot.data(), sizeof(double), plot.size(), f);
    fclose(f);
  }
  return taken;
}

long run_compatibility(const Options &o, const std::string &base, double energy[4]) {
  apply_tension_bc(o);
  dt = ExplicitTimeStepReduction * StableTimeStep();
  GetForce();
  CalculateAccelerations();
  const long n_steps = (long)(o.tmax / dt);
  const long plot_every = n_steps / o.plots > 0 ? n_steps / o.plots : 1;
  FILE_LOG_MASTER(INFO, "initial dt = %3.3e, nSteps = %ld, plot every %ld", dt, n_steps, plot_every);
  int plot = 0;
  long step = 1, taken = 0;
  while (Time < o.tmax && (o.max_steps < 0 || taken < o.max_steps)) {
    const double t_n = Time, t_np1 = Time + dt, t_half = 0.5 * (t_n + t_np1);
    Time = t_np1;
    for (int i = 0; i < nDOF; ++i)
      velocities_half[i] = boundary[i] ? velocities[i] : velocities[i] + (t_half - t_n) * accelerations[i];
    memcpy(displacements_prev, displacements, nDOF * sizeof(double));
    memcpy(accelerations_prev, accelerations, nDOF * sizeof(double));
    memcpy(fi_prev, fi, nDOF * sizeof(double));
    memcpy(fe_prev, fe, nDOF * sizeof(double));
    for (int i = 0; i < nDOF; ++i)
      if (!boundary[i]) displacements[i] += dt * velocities_half[i];
    apply_tension_bc(o);
    GetForce();
    CalculateAccelerations();
    for (int i = 0; i < nDOF; ++i)
      if (!boundary[i]) velocities[i] = velocities_half[i] + (t_np1 - t_half) * accelerations[i];
    const int write_flag = (int)(step % plot_every);
    CheckEnergy(Time, write_flag);
    if (write_flag == 0) {
      ++plot;
      CalculateStrain();
      WriteVTU(base.c_str(), plot);
      if (plot < MAXPLOTSTEPS) {
        stepTime[plot] = Time;
        WritePVD(base.c_str(), plot);
      }
      custom_plot(o, false);
    }
    ++step;
    ++taken;
    dt = ExplicitTimeStepReduction * StableTimeStep();
    MPI_Barrier(MPI_COMM_WORLD);
  }
  custom_plot(o, false);  // the last step (ex9.cpp:179)
  (void)energy;
  return taken;
}

// The same problem on the device-resident loop: boundary conditions as data
// (motion 0 fixed, motion 1 linear in time).
long run_resident(const Options &o, double energy[4]) {
  std::vector<int> dof_motion(nDOF, -1);
  for (int n = 0; n < nNodes; ++n) {
    for (int d = 0; d < 3; ++d)
      if (fabs(coordinates[3 * n + d]) < kTol) dof_motion[3 * n + d] = 0;
    if (fabs(coordinates[3 * n + 1] - o.length) < kTol) dof_motion[3 * n + 1] = 1;
  }
  double motions[2 * FT_MOTION_STRIDE] = {0};
  motions[0] = FT_MOTION_FIXED;
  motions[FT_MOTION_STRIDE] = FT_MOTION_LINEAR;
  motions[FT_MOTION_STRIDE + 1] = o.dmax / o.tmax;
  ft_engine *eng = nullptr;
  int rc = ft_engine_create(nNodes, nelements, coordinates, connectivity, pid, nPIDglobal, materialID, properties,
                            getenv("FEMTECH_PRECISION") && !strcmp(getenv("FEMTECH_PRECISION"), "fp32")
                                ? FT_FP32
                                : FT_FP64,
                            &eng);
  int done = 0, k = 0;
  long taken = 0;
  if (!rc) rc = ft_set_boundary(eng, dof_motion.data(), 2, motions);
  if (!rc) rc = ft_set_state(eng, nullptr, nullptr);
  if (!rc) rc = ft_assemble_lumped_mass(eng, nullptr);
  if (!rc) rc = ft_run_init(eng, o.tmax, ExplicitTimeStepReduction, FailureTimeStep);
  // plot.dat rows every nSteps / plots steps, as the compatibility loop
  // writes them: the displacement at each chunk boundary reaches the host
  // while the next chunk computes (ft_snapshot_begin / ft_snapshot_end)
  double t0 = 0.0, dt0 = 0.0;
  if (!rc) rc = ft_get_time(eng, &t0, &dt0, &k);
  const long n_steps = dt0 > 0.0 ? (long)(o.tmax / dt0) : 0;
  const long chunk = o.plots <= 0 ? 1L << 20 : n_steps / o.plots > 0 ? n_steps / o.plots : 1;
  if (!rc) custom_plot(o, true);
  bool pending = false;
  while (!rc) {
    const long left = o.max_steps < 0 ? chunk : std::min(chunk, o.max_steps - taken);
    if (left <= 0) break;
    rc = ft_run_steps(eng, (int)left, &done);
    taken += done;
    if (!rc && pending) {  // the previous boundary's rows, copied meanwhile
      rc = ft_snapshot_end(eng, displacements, &Time, &k);
      if (!rc) custom_plot(o, false);
      pending = false;
    }
    if (rc || done == 0) break;
    if (done == left && (rc = ft_snapshot_begin(eng)) == 0) pending = true;
  }
  if (!rc && pending) {
    rc = ft_snapshot_end(eng, displacements, &Time, &k);
    if (!rc) custom_plot(o, false);
  }
  if (!rc) rc = ft_get_state(eng, displacements, velocities, accelerations, fi);
  if (!rc) rc = ft_get_time(eng, &Time, &dt, &k);
  if (!rc) rc = ft_get_energy(eng, energy);
  if (!rc) custom_plot(o, false);  // the last step (ex9.cpp:179)
  if (rc) {
    FILE_LOG_SINGLE(ERROR, "resident run: %s", ft_last_error());
    ft_engine_destroy(eng);
    TerminateFemTech(rc);
  }
  ft_engine_destroy(eng);
  return taken;
}

}  // namespace

int main(int argc, char **argv) {
  Options o;
  for (int a = 1; a < argc; ++a) {
    const std::string s = argv[a];
    auto next = [&](void) -> const char * {
      if (a + 1 >= argc) {
        fprintf(stderr, "missing value for %s\n", s.c_str());
        exit(2);
      }
      return argv[++a];
    };
    if (s == "--tmax") o.tmax = atof(next());
    else if (s == "--dmax") o.dmax = atof(next());
    else if (s == "--length") o.length = atof(next());
    else if (s == "--plots") o.plots = atoi(next());
    else if (s == "--max-steps") o.max_steps = atol(next());
    else if (s == "--resident") o.resident = true;
    else if (s == "--loop") {
      const std::string l = next();
      if (l != "ex7" && l != "ex9") {
        fprintf(stderr, "--loop takes ex9 or ex7\n");
        return 2;
      }
      o.ex7 = l == "ex7";
    }
    else if (s == "--out") o.out = next();
    else if (!o.mesh) o.mesh = argv[a];
    else {
      fprintf(stderr, "unexpected argument %s\n", s.c_str());
      return 2;
    }
  }
  if (!o.mesh) {
    fprintf(stderr, "usage: %s MESH [--tmax T] [--dmax D] [--length L] [--plots N] "
                    "[--max-steps K] [--resident] [--loop ex9|ex7] [--out PREFIX]\n", argv[0]);
    return 2;
  }
  if (o.ex7 && o.resident) {
    fprintf(stderr, "--loop ex7 runs in compatibility mode only\n");
    return 2;
  }
  InitFemTechWoInput(argc, argv);
  ReadInputFile(o.mesh);
  ReadMaterials();
  PartitionMesh();
  AllocateArrays();
  std::string base(o.mesh);
  const size_t slash = base.find_last_of('/');
  if (slash != std::string::npos) base = base.substr(slash + 1);
  base = base.substr(0, base.find_last_of('.'));
  Time = 0.0;
  double energy[4] = {0, 0, 0, 0};
  long steps;
  if (o.resident) {
    steps = run_resident(o, energy);
  } else {
    WriteVTU(base.c_str(), 0);
    stepTime[0] = Time;
    custom_plot(o, true);
    ShapeFunctions();
    AssembleLumpedMass();
    steps = o.ex7 ? run_ex7(o, base) : run_compatibility(o, base, energy);
  }
  FILE_LOG_MASTER(INFO, "End of Iterative Loop: %ld steps, time %.6e", steps, Time);
  write_outputs(o, steps, energy);
  FinalizeFemTech();
  return 0;
}
