// FemTech's C++ call surface (include/FemTech.h) over the MI355X engine.
//
// Compatibility mode: the host arrays FemTech drivers read and write stay
// authoritative.  Every routine of the explicit path uploads what it reads,
// runs on the GPU through the C ABI (femtech_amd.h) and downloads what it
// writes; the host side only parses files, owns arrays, formats output and
// does FemTech's own bookkeeping (energy accumulators, logging).  There is no
// host implementation of any kernel: without a GPU, ShapeFunctions() fails
// and the process terminates with the engine's error code.
//
// Reference routines are cited per function (paths relative to the FemTech
// source tree).
#include "FemTech.h"

#include <sys/stat.h>
#include <sys/time.h>
#include <time.h>

#include <vector>

#include "femtech_amd.h"

// ---- globals owned by the library (GlobalVariables.h:18-127) ---------------
std::string uid;
int nparts = 0, nelements = 0, nallelements = 0, nNodes = 0, ndim = 0, nDOF = 0;
int nPID = 0, nPIDglobal = 0;
int world_rank = 0, world_size = 1;
double *coordinates = nullptr;
int *connectivity = nullptr;
int *globalNodeID = nullptr;
int *GaussPoints = nullptr, *gptr = nullptr, *dsptr = nullptr, *gpPtr = nullptr, *fptr = nullptr,
    *pk2ptr = nullptr, *nShapeFunctions = nullptr;
double *shp = nullptr, *dshp = nullptr;
int *eptr = nullptr, *detFptr = nullptr, *pid = nullptr, *global_eid = nullptr;
double *mass = nullptr, *stiffness = nullptr, *rhs = nullptr;
char **ElementType = nullptr;
double *C = nullptr, *gaussWeights = nullptr, *detJacobian = nullptr;
int nSpecifiedDispBC = 0;
double *displacements = nullptr;
int *boundary = nullptr;
double *velocities = nullptr, *velocities_half = nullptr, *accelerations = nullptr;
double *fe = nullptr, *fe_prev = nullptr, *fi = nullptr, *fi_prev = nullptr, *f_net = nullptr,
       *fr_prev = nullptr, *fr_curr = nullptr;
double *f_damp_prev = nullptr, *f_damp_curr = nullptr, *displacements_prev = nullptr,
       *accelerations_prev = nullptr;
double *F = nullptr, *detF = nullptr, *invF = nullptr, *pk2 = nullptr, *Eavg = nullptr;
int *materialID = nullptr;
double *properties = nullptr;
double *internals = nullptr;
int *InternalsPtr = nullptr;
double *stepTime = nullptr;
int sendProcessCount = 0;
int *sendProcessID = nullptr, *sendNeighbourCount = nullptr, *sendNeighbourCountCum = nullptr,
    *sendNodeIndex = nullptr;
double *sendNodeDisplacement = nullptr, *recvNodeDisplacement = nullptr;
FILE *energyFile = nullptr;
double *mat1 = nullptr, *mat2 = nullptr, *mat3 = nullptr, *mat4 = nullptr;
double *fintGQ = nullptr, *B = nullptr;
double *Hn_1 = nullptr, *Hn_2 = nullptr, *S0n = nullptr;

namespace {

ft_engine *g_engine = nullptr;
std::vector<double> g_principal;  // 3 per element, from the last CalculateStrain
bool g_hist_host_newer = false;   // Hn_1 / Hn_2 / S0n changed by a host-side StressUpdate
double g_wint = 0.0, g_wext = 0.0;  // CheckEnergy.cpp:4-5 accumulators
FILE *g_log = nullptr;

// FemTech codes pass through; engine-specific codes (>= 50) are kept as the
// process status.
[[noreturn]] void die(int rc, const char *where) {
  fileLogSingle(ERROR, "%s: %s", where, ft_last_error());
  TerminateFemTech(rc);
  exit(rc);  // not reached
}

void check(int rc, const char *where) {
  if (rc) die(rc, where);
}

template <typename T>
T *zalloc(size_t n, const char *name) {
  T *p = (T *)calloc(n ? n : 1, sizeof(T));
  if (!p) {
    fileLogSingle(ERROR, "Error in allocating %s array", name);
    TerminateFemTech(12);
  }
  return p;
}

template <typename T>
void release(T *&p) {
  free((void *)p);
  p = nullptr;
}

int precision_from_env() {
  const char *s = getenv("FEMTECH_PRECISION");
  if (s && (!strcmp(s, "fp32") || !strcmp(s, "FP32"))) return FT_FP32;
  return FT_FP64;
}

void vlog(enum logLevel level, const char *fmt, va_list ap) {
  static const char *const names[] = {"ERROR  ", "WARNING", "INFO   ", "DEBUG  ", "DEBUG  "};
  char stamp[32];
  struct timeval tv;
  gettimeofday(&tv, nullptr);
  struct tm r;
  time_t t = tv.tv_sec;
  strftime(stamp, sizeof(stamp), "%X", localtime_r(&t, &r));
  char msg[4096];
  vsnprintf(msg, sizeof(msg), fmt, ap);
  FILE *outs[2] = {g_log, level == ERROR ? stderr : nullptr};
  for (FILE *o : outs)
    if (o) {
      fprintf(o, "-%s.%03ld %s[%3d]: %s\n", stamp, (long)tv.tv_usec / 1000,
              names[(int)level < 5 ? (int)level : 4], world_rank, msg);
      fflush(o);
    }
}

void log_matrix(enum logLevel level, const double *mat, int n, int m, bool row_major,
                const char *fmt, va_list ap) {
  vlog(level, fmt, ap);
  if (!g_log) return;
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < m; ++j) fprintf(g_log, "%15.6E  ", row_major ? mat[j + i * m] : mat[i + j * n]);
    fprintf(g_log, "\n");
  }
}

template <typename T>
void log_array(enum logLevel level, const T *arr, int n, const char *fmt, va_list ap) {
  vlog(level, fmt, ap);
  if (!g_log) return;
  for (int i = 0; i < n; ++i) {
    if constexpr (sizeof(T) == sizeof(int))
      fprintf(g_log, "%d  ", (int)arr[i]);
    else
      fprintf(g_log, "%15.6E  ", (double)arr[i]);
  }
  fprintf(g_log, "\n");
}

void make_dir(const char *path) { (void)mkdir(path, 0755); }

}  // namespace

// ---- logging (parallel_log.cpp) ----------------------------------------------
const char *levelToString(enum logLevel level) {
  static const char *const names[] = {"ERROR", "WARNING", "INFO", "DEBUG", "DEBUG"};
  return names[(int)level < 5 ? (int)level : 4];
}
void initLog(const char *outFileName) {
  if (g_log) fclose(g_log);
  g_log = fopen(outFileName, "w");
  if (!g_log) fprintf(stdout, "ERROR(%d): Unable to open output log file\n", world_rank);
}
void finaliseLog() {
  if (g_log) fclose(g_log);
  g_log = nullptr;
}
#define FT_VLOG_BODY(call) \
  va_list ap;              \
  va_start(ap, fmt);       \
  call;                    \
  va_end(ap)
void fileLog(enum logLevel level, const char *fmt, ...) { FT_VLOG_BODY(vlog(level, fmt, ap)); }
void fileLogMaster(enum logLevel level, const char *fmt, ...) { FT_VLOG_BODY(vlog(level, fmt, ap)); }
void fileLogSingle(enum logLevel level, const char *fmt, ...) { FT_VLOG_BODY(vlog(level, fmt, ap)); }
void fileLogMatrix(enum logLevel level, const double *mat, const int n, const int m, const char *fmt, ...) {
  FT_VLOG_BODY(log_matrix(level, mat, n, m, false, fmt, ap));
}
void fileLogMatrixRM(enum logLevel level, const double *mat, const int n, const int m, const char *fmt, ...) {
  FT_VLOG_BODY(log_matrix(level, mat, n, m, true, fmt, ap));
}
void fileLogMatrixSingle(enum logLevel level, const double *mat, const int n, const int m, const char *fmt,
                         ...) {
  FT_VLOG_BODY(log_matrix(level, mat, n, m, false, fmt, ap));
}
void fileLogMatrixRMSingle(enum logLevel level, const double *mat, const int n, const int m,
                           const char *fmt, ...) {
  FT_VLOG_BODY(log_matrix(level, mat, n, m, true, fmt, ap));
}
void fileLogArray(enum logLevel level, const double *arr, const int n, const char *fmt, ...) {
  FT_VLOG_BODY(log_array(level, arr, n, fmt, ap));
}
void fileLogArrayMaster(enum logLevel level, const double *arr, const int n, const char *fmt, ...) {
  FT_VLOG_BODY(log_array(level, arr, n, fmt, ap));
}
void fileLogArraySingle(enum logLevel level, const double *arr, const int n, const char *fmt, ...) {
  FT_VLOG_BODY(log_array(level, arr, n, fmt, ap));
}
void fileLogArrayInt(enum logLevel level, const int *arr, const int n, const char *fmt, ...) {
  FT_VLOG_BODY(log_array(level, arr, n, fmt, ap));
}
void fileLogArrayIntMaster(enum logLevel level, const int *arr, const int n, const char *fmt, ...) {
  FT_VLOG_BODY(log_array(level, arr, n, fmt, ap));
}
void fileLogArrayIntSingle(enum logLevel level, const int *arr, const int n, const char *fmt, ...) {
  FT_VLOG_BODY(log_array(level, arr, n, fmt, ap));
}

// ---- start-up and shut-down (InitFinalizeFemTech.cpp:47-86) -----------------
void InitFemTechWoInput(int argc, char **argv) {
  world_rank = 0;
  world_size = 1;
  if (argc < 1) {
    fprintf(stdout, "ERROR(%d): Please provide an input file\n", world_rank);
    fprintf(stdout, "ERROR(%d): Usage %s inputFile\n", world_rank, argv ? argv[0] : "driver");
    exit(EXIT_FAILURE);
  }
  // Simulation id from the wall clock, as the reference; FEMTECH_UID fixes it
  // (e.g. for tests that look for the output files).
  const char *fixed = getenv("FEMTECH_UID");
  if (fixed && *fixed) {
    uid = fixed;
  } else {
    time_t t = time(nullptr);
    struct tm r;
    char buf[80];
    strftime(buf, sizeof(buf), "%d_%m_%Y_%H_%M_%S", localtime_r(&t, &r));
    uid = buf;
  }
  initLog(("femtech_" + uid + ".log").c_str());
  FILE_LOG_MASTER(INFO, "MI355X engine, C ABI version %d", ft_version());
}

void FinalizeFemTech() {
  FILE_LOG_MASTER(INFO, "Terminating Application Normally");
  finaliseLog();
  FreeArrays();
}

void TerminateFemTech(int errorCode) {
  finaliseLog();
  FreeArrays();
  exit(errorCode);
}

// ---- input (ReadInputFile.cpp:26-78, ReadMaterials.cpp:8-138) ----------------
void ReadInputFile(const char *FileName) {
  if (world_size < 1 || world_rank < 0 || world_rank >= world_size) {
    FILE_LOG_SINGLE(ERROR, "'World_size' and/or 'world_rank' variable is not valid");
    TerminateFemTech(3);
  }
  if (FileName == nullptr || strlen(FileName) == 0) {
    FILE_LOG_SINGLE(ERROR, "Input file name is empty");
    TerminateFemTech(3);
  }
  const char *ext = strrchr(FileName, '.');
  if (!ext || (strcasecmp(ext, ".k") != 0 && strcasecmp(ext, ".inp") != 0)) {
    FILE_LOG_SINGLE(ERROR, "Input file type is unknown");
    FILE_LOG_SINGLE(ERROR, "Input mesh file read failed");
    TerminateFemTech(3);
  }
  int nn = 0, ne = 0;
  if (ft_read_mesh(FileName, &nn, &ne, nullptr, nullptr, nullptr, nullptr)) {
    FILE_LOG_SINGLE(ERROR, "%s", ft_last_error());
    FILE_LOG_SINGLE(ERROR, "Input mesh file read failed");
    TerminateFemTech(3);
  }
  coordinates = zalloc<double>(3 * (size_t)nn, "coordinates");
  connectivity = zalloc<int>(8 * (size_t)ne, "connectivity");
  pid = zalloc<int>(ne, "pid");
  globalNodeID = zalloc<int>(nn, "globalNodeID");
  if (ft_read_mesh(FileName, &nn, &ne, coordinates, connectivity, pid, globalNodeID)) {
    FILE_LOG_SINGLE(ERROR, "%s", ft_last_error());
    FILE_LOG_SINGLE(ERROR, "Input mesh file read failed");
    TerminateFemTech(3);
  }
  nNodes = nn;
  nelements = ne;
  nallelements = ne;
  ndim = 3;
  nDOF = 3 * nn;
  eptr = zalloc<int>((size_t)ne + 1, "eptr");
  global_eid = zalloc<int>(ne, "global_eid");
  nShapeFunctions = zalloc<int>(ne, "nShapeFunctions");
  GaussPoints = zalloc<int>(ne, "GaussPoints");
  ElementType = zalloc<char *>(ne, "ElementType");
  for (int e = 0; e < ne; ++e) {
    eptr[e + 1] = 8 * (e + 1);
    global_eid[e] = e;
    nShapeFunctions[e] = 8;
    GaussPoints[e] = 8;
    ElementType[e] = zalloc<char>(MAX_ELEMENT_TYPE_SIZE, "ElementType");
    strcpy(ElementType[e], "C3D8");
  }
  // distinct part ids, and the global count = largest id + 1
  std::vector<char> seen;
  int pmax = -1;
  nPID = 0;
  for (int e = 0; e < ne; ++e) {
    if (pid[e] >= (int)seen.size()) seen.resize(pid[e] + 1, 0);
    if (!seen[pid[e]]) {
      seen[pid[e]] = 1;
      ++nPID;
    }
    pmax = pid[e] > pmax ? pid[e] : pmax;
  }
  nPIDglobal = pmax + 1;
  nparts = 1;
}

void ReadMaterials() {
  FILE_LOG(INFO, "Number of parts = %d", nPIDglobal);
  materialID = zalloc<int>(nPIDglobal, "materialID");
  properties = zalloc<double>((size_t)nPIDglobal * MAXMATPARAMS, "properties");
  int rc = ft_read_materials("materials.dat", nPIDglobal, materialID, properties);
  if (rc) die(rc, "ReadMaterials");
}

// One GPU holds the whole mesh: a single partition, no halo exchange.  The
// one-rank node numbering of PartitionMesh.cpp:485-536 is already applied by
// ReadInputFile.
void PartitionMesh() {
  sendProcessCount = 0;
  sendNeighbourCountCum = zalloc<int>(1, "sendNeighbourCountCum");
  FILE_LOG_MASTER(INFO, "Single partition: %d elements, %d nodes", nelements, nNodes);
}

// ---- arrays (AllocateArrays.cpp:29-153, FreeArrays.cpp:4-83) ----------------
void AllocateArrays() {
  displacements = zalloc<double>(nDOF, "displacements");
  boundary = zalloc<int>(nDOF, "boundary");
  if (ImplicitDynamic || ExplicitDynamic) {
    velocities = zalloc<double>(nDOF, "velocities");
    accelerations = zalloc<double>(nDOF, "accelerations");
    velocities_half = zalloc<double>(nDOF, "velocities_half");
    displacements_prev = zalloc<double>(nDOF, "displacements_prev");
    accelerations_prev = zalloc<double>(nDOF, "accelerations_prev");
    Eavg = zalloc<double>((size_t)nelements * 9, "Eavg");
    fe = zalloc<double>(nDOF, "fe");
    fe_prev = zalloc<double>(nDOF, "fe_prev");
    fi = zalloc<double>(nDOF, "fi");
    fi_prev = zalloc<double>(nDOF, "fi_prev");
    f_net = zalloc<double>(nDOF, "f_net");
    fr_prev = zalloc<double>(nDOF, "fr_prev");
    fr_curr = zalloc<double>(nDOF, "fr_curr");
    f_damp_curr = zalloc<double>(nDOF, "f_damp_curr");
    f_damp_prev = zalloc<double>(nDOF, "f_damp_prev");
    energyFile = fopen(("energy_" + uid + ".dat").c_str(), "w");
    if (!energyFile) {
      FILE_LOG_SINGLE(ERROR, "Cannot open energy file");
      TerminateFemTech(3);
    }
    fprintf(energyFile, "# Energy for FEM\n");
    fprintf(energyFile, "# Time  Winternal   Wexternal   WKE   total\n");
    stepTime = zalloc<double>(MAXPLOTSTEPS, "stepTime");
    g_principal.assign(3 * (size_t)nelements, 0.0);
  }
}

void FreeArrays() {
  if (g_engine) {
    ft_engine_destroy(g_engine);
    g_engine = nullptr;
  }
  release(coordinates), release(connectivity), release(globalNodeID), release(pid), release(global_eid);
  release(eptr), release(shp), release(dshp), release(dsptr), release(gptr), release(nShapeFunctions);
  release(C), release(gaussWeights), release(gpPtr), release(detJacobian), release(mass);
  release(stiffness), release(rhs), release(displacements), release(velocities), release(accelerations);
  release(accelerations_prev), release(boundary), release(velocities_half), release(fe), release(fi);
  release(f_net), release(fr_prev), release(fr_curr), release(fi_prev), release(fe_prev);
  release(f_damp_prev), release(f_damp_curr), release(displacements_prev), release(F), release(pk2);
  release(pk2ptr), release(materialID), release(properties), release(detF), release(invF), release(Eavg);
  release(detFptr), release(InternalsPtr), release(internals), release(GaussPoints), release(fptr);
  release(recvNodeDisplacement), release(sendProcessID), release(sendNeighbourCount);
  release(sendNeighbourCountCum), release(sendNodeIndex), release(sendNodeDisplacement), release(stepTime);
  release(mat1), release(mat2), release(mat3), release(mat4), release(fintGQ), release(B);
  release(Hn_1), release(Hn_2), release(S0n);
  if (ElementType) {
    for (int e = 0; e < nelements; ++e) free(ElementType[e]);
    free(ElementType);
    ElementType = nullptr;
  }
  if (energyFile) fclose(energyFile);
  energyFile = nullptr;
  g_principal.clear();
}

// ---- the explicit path ---------------------------------------------------------
// ShapeFunctions.cpp:32-255: the engine uploads the mesh and materials once
// (its element kernel recomputes shape-function gradients in registers every
// step), and FemTech's per-Gauss-point arrays are set up as the reference
// leaves them: pointer tables, N, dN/dX and det J at the 8 Gauss points
// (evaluated on the device), unit weights, F = invF = I, det F = 1, zero PK2
// and internals, and the Prony history arrays when a part uses material 5
// (or this engine's Ogden + Prony law, 7).
void ShapeFunctions() {
  if (g_engine) ft_engine_destroy(g_engine);
  g_engine = nullptr;
  if (!materialID || !properties) {
    FILE_LOG_SINGLE(ERROR, "ShapeFunctions: materials not read");
    TerminateFemTech(3);
  }
  check(ft_engine_create(nNodes, nelements, coordinates, connectivity, pid, nPIDglobal, materialID, properties,
                         precision_from_env(), &g_engine),
        "ShapeFunctions");
  const size_t ne = (size_t)nelements;
  release(GaussPoints), release(nShapeFunctions), release(gptr), release(dsptr), release(gpPtr);
  release(fptr), release(pk2ptr), release(detFptr), release(InternalsPtr), release(shp), release(dshp);
  release(detJacobian), release(gaussWeights), release(F), release(detF), release(invF), release(internals);
  release(pk2), release(fintGQ), release(B), release(Hn_1), release(Hn_2), release(S0n);
  GaussPoints = zalloc<int>(ne, "GaussPoints");
  nShapeFunctions = zalloc<int>(ne, "nShapeFunctions");
  gptr = zalloc<int>(ne + 1, "gptr");
  dsptr = zalloc<int>(ne + 1, "dsptr");
  gpPtr = zalloc<int>(ne + 1, "gpPtr");
  fptr = zalloc<int>(ne + 1, "fptr");
  pk2ptr = zalloc<int>(ne + 1, "pk2ptr");
  detFptr = zalloc<int>(ne + 1, "detFptr");
  InternalsPtr = zalloc<int>(ne + 1, "InternalsPtr");
  for (size_t e = 0; e < ne; ++e) {  // C3D8 only: 8 Gauss points, 8 shape functions
    GaussPoints[e] = 8;
    nShapeFunctions[e] = 8;
    gptr[e + 1] = gptr[e] + 64;
    dsptr[e + 1] = dsptr[e] + 192;
    gpPtr[e + 1] = gpPtr[e] + 8;
    fptr[e + 1] = fptr[e] + 72;
    pk2ptr[e + 1] = pk2ptr[e] + 48;
    detFptr[e + 1] = detFptr[e] + 8;
    InternalsPtr[e + 1] = InternalsPtr[e] + MAXINTERNALVARS * 8;
  }
  shp = zalloc<double>(64 * ne, "shp");
  dshp = zalloc<double>(192 * ne, "dshp");
  detJacobian = zalloc<double>(8 * ne, "detJacobian");
  gaussWeights = zalloc<double>(8 * ne, "gaussWeights");
  F = zalloc<double>(72 * ne, "F");
  detF = zalloc<double>(8 * ne, "detF");
  invF = zalloc<double>(72 * ne, "invF");
  for (size_t q = 0; q < 8 * ne; ++q) {
    F[9 * q] = F[9 * q + 4] = F[9 * q + 8] = 1.0;
    invF[9 * q] = invF[9 * q + 4] = invF[9 * q + 8] = 1.0;
    detF[q] = 1.0;
  }
  internals = zalloc<double>((size_t)MAXINTERNALVARS * 8 * ne, "internals");
  pk2 = zalloc<double>(48 * ne, "pk2");
  fintGQ = zalloc<double>(24, "fintGQ");
  B = zalloc<double>(144, "B");
  for (int p = 0; p < nPIDglobal; ++p)
    if (materialID[p] == 5 || materialID[p] == 7) {  // Prony laws
      Hn_1 = zalloc<double>(72 * ne, "Hn_1");
      Hn_2 = zalloc<double>(72 * ne, "Hn_2");
      S0n = zalloc<double>(72 * ne, "S0n");
      break;
    }
  // N, dN/dX, det J at the Gauss points of every element, one device pass
  std::vector<double> in(27 * 8 * ne), out(33 * 8 * ne);
  double chi[24], w[8];
  GaussQuadrature3D(0, 8, chi, w);
  for (size_t e = 0; e < ne; ++e) {
    for (int gp = 0; gp < 8; ++gp) {
      double *r = &in[27 * (8 * e + gp)];
      for (int k = 0; k < 3; ++k) r[k] = chi[3 * gp + k];
      for (int a = 0; a < 8; ++a)
        for (int k = 0; k < 3; ++k) r[3 + 3 * a + k] = coordinates[3 * (size_t)connectivity[8 * e + a] + k];
      gaussWeights[8 * e + gp] = w[gp];
    }
  }
  check(ft_point_eval(FT_PT_SHAPE, (int)(8 * ne), in.data(), out.data()), "ShapeFunctions");
  for (size_t q = 0; q < 8 * ne; ++q) {
    const double *r = &out[33 * q];
    for (int a = 0; a < 8; ++a) shp[8 * q + a] = r[a];
    for (int k = 0; k < 24; ++k) dshp[24 * q + k] = r[8 + k];
    detJacobian[q] = r[32];
  }
  g_hist_host_newer = false;
}

// GaussQuadrature3D.cpp:3-60 (C3D8, 8 points): the 2x2x2 Gauss table in the
// reference's point order.
void GaussQuadrature3D(int element, int nGaussPoint, double *Chi, double *GaussWeights) {
  (void)element;
  if (nGaussPoint != 8) return;
  const double g = 0.577350269189626;
  for (int p = 0; p < 8; ++p) {
    const int q = p & 3;
    Chi[3 * p + 0] = (q == 1 || q == 2) ? g : -g;
    Chi[3 * p + 1] = (q >= 2) ? g : -g;
    Chi[3 * p + 2] = p < 4 ? g : -g;
    GaussWeights[p] = 1.0;
  }
}

void AssembleLumpedMass(void) {
  if (!g_engine) {
    FILE_LOG_SINGLE(ERROR, "AssembleLumpedMass before ShapeFunctions");
    TerminateFemTech(3);
  }
  release(mass);
  mass = zalloc<double>(nDOF, "mass");
  check(ft_assemble_lumped_mass(g_engine, mass), "AssembleLumpedMass");
}

double StableTimeStep() {
  double dt_min = 0.0;
  int rc = ft_stable_time_step(g_engine, displacements, boundary, FailureTimeStep, &dt_min);
  if (rc == FT_ERR_DT) {
    FILE_LOG_MASTER(ERROR, "Timestep too small, dt = %15.9e", dt_min);
    TerminateFemTech(19);
  }
  check(rc, "StableTimeStep");
  return dt_min;
}

void GetForce() {
  if (ndim != 3) {
    FILE_LOG_SINGLE(ERROR, "GetForce function not yet implemented for %dD", ndim);
    TerminateFemTech(3);
  }
  GetForce_3D();
}

// GetForce_3D.cpp:5-53: fi from the element kernel and the node gather (the
// Prony history advances with the driver's dt); f_net = fe - fi.  The
// per-Gauss-point F, det F and PK2 the reference's loop leaves in its arrays
// come back from the same pass, and so does the Prony history (Hn_1, Hn_2,
// S0n as full 3 x 3 matrices); a history changed on the host by the
// per-point StressUpdate goes to the device first.
void GetForce_3D() {
  const size_t ne = (size_t)nelements;
  if (Hn_1 && g_hist_host_newer) {
    std::vector<double> h(144 * ne);
    static const int vi[6] = {0, 4, 8, 7, 6, 3};
    for (size_t e = 0; e < ne; ++e)
      for (int gp = 0; gp < 8; ++gp)
        for (int t = 0; t < 3; ++t) {
          const double *m = (t == 0 ? Hn_1 : t == 1 ? Hn_2 : S0n) + fptr[e] + 9 * gp;
          for (int k = 0; k < 6; ++k) h[144 * e + 18 * gp + 6 * t + k] = m[vi[k]];
        }
    check(ft_set_visco_history(g_engine, h.data()), "GetForce_3D");
    g_hist_host_newer = false;
  }
  std::vector<double> stress(F ? 48 * ne : 0);
  check(ft_get_force_fields(g_engine, displacements, dt, fi, F, detF, F ? stress.data() : nullptr), "GetForce_3D");
  if (F)  // rigid parts: StressUpdate leaves their PK2 untouched (StressUpdate.cpp:8-9)
    for (size_t e = 0; e < ne; ++e)
      if (materialID[pid[e]] != 0) memcpy(pk2 + pk2ptr[e], &stress[48 * e], 48 * sizeof(double));
  if (Hn_1) {
    std::vector<double> h(144 * ne);
    check(ft_get_visco_history(g_engine, h.data()), "GetForce_3D");
    for (size_t e = 0; e < ne; ++e) {
      if (materialID[pid[e]] != 5 && materialID[pid[e]] != 7) continue;
      for (int gp = 0; gp < 8; ++gp)
        for (int t = 0; t < 3; ++t) {
          double *m = (t == 0 ? Hn_1 : t == 1 ? Hn_2 : S0n) + fptr[e] + 9 * gp;
          const double *v = &h[144 * e + 18 * gp + 6 * t];
          m[0] = v[0], m[4] = v[1], m[8] = v[2];
          m[5] = m[7] = v[3];
          m[2] = m[6] = v[4];
          m[1] = m[3] = v[5];
        }
    }
  }
  for (int i = 0; i < nDOF; ++i) f_net[i] = fe[i] - fi[i];
}

// ---- per-element, per-Gauss-point routines ------------------------------------
// Each gathers its inputs from the global arrays the reference routine reads,
// evaluates on the device (ft_point_eval) and writes the outputs where the
// reference writes them.  Calls need ShapeFunctions first (it sets up the
// arrays), as in the reference.
namespace {

void need_arrays(const char *who) {
  if (!dshp || !F) {
    FILE_LOG_SINGLE(ERROR, "%s before ShapeFunctions", who);
    TerminateFemTech(3);
  }
}

void point(int op, const double *in, double *out, const char *who) {
  check(ft_point_eval(op, 1, in, out), who);
}

// X + u of element e's nodes in local order (CalculateDeformationGradient.cpp:20,
// CalculateCharacteristicLength_C3D8.cpp:6-11).
void deformed_nodes(int e, double *x) {
  for (int a = 0; a < 8; ++a)
    for (int k = 0; k < 3; ++k) {
      const size_t i = 3 * (size_t)connectivity[eptr[e] + a] + k;
      x[3 * a + k] = coordinates[i] + displacements[i];
    }
}

// The material routines with the material fixed (StressUpdate.cpp:7-23).
void material_update(int e, int gp, int mat, const char *who) {
  need_arrays(who);
  double in[20], out[34] = {0};
  memcpy(in, F + fptr[e] + 9 * gp, 9 * sizeof(double));
  in[9] = mat;
  memcpy(in + 10, properties + MAXMATPARAMS * pid[e], 9 * sizeof(double));
  in[19] = dt;
  const bool visco = mat == 5 || mat == 7;
  if (visco) {
    if (!Hn_1) {
      FILE_LOG_SINGLE(ERROR, "%s: no viscoelastic part was set up", who);
      TerminateFemTech(3);
    }
    memcpy(out + 6, Hn_1 + fptr[e] + 9 * gp, 9 * sizeof(double));
    memcpy(out + 15, Hn_2 + fptr[e] + 9 * gp, 9 * sizeof(double));
    memcpy(out + 24, S0n + fptr[e] + 9 * gp, 9 * sizeof(double));
  }
  point(FT_PT_STRESS, in, out, who);
  memcpy(pk2 + pk2ptr[e] + 6 * gp, out, 6 * sizeof(double));
  if (visco) {
    memcpy(Hn_1 + fptr[e] + 9 * gp, out + 6, 9 * sizeof(double));
    memcpy(Hn_2 + fptr[e] + 9 * gp, out + 15, 9 * sizeof(double));
    memcpy(S0n + fptr[e] + 9 * gp, out + 24, 9 * sizeof(double));
    g_hist_host_newer = true;
  }
}

}  // namespace

// ShapeFunction_C3D8.cpp:4-128
void ShapeFunction_C3D8(int e, int gp, double *Chi, double *detJ) {
  need_arrays("ShapeFunction_C3D8");
  double in[27], out[33];
  for (int k = 0; k < 3; ++k) in[k] = Chi[3 * gp + k];
  for (int a = 0; a < 8; ++a)
    for (int k = 0; k < 3; ++k) in[3 + 3 * a + k] = coordinates[3 * (size_t)connectivity[eptr[e] + a] + k];
  point(FT_PT_SHAPE, in, out, "ShapeFunction_C3D8");
  memcpy(shp + gptr[e] + 8 * gp, out, 8 * sizeof(double));
  memcpy(dshp + dsptr[e] + 24 * gp, out + 8, 24 * sizeof(double));
  detJ[gp] = out[32];
}

// CalculateDeformationGradient.cpp:11-25
void CalculateDeformationGradient(int e, int gp) {
  need_arrays("CalculateDeformationGradient");
  double in[48];
  deformed_nodes(e, in);
  memcpy(in + 24, dshp + dsptr[e] + 24 * gp, 24 * sizeof(double));
  point(FT_PT_DEFGRAD, in, F + fptr[e] + 9 * gp, "CalculateDeformationGradient");
}

// DeterminateF.cpp:38-57 (writes at gpPtr, as the reference does)
void DeterminateF(int e, int gp) {
  need_arrays("DeterminateF");
  point(FT_PT_DETF, F + fptr[e] + 9 * gp, detF + gpPtr[e] + gp, "DeterminateF");
}

// InverseF.cpp (3-D branch): F and det F from the global arrays
void InverseF(int e, int gp, double *fInv) {
  need_arrays("InverseF");
  double in[10];
  memcpy(in, F + fptr[e] + 9 * gp, 9 * sizeof(double));
  in[9] = detF[detFptr[e] + gp];
  point(FT_PT_INVF, in, fInv, "InverseF");
}

// StressUpdate.cpp:5-29
void StressUpdate(int e, int gp) {
  const int mat = materialID[pid[e]];
  if (mat == 0) return;  // rigid-body motion
  if (mat < 0 || mat > 8) {
    FILE_LOG_SINGLE(ERROR, "Unknown material type");
    TerminateFemTech(1);
  }
  material_update(e, gp, mat, "StressUpdate");
}

void CompressibleNeoHookean(int e, int gp) { material_update(e, gp, 1, "CompressibleNeoHookean"); }
void StVenantKirchhoff(int e, int gp) { material_update(e, gp, 2, "StVenantKirchhoff"); }
void LinearElastic(int e, int gp) { material_update(e, gp, 3, "LinearElastic"); }
void HGOIsotropic(int e, int gp) { material_update(e, gp, 4, "HGOIsotropic"); }
void HGOIsotropicViscoelastic(int e, int gp) { material_update(e, gp, 5, "HGOIsotropicViscoelastic"); }

// StrainDisplacementMatrix.cpp:25-77
void StrainDisplacementMatrix(int e, int gp, int nI, double *Bmat) {
  need_arrays("StrainDisplacementMatrix");
  double in[12];
  memcpy(in, dshp + dsptr[e] + 24 * gp + 3 * nI, 3 * sizeof(double));
  memcpy(in + 3, F + fptr[e] + 9 * gp, 9 * sizeof(double));
  point(FT_PT_BMAT, in, Bmat, "StrainDisplacementMatrix");
}

// InternalForceUpdate.cpp:4-28 (also leaves B and fintGQ as the reference does)
void InternalForceUpdate(int e, int gp, double *force) {
  need_arrays("InternalForceUpdate");
  double in[64], out[192];
  memcpy(in, dshp + dsptr[e] + 24 * gp, 24 * sizeof(double));
  memcpy(in + 24, F + fptr[e] + 9 * gp, 9 * sizeof(double));
  memcpy(in + 33, pk2 + pk2ptr[e] + 6 * gp, 6 * sizeof(double));
  in[39] = gaussWeights[gpPtr[e] + gp] * detJacobian[gpPtr[e] + gp];
  memcpy(in + 40, force, 24 * sizeof(double));
  point(FT_PT_FINT, in, out, "InternalForceUpdate");
  memcpy(force, out, 24 * sizeof(double));
  memcpy(fintGQ, out + 24, 24 * sizeof(double));
  memcpy(B, out + 48, 144 * sizeof(double));
}

// CalculateCharacteristicLength_C3D8.cpp:3-30 at the current displacements
double CalculateCharacteristicLength_C3D8(int e) {
  double in[24], out = 0.0;
  deformed_nodes(e, in);
  point(FT_PT_CHARLEN, in, &out, "CalculateCharacteristicLength_C3D8");
  return out;
}

// CalculateCharacteristicLength.cpp (C3D8 is the only element type here)
double CalculateCharacteristicLength(int e) { return CalculateCharacteristicLength_C3D8(e); }

// CalculateTimeStep.cpp:7-21
double CalculateTimeStep(int e) {
  double in[27], out = 0.0;
  deformed_nodes(e, in);
  const double *pr = properties + MAXMATPARAMS * pid[e];
  in[24] = pr[0];
  in[25] = pr[1];
  in[26] = pr[2];
  point(FT_PT_TIMESTEP, in, &out, "CalculateTimeStep");
  return out;
}

// math.cpp:8-28
void inverse3x3Matrix(double *mat, double *invMat, double *det) {
  double out[10];
  point(FT_PT_INV3, mat, out, "inverse3x3Matrix");
  memcpy(invMat, out, 9 * sizeof(double));
  *det = out[9];
}

void CalculateAccelerations() {
  check(ft_accelerations(g_engine, f_net, boundary, accelerations), "CalculateAccelerations");
}

// CheckEnergy.cpp:3-85 (one rank): the device returns the step's sums; the
// running works are kept here as in the reference's static accumulators.
void CheckEnergy(double time, int writeFlag) {
  double s[3];
  check(ft_energy_terms(g_engine, displacements, displacements_prev, velocities, accelerations,
                        accelerations_prev, fi, fi_prev, fe, fe_prev, boundary, s),
        "CheckEnergy");
  const double wke = 0.5 * s[2];
  g_wint += 0.5 * s[0];
  g_wext += 0.5 * s[1];
  const double total = fabs(wke + g_wint - g_wext);
  FILE_LOG_SINGLE(DEBUGLOG, "Internal Work : %15.9e", g_wint);
  FILE_LOG_SINGLE(DEBUGLOG, "External Work : %15.9e", g_wext);
  FILE_LOG_SINGLE(DEBUGLOG, "Kinetic Energy : %15.9e", wke);
  FILE_LOG_SINGLE(DEBUGLOG, "Total Energy : %15.9e", total);
  double mx = fabs(g_wint);
  if (mx < fabs(g_wext)) mx = fabs(g_wext);
  if (mx < fabs(wke)) mx = fabs(wke);
  if (total > 0.01 * mx)
    FILE_LOG_MASTER(WARNING, "Energy Violation. Total = %15.9e, Max = %15.9e, Error%% : %10.2f", total, mx,
                    total * 100.0 / mx);
  if (writeFlag == 0 && energyFile)
    fprintf(energyFile, "%12.6e %12.6e  %12.6e  %12.6e %12.6e\n", time, g_wint, g_wext, wke, total);
}

// CalculateStrain.cpp:77-97 into Eavg; the principal values of
// CalculateStrain.cpp:8-75 come from the same device pass.
void CalculateStrain() {
  g_principal.resize(3 * (size_t)nelements);
  check(ft_element_strain(g_engine, displacements, Eavg, g_principal.data()), "CalculateStrain");
}

void CalculateMaximumPrincipalStrain(int elm, double *currentStrainMax, double *currentStrainMin,
                                     double *currentShearMax) {
  const bool have = elm >= 0 && (size_t)(3 * elm + 2) < g_principal.size();
  *currentStrainMax = have ? g_principal[3 * (size_t)elm] : 0.0;
  *currentStrainMin = have ? g_principal[3 * (size_t)elm + 1] : 0.0;
  *currentShearMax = have ? g_principal[3 * (size_t)elm + 2] : 0.0;
}

// ---- output (WriteVTU.cpp:3-263, WritePVD.cpp:6-49) ---------------------------
// Same VTK XML layout, field names and number formats as the reference, so
// its post-processing reads these files unchanged.  results/ and results/vtu/
// are created when missing.
void WriteVTU(const char *FileName, int step, int **intCellData, int cellDataCount,
              const char **cellDataNames) {
  make_dir("results");
  make_dir("results/vtu");
  char path[1200];
  snprintf(path, sizeof(path), "./results/vtu/%s.vtu.%04d.%04d", FileName, step, world_rank);
  FILE *fp = fopen(path, "w");
  if (!fp) {
    FILE_LOG_SINGLE(ERROR, "Cannot open %s", path);
    TerminateFemTech(3);
  }
  const char *vec3 = "NumberOfComponents=\"%d\" ComponentName0=\"X\" ComponentName1=\"Y\" "
                     "ComponentName2=\"Z\" format=\"ascii\"";
  const char *strain9 = "NumberOfComponents=\"%d\" ComponentName0=\"E11\" ComponentName1=\"E21\" "
                        "ComponentName2=\"E31\" ComponentName3=\"E12\" ComponentName4=\"E22\" "
                        "ComponentName5=\"E32\" ComponentName6=\"E13\" ComponentName7=\"E23\" "
                        "ComponentName8=\"E33\" format=\"ascii\"";
  auto open_array = [&](const char *type, const char *name, const char *attrs, int ncomp) {
    fprintf(fp, "\t\t\t\t<DataArray type=\"%s\"", type);
    if (name) fprintf(fp, " Name=\"%s\"", name);
    fprintf(fp, " ");
    fprintf(fp, attrs, ncomp);
    fprintf(fp, ">\n");
  };
  fprintf(fp, "<?xml version=\"1.0\"?>\n");
  fprintf(fp, "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n");
  fprintf(fp, "\t<UnstructuredGrid>\n");
  fprintf(fp, "\t\t<Piece NumberOfPoints=\"%d\" NumberOfCells=\"%d\">\n", nNodes, nelements);
  fprintf(fp, "\t\t\t<Points>\n");
  open_array("Float64", nullptr, "NumberOfComponents=\"%d\" format=\"ascii\"", ndim);
  for (int i = 0; i < nNodes; ++i) {
    fprintf(fp, "\t\t\t\t\t");
    for (int j = 0; j < ndim; ++j) {
      double &d = displacements[ndim * i + j];
      if (fabs(d) < 1e-20) d = 0.0;  // the reference flushes tiny displacements in place
      fprintf(fp, "%10.8e ", coordinates[ndim * i + j] + d);
    }
    fprintf(fp, "\n");
  }
  fprintf(fp, "\t\t\t\t</DataArray>\n\t\t\t</Points>\n");
  fprintf(fp, "\t\t\t<Cells>\n");
  fprintf(fp, "\t\t\t\t<DataArray type=\"Int32\" Name=\"connectivity\" format=\"ascii\">\n");
  for (int i = 0; i < nelements; ++i) {
    fprintf(fp, "\t\t\t\t\t");
    for (int j = eptr[i]; j < eptr[i + 1]; ++j) fprintf(fp, "%d ", connectivity[j]);
    fprintf(fp, "\n");
  }
  fprintf(fp, "\t\t\t\t</DataArray>\n");
  fprintf(fp, "\t\t\t\t<DataArray type=\"Int32\" Name=\"offsets\" format=\"ascii\">\n");
  for (int i = 0; i < nelements; ++i) fprintf(fp, "\t\t\t\t\t%d\n", eptr[i + 1]);
  fprintf(fp, "\t\t\t\t</DataArray>\n");
  fprintf(fp, "\t\t\t\t<DataArray type=\"Int32\" Name=\"types\" format=\"ascii\">\n");
  for (int i = 0; i < nelements; ++i) fprintf(fp, "\t\t\t\t\t%d\n", 12);  // VTK_HEXAHEDRON
  fprintf(fp, "\t\t\t\t</DataArray>\n\t\t\t</Cells>\n");
  fprintf(fp, "\t\t\t<PointData>\n");
  const struct {
    const char *name;
    const double *v;
  } vecs[2] = {{"Displacements", displacements}, {"Accelerations", accelerations}};
  for (const auto &f : vecs) {
    open_array("Float64", f.name, vec3, ndim);
    for (int i = 0; i < nNodes; ++i) {
      fprintf(fp, "\t\t\t\t\t");
      for (int j = 0; j < ndim; ++j) fprintf(fp, "%10.8e ", f.v ? f.v[ndim * i + j] : 0.0);
      fprintf(fp, "\n");
    }
    fprintf(fp, "\t\t\t\t</DataArray>\n");
  }
  open_array("Int32", "Boundary", vec3, ndim);
  for (int i = 0; i < nNodes; ++i) {
    fprintf(fp, "\t\t\t\t\t");
    for (int j = 0; j < ndim; ++j) fprintf(fp, "%d ", boundary[ndim * i + j]);
    fprintf(fp, "\n");
  }
  fprintf(fp, "\t\t\t\t</DataArray>\n");
  fprintf(fp, "\t\t\t</PointData>\n");
  fprintf(fp, "\t\t\t<CellData>\n");
  fprintf(fp, "\t\t\t\t<DataArray type=\"Int32\" Name=\"PartID\" format=\"ascii\">\n");
  for (int i = 0; i < nelements; ++i) fprintf(fp, "\t\t\t\t\t%d\n", pid[i]);
  fprintf(fp, "\t\t\t\t</DataArray>\n");
  open_array("Float64", "AvgStrain", strain9, ndim * ndim);
  for (int i = 0; i < nelements; ++i) {
    fprintf(fp, "\t\t\t\t\t");
    for (int j = 0; j < ndim * ndim; ++j) fprintf(fp, "%10.8e ", Eavg ? Eavg[i * ndim * ndim + j] : 0.0);
    fprintf(fp, "\n");
  }
  fprintf(fp, "\t\t\t\t</DataArray>\n");
  fprintf(fp, "\t\t\t\t<DataArray type=\"Int32\" Name=\"ProcID\" format=\"ascii\">\n");
  for (int i = 0; i < nelements; ++i) fprintf(fp, "\t\t\t\t\t%d\n", world_rank);
  fprintf(fp, "\t\t\t\t</DataArray>\n");
  for (int c = 0; c < cellDataCount; ++c) {
    fprintf(fp, "\t\t\t\t<DataArray type=\"Int32\" Name=\"%s\" format=\"ascii\">\n", cellDataNames[c]);
    for (int i = 0; i < nelements; ++i) fprintf(fp, "\t\t\t\t\t%d\n", intCellData[c][i]);
    fprintf(fp, "\t\t\t\t</DataArray>\n");
  }
  fprintf(fp, "\t\t\t</CellData>\n\t\t</Piece>\n\t</UnstructuredGrid>\n</VTKFile>\n");
  fclose(fp);

  // the parallel index file (one piece: this rank)
  snprintf(path, sizeof(path), "./results/%s.%04d.pvtu", FileName, step);
  fp = fopen(path, "w");
  if (!fp) {
    FILE_LOG_SINGLE(ERROR, "Cannot open %s", path);
    TerminateFemTech(3);
  }
  fprintf(fp, "<?xml version=\"1.0\"?>\n");
  fprintf(fp, "<VTKFile type=\"PUnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n");
  fprintf(fp, "\t<PUnstructuredGrid GhostLevel=\"0\">\n");
  fprintf(fp, "\t\t<PPoints>\n");
  fprintf(fp, "\t\t\t<PDataArray type=\"Float64\" NumberOfComponents=\"%d\" format=\"ascii\"/>\n", ndim);
  fprintf(fp, "\t\t</PPoints>\n\t\t<PCells>\n");
  for (const char *n : {"connectivity", "offsets", "types"})
    fprintf(fp, "\t\t\t<PDataArray type=\"Int32\" Name=\"%s\" NumberOfComponents=\"1\"/>\n", n);
  fprintf(fp, "\t\t</PCells>\n");
  fprintf(fp, "\t\t<PPointData  Vectors=\"Displacements Accelerations\" >\n");
  for (const char *n : {"Displacements", "Accelerations"}) {
    fprintf(fp, "\t\t\t<PDataArray type=\"Float64\" Name=\"%s\" ", n);
    fprintf(fp, vec3, ndim);
    fprintf(fp, " />\n");
  }
  fprintf(fp, "\t\t\t<PDataArray type=\"Int32\" Name=\"Boundary\" ");
  fprintf(fp, vec3, ndim);
  fprintf(fp, " />\n");
  fprintf(fp, "\t\t</PPointData>\n");
  fprintf(fp, "\t\t<PCellData Scalars=\"PartID\" Tensors=\"AvgStrain\">\n");
  fprintf(fp, "\t\t\t<PDataArray type=\"Int32\" Name=\"PartID\"/>\n");
  fprintf(fp, "\t\t\t<PDataArray type=\"Float64\" Name=\"AvgStrain\" ");
  fprintf(fp, strain9, ndim * ndim);
  fprintf(fp, "/>\n");
  fprintf(fp, "\t\t\t<PDataArray type=\"Int32\" Name=\"ProcID\"/>\n");
  for (int c = 0; c < cellDataCount; ++c)
    fprintf(fp, "\t\t\t<PDataArray type=\"Int32\" Name=\"%s\"/>\n", cellDataNames[c]);
  fprintf(fp, "\t\t</PCellData>\n");
  for (int r = 0; r < world_size; ++r)
    fprintf(fp, "\t\t<Piece Source=\"vtu/%s.vtu.%04d.%.4d\"/>\n", FileName, step, r);
  fprintf(fp, "\t</PUnstructuredGrid>\n</VTKFile>\n");
  fclose(fp);
}

void WritePVD(const char *FileName, int step) {
  std::string base(FileName);
  const size_t dot = base.find_last_of("./\\");
  if (dot != std::string::npos && dot > 0 && base[dot] == '.') base.resize(dot);
  FILE *fp = fopen((base + ".pvd").c_str(), "w");
  if (!fp) {
    FILE_LOG_SINGLE(ERROR, "Cannot open %s.pvd", base.c_str());
    TerminateFemTech(3);
  }
  fprintf(fp, "<?xml version=\"1.0\"?>\n");
  fprintf(fp, "<VTKFile type=\"Collection\" version=\"0.1\" byte_order=\"LittleEndian\" "
              "compressor=\"vtkZLibDataCompressor\" >\n");
  fprintf(fp, "\t<Collection>\n");
  for (int j = 0; j <= step; ++j)
    fprintf(fp, "\t\t<DataSet timestep=\"%f\" file=\"results/%s.%.04d.pvtu\" />\n", stepTime[j], base.c_str(), j);
  fprintf(fp, "\t</Collection>\n</VTKFile>\n");
  fclose(fp);
}

// ---- outside the explicit engine -------------------------------------------------
namespace {
[[noreturn]] void unsupported(const char *what) {
  FILE_LOG_SINGLE(ERROR, "%s is not provided by the MI355X explicit engine", what);
  TerminateFemTech(3);
  exit(3);
}
}  // namespace
void CreateLinearElasticityCMatrix() { unsupported("CreateLinearElasticityCMatrix (implicit solver)"); }
void Assembly(char *) { unsupported("Assembly (implicit solver)"); }
void ApplySteadyBoundaryConditions(void) { unsupported("ApplySteadyBoundaryConditions (implicit solver)"); }
void SolveSteadyImplicit(void) { unsupported("SolveSteadyImplicit (implicit solver)"); }
void SolveUnsteadyNewmarkImplicit(double, double, double, double, char *) {
  unsupported("SolveUnsteadyNewmarkImplicit (implicit solver)");
}
