/*
 * FemTech.h -- FemTech's C++ call surface, served by the MI355X engine.
 *
 * Drivers written against FemTech (e.g. examples/ex9/ex9.cpp,
 * examples/ex7/ex7.cpp, examples/Benchmarking-Parallel/
 * Benchmarking-Parallel.cpp) compile unmodified against this header and link
 * against libFemTech.so from this package instead of the reference library.
 * Names, signatures, globals and logging macros follow the reference's
 * include/FemTech.h:22-129, include/GlobalVariables.h:18-127 and
 * include/parallel_log.h:18-105.
 *
 * Semantics ("compatibility mode"): the host arrays below stay authoritative,
 * exactly as drivers expect (they update velocities and displacements and
 * write boundary conditions themselves).  Each library routine uploads what
 * it reads, runs on the GPU through the C ABI in femtech_amd.h and downloads
 * what it writes.  The device-resident loop (ft_run_init / ft_run_steps) is
 * the fast path; this layer is for existing drivers.
 *
 * One process drives one GPU that holds the whole mesh, so the MPI
 * decomposition collapses to a single rank.  Unless FT_HAVE_MPI is defined
 * (the driver then includes a real mpi.h first), the few MPI names drivers use
 * are given one-rank definitions here.
 *
 * Per-Gauss-point fields (F, detF, invF, pk2, shp, dshp, ...): the fused
 * element kernel keeps them in registers, but ShapeFunctions sets the arrays
 * up as the reference does, GetForce_3D fills F, detF and pk2 (and the Prony
 * history) from the same device pass, and the per-element, per-Gauss-point
 * routines (ShapeFunction_C3D8, CalculateDeformationGradient, DeterminateF,
 * InverseF, StressUpdate and the material routines, StrainDisplacementMatrix,
 * InternalForceUpdate, CalculateTimeStep, CalculateCharacteristicLength,
 * inverse3x3Matrix) read and write them like the reference, each call
 * evaluated on the GPU.  Implicit solvers (Assembly, SolveSteadyImplicit,
 * SolveUnsteadyNewmarkImplicit, ...) are outside the engine's scope: they are
 * defined so that drivers link, and terminate with code 3 if called.
 */
#ifndef FEMTECH_H
#define FEMTECH_H

#include <limits.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <map>
#include <string>

/* ---- one-rank MPI names used by drivers ---------------------------------- */
#ifndef FT_HAVE_MPI
typedef int MPI_Comm;
#define MPI_COMM_WORLD ((MPI_Comm)0x44000000)
static inline int MPI_Barrier(MPI_Comm) { return 0; }
static inline int MPI_Comm_rank(MPI_Comm, int *rank) { *rank = 0; return 0; }
static inline int MPI_Comm_size(MPI_Comm, int *size) { *size = 1; return 0; }
#endif

/* ---- logging (parallel_log.h:18-105) ------------------------------------- */
enum logLevel { ERROR, WARNING, INFO, DEBUGLOG, DEBUGLOGIGNORE };

#ifndef FILELOG_MAX_LEVEL
#ifdef DEBUG
#define FILELOG_MAX_LEVEL DEBUGLOG
#else
#define FILELOG_MAX_LEVEL INFO
#endif
#endif

const char *levelToString(enum logLevel level);
void initLog(const char *outFileName);
void finaliseLog();
void fileLog(enum logLevel level, const char *fmt, ...);
void fileLogMaster(enum logLevel level, const char *fmt, ...);
void fileLogSingle(enum logLevel level, const char *fmt, ...);
void fileLogMatrix(enum logLevel level, const double *mat, const int n, const int m,
                   const char *fmt, ...);
void fileLogMatrixRM(enum logLevel level, const double *mat, const int n, const int m,
                     const char *fmt, ...);
void fileLogMatrixSingle(enum logLevel level, const double *mat, const int n, const int m,
                         const char *fmt, ...);
void fileLogMatrixRMSingle(enum logLevel level, const double *mat, const int n, const int m,
                           const char *fmt, ...);
void fileLogArray(enum logLevel level, const double *arr, const int n, const char *fmt, ...);
void fileLogArrayMaster(enum logLevel level, const double *arr, const int n, const char *fmt,
                        ...);
void fileLogArraySingle(enum logLevel level, const double *arr, const int n, const char *fmt,
                        ...);
void fileLogArrayInt(enum logLevel level, const int *arr, const int n, const char *fmt, ...);
void fileLogArrayIntMaster(enum logLevel level, const int *arr, const int n, const char *fmt,
                           ...);
void fileLogArrayIntSingle(enum logLevel level, const int *arr, const int n, const char *fmt,
                           ...);

#define FT_LOG_IF_(level, call) \
  do {                          \
    if ((level) <= FILELOG_MAX_LEVEL) call; \
  } while (0)
#define FILE_LOG(level, ...) FT_LOG_IF_(level, fileLog(level, __VA_ARGS__))
#define FILE_LOG_MASTER(level, ...) FT_LOG_IF_(level, fileLogMaster(level, __VA_ARGS__))
#define FILE_LOG_SINGLE(level, ...) FT_LOG_IF_(level, fileLogSingle(level, __VA_ARGS__))
#define FILE_LOGMatrix(level, mat, n, m, ...) \
  FT_LOG_IF_(level, fileLogMatrix(level, mat, n, m, __VA_ARGS__))
#define FILE_LOGMatrixRM(level, mat, n, m, ...) \
  FT_LOG_IF_(level, fileLogMatrixRM(level, mat, n, m, __VA_ARGS__))
#define FILE_LOGMatrix_SINGLE(level, mat, n, m, ...) \
  FT_LOG_IF_(level, fileLogMatrixSingle(level, mat, n, m, __VA_ARGS__))
#define FILE_LOGMatrixRM_SINGLE(level, mat, n, m, ...) \
  FT_LOG_IF_(level, fileLogMatrixRMSingle(level, mat, n, m, __VA_ARGS__))
#define FILE_LOGArray(level, arr, n, ...) FT_LOG_IF_(level, fileLogArray(level, arr, n, __VA_ARGS__))
#define FILE_LOGArrayMaster(level, arr, n, ...) \
  FT_LOG_IF_(level, fileLogArrayMaster(level, arr, n, __VA_ARGS__))
#define FILE_LOGArraySingle(level, arr, n, ...) \
  FT_LOG_IF_(level, fileLogArraySingle(level, arr, n, __VA_ARGS__))
#define FILE_LOGArrayInt(level, arr, n, ...) \
  FT_LOG_IF_(level, fileLogArrayInt(level, arr, n, __VA_ARGS__))
#define FILE_LOGArrayIntMaster(level, arr, n, ...) \
  FT_LOG_IF_(level, fileLogArrayIntMaster(level, arr, n, __VA_ARGS__))
#define FILE_LOGArrayIntSingle(level, arr, n, ...) \
  FT_LOG_IF_(level, fileLogArrayIntSingle(level, arr, n, __VA_ARGS__))

/* ---- constants (GlobalVariables.h:3-16, FemTech.h:17-18) ----------------- */
static int MAXMATPARAMS = 9;
static int MAXINTERNALVARS = 5;
static int MAXPLOTSTEPS = 1000;
const double huge = 1e20;
#define MAX_FILE_LINE 260
#define MAX_ELEMENT_TYPE_SIZE 10

/* ---- globals (GlobalVariables.h:18-127) ----------------------------------
 * Defined by the library unless marked "driver". */
extern std::string uid;
extern int nparts, nelements, nallelements, nNodes, ndim, nDOF;
extern int nPID, nPIDglobal;
extern int world_rank, world_size;
extern double ExplicitTimeStepReduction; /* driver */
extern double FailureTimeStep;           /* driver */
extern bool ImplicitStatic;              /* driver */
extern bool ImplicitDynamic;             /* driver */
extern bool ExplicitDynamic;             /* driver */
extern double Time;                      /* driver */
extern double dt;                        /* driver */

extern double *coordinates;
extern int *connectivity;
extern int *globalNodeID;
extern int *GaussPoints, *gptr, *dsptr, *gpPtr, *fptr, *pk2ptr, *nShapeFunctions;
extern double *shp, *dshp;
extern int *eptr, *detFptr, *pid, *global_eid;
extern double *mass, *stiffness, *rhs;
extern char **ElementType;
extern double *C, *gaussWeights, *detJacobian;
extern int nSpecifiedDispBC;
extern double *displacements;
extern int *boundary;
extern double *velocities, *velocities_half, *accelerations;
extern double *fe, *fe_prev, *fi, *fi_prev, *f_net, *fr_prev, *fr_curr;
extern double *f_damp_prev, *f_damp_curr, *displacements_prev, *accelerations_prev;
extern double *F, *detF, *invF, *pk2, *Eavg;
extern int *materialID;
extern double *properties;
extern double *internals;
extern int *InternalsPtr;
extern double *stepTime;
extern int sendProcessCount;
extern int *sendProcessID, *sendNeighbourCount, *sendNeighbourCountCum, *sendNodeIndex;
extern double *sendNodeDisplacement, *recvNodeDisplacement;
extern FILE *energyFile;
extern double *mat1, *mat2, *mat3, *mat4;
extern double *fintGQ, *B;
extern double *Hn_1, *Hn_2, *S0n;

/* ---- the explicit path ---------------------------------------------------- */
void InitFemTechWoInput(int argc, char **argv);   /* InitFinalizeFemTech.cpp:47 */
void FinalizeFemTech();                           /* InitFinalizeFemTech.cpp:73 */
void TerminateFemTech(int errorCode);             /* InitFinalizeFemTech.cpp:80 */
void ReadInputFile(const char *FileName);         /* ReadInputFile.cpp:26 */
void ReadMaterials();                             /* ReadMaterials.cpp:8 */
void PartitionMesh();                             /* PartitionMesh.cpp:20 */
void AllocateArrays();                            /* AllocateArrays.cpp:29 */
void ShapeFunctions();                            /* ShapeFunctions.cpp:32 */
void GaussQuadrature3D(int element, int nGaussPoint, double *Chi, double *GaussWeights);
void AssembleLumpedMass(void);                    /* Mass3D.cpp:127 */
double StableTimeStep();                          /* StableTimeStep.cpp:4 */
void GetForce();                                  /* GetForce.cpp:12 */
void GetForce_3D();                               /* GetForce_3D.cpp:5 */
void CalculateAccelerations();                    /* CalculateAcclerations.cpp:4 */
void CheckEnergy(double time, int writeFlag);     /* CheckEnergy.cpp:3 */
void CalculateStrain();                           /* CalculateStrain.cpp:77 */
void CalculateMaximumPrincipalStrain(int elm, double *currentStrainMax,
                                     double *currentStrainMin, double *currentShearMax);
void WriteVTU(const char *FileName, int step, int **intCellData = NULL, int cellDataCount = 0,
              const char **cellDataNames = NULL); /* WriteVTU.cpp:3 */
void WritePVD(const char *FileName, int step);    /* WritePVD.cpp:6 */
void FreeArrays();                                /* FreeArrays.cpp:4 */

/* ---- per-element, per-Gauss-point routines (FemTech.h:29-104) ------------- */
void ShapeFunction_C3D8(int e, int gp, double *Chi, double *detJ);  /* ShapeFunction_C3D8.cpp:4 */
double CalculateTimeStep(int e);                      /* CalculateTimeStep.cpp:7 */
void CalculateDeformationGradient(int e, int gp);     /* CalculateDeformationGradient.cpp:11 */
void StrainDisplacementMatrix(int e, int gp, int nI, double *B);  /* StrainDisplacementMatrix.cpp:3 */
void StressUpdate(int e, int gp);                     /* StressUpdate.cpp:5 */
void DeterminateF(int e, int gp);                     /* DeterminateF.cpp:38 */
void InverseF(int e, int gp, double *fInv);           /* InverseF.cpp:4 */
void InternalForceUpdate(int e, int gp, double *force);  /* InternalForceUpdate.cpp:4 */
void StVenantKirchhoff(int e, int gp);                /* StVenantKirchhoff.cpp:14 */
void CompressibleNeoHookean(int e, int gp);           /* CompressibleNeoHookean.cpp:14 */
void LinearElastic(int e, int gp);                    /* LinearElastic.cpp:14 */
void HGOIsotropic(int e, int gp);                     /* HGOIsotropic.cpp:21 */
void HGOIsotropicViscoelastic(int e, int gp);         /* HGOIsotropicViscoelastic.cpp:27 */
void inverse3x3Matrix(double *mat, double *invMat, double *det);  /* math.cpp:8 */
double CalculateCharacteristicLength(int e);          /* CalculateCharacteristicLength.cpp:3 */
double CalculateCharacteristicLength_C3D8(int e);     /* CalculateCharacteristicLength_C3D8.cpp:3 */

/* ---- outside the engine (link compatibility only; terminate with 3) ------ */
void CreateLinearElasticityCMatrix();
void Assembly(char *operation);
void ApplySteadyBoundaryConditions(void);
void SolveSteadyImplicit(void);
void SolveUnsteadyNewmarkImplicit(double beta, double gamma, double dt, double timeFinal,
                                  char *name);

#endif
