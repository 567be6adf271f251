// Test driver over FemTech's C++ surface (libFemTech.so): the reference's
// per-element, per-Gauss-point routines called one at a time the way
// GetForce_3D.cpp:15-44 calls them, against the library's fused GetForce_3D.
// Test code only.
//
//   per_element_driver MESH MODE DT U1 [U2]
//     manual  one pass of CalculateDeformationGradient, DeterminateF,
//             StressUpdate and InternalForceUpdate over every element and
//             Gauss point at U1, assembling fi on the host; then
//             CalculateTimeStep / CalculateCharacteristicLength per element,
//             InverseF, StrainDisplacementMatrix and inverse3x3Matrix of
//             Gauss point 0 of every element
//     fused   GetForce_3D() at U1
//     mixed   the manual pass at U1, then GetForce_3D() at U2 (the Prony
//             history the host pass wrote continues on the device)
//     twice   the manual pass at U1, then again at U2
// Outputs (binary FP64, in the working directory): shp, dshp, detJ, weights,
// F, detF, pk2, fi, hist (Hn_1, Hn_2, S0n concatenated, if allocated) and,
// for manual, dte, le, invF0, B0, inv0 (inverse and determinant of F0).
#include "FemTech.h"

#include <string>
#include <vector>

double Time, dt;
double ExplicitTimeStepReduction = 0.8;
double FailureTimeStep = 1e-11;
bool ImplicitStatic = false;
bool ImplicitDynamic = false;
bool ExplicitDynamic = true;

namespace {

void dump(const char *name, const double *p, size_t n) {
  FILE *f = fopen(name, "wb");
  if (!f) exit(4);
  fwrite(p, sizeof(double), n, f);
  fclose(f);
}

void load(const char *path, double *p, size_t n) {
  FILE *f = fopen(path, "rb");
  if (!f || fread(p, sizeof(double), n, f) != n) exit(5);
  fclose(f);
}

void manual_pass() {
  memset(fi, 0, nDOF * sizeof(double));
  for (int e = 0; e < nelements; ++e) {
    double local[24] = {0};
    for (int gp = 0; gp < GaussPoints[e]; ++gp) {
      CalculateDeformationGradient(e, gp);
      DeterminateF(e, gp);
      StressUpdate(e, gp);
      InternalForceUpdate(e, gp, local);
    }
    for (int k = 0; k < 8; ++k)
      for (int d = 0; d < 3; ++d) fi[3 * connectivity[eptr[e] + k] + d] += local[3 * k + d];
  }
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 5) return 2;
  const std::string mode = argv[2];
  InitFemTechWoInput(argc, argv);
  ReadInputFile(argv[1]);
  ReadMaterials();
  PartitionMesh();
  AllocateArrays();
  ShapeFunctions();
  dt = atof(argv[3]);
  const size_t ne = nelements;
  dump("shp.f64", shp, 64 * ne);
  dump("dshp.f64", dshp, 192 * ne);
  dump("detJ.f64", detJacobian, 8 * ne);
  dump("weights.f64", gaussWeights, 8 * ne);
  load(argv[4], displacements, nDOF);
  if (mode == "fused") {
    GetForce_3D();
  } else {
    manual_pass();
    if (mode == "mixed" || mode == "twice") {
      if (argc < 6) return 2;
      load(argv[5], displacements, nDOF);
      if (mode == "mixed")
        GetForce_3D();
      else
        manual_pass();
    }
  }
  dump("F.f64", F, 72 * ne);
  dump("detF.f64", detF, 8 * ne);
  dump("pk2.f64", pk2, 48 * ne);
  dump("fi.f64", fi, nDOF);
  if (Hn_1) {
    std::vector<double> h;
    h.insert(h.end(), Hn_1, Hn_1 + 72 * ne);
    h.insert(h.end(), Hn_2, Hn_2 + 72 * ne);
    h.insert(h.end(), S0n, S0n + 72 * ne);
    dump("hist.f64", h.data(), h.size());
  }
  if (mode == "manual") {
    std::vector<double> dte(ne), le(ne), le8(ne), invf(9 * ne), b0(18 * ne), inv0(10 * ne);
    for (size_t e = 0; e < ne; ++e) {
      dte[e] = CalculateTimeStep((int)e);
      le[e] = CalculateCharacteristicLength((int)e);
      le8[e] = CalculateCharacteristicLength_C3D8((int)e);
      InverseF((int)e, 0, &invf[9 * e]);
      StrainDisplacementMatrix((int)e, 0, 0, &b0[18 * e]);
      double det = 0.0;
      inverse3x3Matrix(F + fptr[e], &inv0[10 * e], &det);
      inv0[10 * e + 9] = det;
    }
    dump("dte.f64", dte.data(), ne);
    dump("le.f64", le.data(), ne);
    dump("le8.f64", le8.data(), ne);
    dump("invF0.f64", invf.data(), 9 * ne);
    dump("B0.f64", b0.data(), 18 * ne);
    dump("inv0.f64", inv0.data(), 10 * ne);
  }
  FinalizeFemTech();
  return 0;
}
