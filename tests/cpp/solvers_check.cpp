// solvers_check.cpp -- calls VPFFT::Solvers (include/vpfft_solvers.hpp)
// the way the reference's MaterialGrid calls its own per voxel
// (Src/MaterialGrid.cpp:107-124 for the Schmid tensors, :946-1003 and
// :1098-1157 for the solvers): one orientation from Euler angles, its 12
// Schmid and anti-Schmid tensors in the sample frame, then every solver in
// turn on the previous one's result.  Prints each result with 17 significant
// digits for the Python test to compare with the oracle.
//
// usage: solvers_check <phi> <theta> <psi> <rate_sensitivity>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "vpfft_solvers.hpp"

using VPFFT::Float;
using VPFFT::FCC_CrystalTest::FCC_SchmidtBasis;
using VPFFT::LinearAlgebra::EigenRep;
using VPFFT::LinearAlgebra::SMatrix3x3;
using VPFFT::LinearAlgebra::SMatrix5x5;
namespace Solvers = VPFFT::Solvers;

static void print(const char* tag, const Float* v, int n) {
  std::printf("%s", tag);
  for (int i = 0; i < n; i++) std::printf(" %.17g", v[i]);
  std::printf("\n");
}

int main(int argc, char** argv) {
  if (argc != 5) {
    std::fprintf(stderr, "usage: %s phi theta psi rate_sensitivity\n", argv[0]);
    return 2;
  }
  SMatrix3x3 R;
  R.BuildActiveEulerMatrix(std::atof(argv[1]), std::atof(argv[2]), std::atof(argv[3]));
  SMatrix3x3 RT = R;
  RT.Transpose();
  const FCC_SchmidtBasis& fcc = FCC_SchmidtBasis::Get();
  std::vector<EigenRep> schmid;
  std::vector<SMatrix3x3> anti;
  for (int k = 0; k < FCC_SchmidtBasis::NumSlipSystems(); k++) {
    schmid.push_back(EigenRep(R * fcc(k) * RT));
    anti.push_back(R * fcc.Anti(k) * RT);
  }
  std::vector<Float> crss(12), gdot0(12), tau0(12, 1.0), tau1(12, 0.8), theta0(12, 20.0), theta1(12, 2.0);
  std::vector<int> nexp(12, std::atoi(argv[4]));
  for (int k = 0; k < 12; k++) {
    crss[k] = 1.0 + 0.02 * k;
    gdot0[k] = 1.0 + 0.01 * (k % 3);
  }
  std::vector<std::vector<Float> > H(12, std::vector<Float>(12, 1.0));
  for (int k = 0; k < 12; k++) H[k][k] = 1.4;

  SMatrix3x3 E;
  E.SetZero();
  E.m[0][0] = 1.0;
  E.m[1][1] = -0.5;
  E.m[2][2] = -0.5;
  E.m[0][1] = E.m[1][0] = 0.2;
  const EigenRep edot(E);
  SMatrix5x5 L0;
  L0.SetIdentity();
  L0 *= 3.0;
  L0(0, 1) = 0.4;  // not symmetric: used as given
  L0(3, 2) = -0.2;

  try {
    print("orientation", &R.m[0][0], 9);
    for (int k = 0; k < 12; k++) print("schmid", schmid[k].m, 5);
    const EigenRep sachs = Solvers::SolveSachEquation(schmid, crss, edot);
    print("sachs", sachs.m, 5);
    const EigenRep taylor = Solvers::SolveConstitutiveEquations(sachs, schmid, crss, gdot0, nexp, edot, 5e-4, 2.0,
                                                                1000);
    print("taylor", taylor.m, 5);
    const EigenRep dtilde(0.01, -0.02, 0.005, 0.0, 0.015);
    Float err = 99.0;
    const EigenRep sigma = Solvers::SolveConstrainedConstitutiveEquations(
        taylor, schmid, crss, gdot0, nexp, edot, taylor, dtilde, L0, 5e-4, 1e9, 100, &err);
    print("constrained", sigma.m, 5);
    print("nr_error", &err, 1);
    const EigenRep rate = Solvers::ApplyConstitutiveEquations(sigma, schmid, crss, gdot0, nexp);
    print("apply", rate.m, 5);
    const SMatrix3x3 W = Solvers::CalculateShearRate(sigma, schmid, anti, crss, gdot0, nexp);
    print("shear_rate", &W.m[0][0], 9);
    Float acc = 0.1;
    const std::vector<Float> tau = Solvers::CalculateHardening(acc, 0.005, sigma, schmid, crss, gdot0, nexp, H, tau0,
                                                               tau1, theta0, theta1);
    print("hardening", tau.data(), 12);
    print("accumulated", &acc, 1);
    // the reference's own shape errors are silent; here they throw
    try {
      std::vector<EigenRep> short_list(schmid.begin(), schmid.begin() + 5);
      Solvers::SolveSachEquation(short_list, crss, edot);
      std::printf("shape_error 0\n");
    } catch (const VPFFT::Error& e) {
      std::printf("shape_error %d\n", e.code());
    }
  } catch (const VPFFT::Error& e) {
    std::fprintf(stderr, "VPFFT::Error %d: %s\n", e.code(), e.what());
    return 1;
  }
  return 0;
}
