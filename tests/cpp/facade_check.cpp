// facade_check.cpp -- drives VPFFT::DataStructures::MaterialGrid
// (include/vpfft_materialgrid.hpp) the way the reference's MPI_VPFFT_Test
// drives its own class (Src/UnitTester.cpp:384-457), with the same calls on
// the grid and on the value types: FCC_SchmidtBasis::Get, the three-integer
// constructor (the reference's material), SendDataAndInitialize,
// SetVoxelLength, a strain rate built with SMatrix3x3::SetZero and
// EigenRep(const SMatrix3x3&), consecutive strain steps.  The orientations
// are a seeded polycrystal.  Prints one line per result with 17 significant
// digits for the Python test to compare with the same run through the Python
// mirror.
//
// usage: facade_check <edge> <grains> <seed> <steps>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "vpfft_materialgrid.hpp"

using VPFFT::Error;
using VPFFT::DataStructures::MaterialGrid;
using VPFFT::FCC_CrystalTest::FCC_SchmidtBasis;
using VPFFT::LinearAlgebra::EigenRep;
using VPFFT::LinearAlgebra::SMatrix3x3;

static void print5(const char* tag, const double* v) {
  std::printf("%s %.17g %.17g %.17g %.17g %.17g\n", tag, v[0], v[1], v[2], v[3], v[4]);
}

int main(int argc, char** argv) {
  if (argc != 5) {
    std::fprintf(stderr, "usage: %s edge grains seed steps\n", argv[0]);
    return 2;
  }
  const int e = std::atoi(argv[1]), grains = std::atoi(argv[2]), steps = std::atoi(argv[4]);
  const unsigned long long seed = std::strtoull(argv[3], nullptr, 10);
  const long n = (long)e * e * e;
  std::vector<double> R(9 * n);
  if (vpfft_generate_voronoi(e, e, e, grains, seed, R.data(), nullptr, nullptr)) {
    std::fprintf(stderr, "%s\n", vpfft_last_error());
    return 1;
  }
  std::vector<SMatrix3x3> orient(n);
  std::memcpy(orient.data(), R.data(), sizeof(double) * 9 * n);
  try {
    const FCC_SchmidtBasis& fcc = FCC_SchmidtBasis::Get();
    std::printf("slip_systems %d %.17g\n", FCC_SchmidtBasis::NumSlipSystems(), fcc(0).m[0][1]);
    MaterialGrid grid(e, e, e);  // the reference's hard-coded material
    grid.SendDataAndInitialize(orient.data(), nullptr);
    grid.SetVoxelLength(1, 1, 1);
    // uniaxial tension along x
    SMatrix3x3 StrainRateM;
    StrainRateM.SetZero();
    StrainRateM.m[0][0] = 1;
    StrainRateM.m[1][1] = -0.5;
    StrainRateM.m[2][2] = -0.5;
    const EigenRep edot(StrainRateM);
    print5("edot", edot.m);
    for (int s = 0; s < steps; s++) {
      const EigenRep avg = grid.RunSingleStrainStep(edot, 100, 0.005, 1e-5);
      std::printf("iterations %d\n", grid.LastStepInfo().iterations);
      print5("avg", avg.m);
    }
    print5("stress", grid.Stress(1, 2, 3).m);
    print5("lambda", grid.LagrangeMultiplier(1, 2, 3).m);
    print5("schmid", grid.SchmidtTensor(1, 2, 3, 7).m);
    const SMatrix3x3& r = grid.Orientation(1, 2, 3);
    std::printf("orientation %.17g %.17g %.17g\n", r.m[0][0], r.m[1][1], r.m[2][2]);
    // a write through the non-const accessor reaches the device before the
    // next step: set one voxel's orientation to the identity and read back
    SMatrix3x3& w = grid.Orientation(0, 0, 0);
    std::memset(&w, 0, sizeof w);
    w.m[0][0] = w.m[1][1] = w.m[2][2] = 1.0;
    grid.RunSingleStrainStep(edot, 1, 0.0, 1e-5);  // dt = 0: the epilogue leaves R as uploaded
    const SMatrix3x3& back = grid.Orientation(0, 0, 0);
    std::printf("roundtrip %d\n", back.m[0][0] == 1.0 && back.m[0][1] == 0.0 && back.m[2][2] == 1.0);
    // the reference exits on degenerate loading; the class throws
    try {
      grid.RunSingleStrainStep(EigenRep(), 10, 0.005, 1e-5);
      std::printf("degenerate none\n");
    } catch (const Error& err) {
      std::printf("degenerate %d\n", err.code());
    }
  } catch (const Error& err) {
    std::fprintf(stderr, "error %d: %s\n", err.code(), err.what());
    return 1;
  }
  return 0;
}
