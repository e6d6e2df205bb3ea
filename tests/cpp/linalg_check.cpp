// linalg_check.cpp -- exercises the host value types of
// include/vpfft_linalg.hpp (VPFFT::LinearAlgebra, Src/LinearAlgebra.h) for
// the Python test: conversions between matrices and EigenRep, the Euler
// matrix, products, MovingAverage and NormalizeValues, printed with 17
// significant digits.
//
// usage: linalg_check <phi> <theta> <psi>
#include <cstdio>
#include <cstdlib>

#include "vpfft_linalg.hpp"

using VPFFT::Float;
using VPFFT::FCC_CrystalTest::FCC_SchmidtBasis;
using namespace VPFFT::LinearAlgebra;

static void print(const char* tag, const Float* v, int n) {
  std::printf("%s", tag);
  for (int i = 0; i < n; i++) std::printf(" %.17g", v[i]);
  std::printf("\n");
}

int main(int argc, char** argv) {
  if (argc != 4) return 2;
  SMatrix3x3 R;
  R.BuildActiveEulerMatrix(std::atof(argv[1]), std::atof(argv[2]), std::atof(argv[3]));
  print("R", &R.m[0][0], 9);
  Float M0[3][3] = {{0.3, -1.2, 0.7}, {0.4, 2.0, -0.1}, {1.1, 0.25, -0.9}};  // neither symmetric nor traceless
  const SMatrix3x3 M(M0);
  const EigenRep v(M);
  print("eigenrep", v.m, 5);
  const SMatrix3x3 back = v.ToMatrixRep();
  print("matrix", &back.m[0][0], 9);
  const SMatrix3x3 RMRt = [&] {
    SMatrix3x3 RT = R;
    RT.Transpose();
    return R * M * RT;
  }();
  print("rotated", &RMRt.m[0][0], 9);
  const EigenRep w(0.5, -0.25, 1.5, 2.0, -1.0);
  const Float inner = InnerProduct(v, w);
  print("inner", &inner, 1);
  const EigenRep sym = v.Symmetrize(), antisym = v.AntiSymmetrize();
  print("sym", sym.m, 5);
  print("antisym", antisym.m, 5);
  const Float nrm = w.Norm();
  print("norm", &nrm, 1);
  SMatrix5x5 A = OuterProduct(v, w);
  A(2, 2) += 4.0;
  print("outer", A.m, 25);
  const EigenRep Av = A * w, vA = w * A;
  print("Av", Av.m, 5);
  print("vA", vA.m, 5);
  SMatrix5x5 B;
  B.SetIdentity();
  B(0, 4) = 1e-12;
  B(1, 3) = 3.0;
  const SMatrix5x5 AB = A * B;
  print("AB", AB.m, 25);
  SMatrix5x5 avg;
  avg.SetZero();
  avg.MovingAverage(A, 0);
  avg.MovingAverage(B, 1);
  print("moving_average", avg.m, 25);
  B.NormalizeValues();
  print("normalized", B.m, 25);
  for (int k = 0; k < 12; k++) print("schmid_basis", &FCC_SchmidtBasis::Get()(k).m[0][0], 9);
  for (int k = 0; k < 12; k++) print("anti_basis", &FCC_SchmidtBasis::Get().Anti(k).m[0][0], 9);
  return 0;
}
