// device_common.h -- per-voxel crystal plasticity on gfx950, FP64 in registers.
//
// One voxel per lane.  Everything here is a __device__ restatement of the
// reference's per-voxel solvers (Src/Solvers.cpp) written for registers:
//   * Schmid tensors are rebuilt from the voxel's orientation R instead of
//     being stored (MaterialGrid.cpp:107-124 keeps 12x5 doubles per voxel);
//   * the 5x5 Newton system J' d = R of SolveConstrainedConstitutiveEquations
//     (Solvers.cpp:202-210, J' = -I - L0 J) is solved in the equivalent
//     symmetric form (L0^-1 + J) d = -(L0^-1 R), where L0^-1 + J is symmetric
//     positive definite (L0 is an average of SPD secant moduli and J a sum of
//     c_k m_k m_k^T with c_k >= 0), so an unpivoted LDL^T replaces Eigen's
//     full-pivot LU; the Taylor system (Solvers.cpp:77-89) is -J d = R with the
//     same SPD form.  Both agree with the literal solve to rounding level.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vpfft {

constexpr int kSlips = 12;

// Material + per-step constants, uploaded once per strain step.
struct StepConst {
  double L0[25];      // SMatrix5x5 column-major, symmetrised
  double L0inv[25];   // inverse of L0 (column-major, symmetric)
  double L0inv_w[15]; // W lower triangular with L0inv = W W^T, packed by rows (W[sidx(i, j)], i >= j)
  double edot[5];     // macroscopic strain rate (EigenRep)
  double gdot0[kSlips];
  int nexp[kSlips];
  int uniform_n;      // 1 if all nexp equal (then nexp[0] is used for all slips)
  double imposed_max_rss;  // Taylor admissibility bound, MaterialGrid.cpp:162-166
  double inv_n;            // 1/N
  double dt;
  double tau0[kSlips], tau1[kSlips], theta0[kSlips], theta1[kSlips];
  double hmat[kSlips * kSlips];
  // Green operator: K(k)_ik = sum over monomials k_j k_l of Kc[ik][jl]
  double Kc[6][6];
  double kscale[3];   // 1/(N_a * voxel length_a)
  int nx, ny, nz, nzh;
};

// Read-only view of per-step constants: loads through the constant address
// space are invariant for the kernel, so uniform ones become scalar loads
// the compiler may repeat instead of holding the values in registers.  The
// pointer is rebuilt from its integer value: a generic pointer cast to the
// constant space and straight back is folded back to generic.
template <typename T>
__device__ __forceinline__ const T* as_const_mem(const T* p) {
  return (const T*)(const __attribute__((address_space(4))) T*)(unsigned long)p;
}

// FCC {111}<110> systems in the reference's order (BurgerVectorBasis.h:16-72).
__device__ __constant__ static const signed char kB[kSlips][3] = {
    {0, 1, 1}, {1, 0, 1}, {1, -1, 0}, {0, 1, -1}, {1, 0, 1}, {1, 1, 0},
    {0, 1, 1}, {1, 0, -1}, {1, 1, 0}, {0, 1, -1}, {1, 0, -1}, {1, -1, 0}};
__device__ __constant__ static const signed char kN[kSlips][3] = {
    {1, 1, -1}, {1, 1, -1}, {1, 1, -1}, {1, -1, -1}, {1, -1, -1}, {1, -1, -1},
    {1, -1, 1}, {1, -1, 1}, {1, -1, 1}, {1, 1, 1},   {1, 1, 1},   {1, 1, 1}};

// Deviatoric-basis projection of the symmetric matrix M (LinearAlgebra.h:207-247,
// EigenRep(const SMatrix3x3&) :307-312).
__device__ __forceinline__ void proj5(double M00, double M11, double M22, double M12,
                                      double M02, double M01, double v[5]) {
  const double s2 = 0.70710678118654752440;   // 1/sqrt(2)
  const double s6 = 0.40824829046386301637;   // 1/sqrt(6)
  v[0] = s2 * (M11 - M00);
  v[1] = s6 * (2.0 * M22 - M00 - M11);
  v[2] = 2.0 * s2 * M12;
  v[3] = 2.0 * s2 * M02;
  v[4] = 2.0 * s2 * M01;
}

// m_k = proj5(R m0_k R^T), m0_k = (b n^T + n b^T)/(2 sqrt 6)
// (MaterialGrid.cpp:115-122, BurgerVectorBasis.h:122-132).  With u = R b and
// w = R n, R m0 R^T = (u w^T + w u^T)/(2 sqrt 6).
__device__ __forceinline__ void schmid_from_R(const double R[9], double m[kSlips][5]) {
  const double c = 0.20412414523193150819;  // 1/(2 sqrt 6)
#pragma unroll
  for (int k = 0; k < kSlips; k++) {
    double u[3], w[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
      u[i] = kB[k][0] * R[3 * i + 0] + kB[k][1] * R[3 * i + 1] + kB[k][2] * R[3 * i + 2];
      w[i] = kN[k][0] * R[3 * i + 0] + kN[k][1] * R[3 * i + 1] + kN[k][2] * R[3 * i + 2];
    }
    const double M00 = 2.0 * u[0] * w[0] * c, M11 = 2.0 * u[1] * w[1] * c,
                 M22 = 2.0 * u[2] * w[2] * c;
    const double M12 = (u[1] * w[2] + w[1] * u[2]) * c, M02 = (u[0] * w[2] + w[0] * u[2]) * c,
                 M01 = (u[0] * w[1] + w[0] * u[1]) * c;
    proj5(M00, M11, M22, M12, M02, M01, m[k]);
  }
}

// Antisymmetric part: R a0_k R^T, a0_k = (b n^T - n b^T)/(2 sqrt 6)
// (BurgerVectorBasis.h:137-150, MaterialGrid.cpp:1105-1111).
__device__ __forceinline__ void anti_from_R(const double R[9], int k, double W[3]) {
  // returns the three independent entries (01, 02, 12) of the skew matrix
  const double c = 0.20412414523193150819;
  double u[3], w[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    u[i] = kB[k][0] * R[3 * i + 0] + kB[k][1] * R[3 * i + 1] + kB[k][2] * R[3 * i + 2];
    w[i] = kN[k][0] * R[3 * i + 0] + kN[k][1] * R[3 * i + 1] + kN[k][2] * R[3 * i + 2];
  }
  W[0] = (u[0] * w[1] - w[0] * u[1]) * c;
  W[1] = (u[0] * w[2] - w[0] * u[2]) * c;
  W[2] = (u[1] * w[2] - w[1] * u[2]) * c;
}

__device__ __forceinline__ double dot5(const double a[5], const double b[5]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3] + a[4] * b[4];
}

// x^e for integer e >= 0 by binary powering (stands in for std::pow(double,int),
// Solvers.cpp:67,178,268; differs from a correctly rounded pow by a few ulp).
__device__ __forceinline__ double ipow(double x, int e) {
  double r = 1.0;
  while (e > 0) {
    if (e & 1) r *= x;
    x *= x;
    e >>= 1;
  }
  return r;
}

// Symmetric 5x5 stored as 15 entries, packed lower triangle by rows:
// (0,0) (1,0) (1,1) (2,0) (2,1) (2,2) ...
__device__ __forceinline__ constexpr int sidx(int i, int j) {
  return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i;
}

// Solve A x = b for symmetric positive (semi)definite A by LDL^T without
// pivoting.  A zero/negative pivot marks a direction outside the range: that
// component of the solution is set to 0 (the rank-deficient behaviour of a
// full-pivot LU solve).
__device__ __forceinline__ void ldlt_solve5(double A[15], const double b[5], double x[5]) {
  double L[15];
  double dinv[5];
  double D[5];
#pragma unroll
  for (int j = 0; j < 5; j++) {
    double d = A[sidx(j, j)];
#pragma unroll
    for (int k = 0; k < j; k++) d -= L[sidx(j, k)] * L[sidx(j, k)] * D[k];
    const double scale = fabs(A[sidx(j, j)]);
    const bool ok = d > 1e-300 && d > 1e-14 * scale;
    D[j] = ok ? d : 0.0;
    dinv[j] = ok ? 1.0 / d : 0.0;
    L[sidx(j, j)] = 1.0;
#pragma unroll
    for (int i = j + 1; i < 5; i++) {
      double s = A[sidx(i, j)];
#pragma unroll
      for (int k = 0; k < j; k++) s -= L[sidx(i, k)] * L[sidx(j, k)] * D[k];
      L[sidx(i, j)] = s * dinv[j];
    }
  }
  double y[5];
#pragma unroll
  for (int i = 0; i < 5; i++) {
    double s = b[i];
#pragma unroll
    for (int k = 0; k < i; k++) s -= L[sidx(i, k)] * y[k];
    y[i] = s;
  }
#pragma unroll
  for (int i = 4; i >= 0; i--) {
    double s = y[i] * dinv[i];
#pragma unroll
    for (int k = i + 1; k < 5; k++) s -= L[sidx(k, i)] * x[k];
    x[i] = s;
  }
}

// 1/d within an ulp: the hardware reciprocal estimate refined by two Newton
// steps (5 instructions against the 10 of an IEEE division).  For the Newton
// kernel's own solve, where a last-bit difference only moves the iterate.
__device__ __forceinline__ double rcp_nr(double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = __builtin_fma(-d, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-d, r, 1.0);
  return __builtin_fma(r, e, r);
}

// In-place variant: A's packed lower triangle is overwritten by L (unit
// diagonal implied), the solution replaces b.  The factorisation of
// ldlt_solve5 with the same pivot rule, but each column is formed from the
// unscaled entries U = L D of the rows above it (one FMA per term instead of
// two products), and the pivots' reciprocals come from rcp_nr.
__device__ __forceinline__ void ldlt_solve5_inplace(double A[15], double b[5]) {
  double dinv[5], U[15];
#pragma unroll
  for (int j = 0; j < 5; j++) {
    double d = A[sidx(j, j)];
#pragma unroll
    for (int k = 0; k < j; k++) d = __builtin_fma(-A[sidx(j, k)], U[sidx(j, k)], d);
    const double scale = fabs(A[sidx(j, j)]);
    const bool ok = d > 1e-300 && d > 1e-14 * scale;
    const double dj = ok ? d : 0.0;
    dinv[j] = ok ? rcp_nr(d) : 0.0;
#pragma unroll
    for (int i = j + 1; i < 5; i++) {
      double t = A[sidx(i, j)];
#pragma unroll
      for (int k = 0; k < j; k++) t = __builtin_fma(-A[sidx(i, k)], U[sidx(j, k)], t);
      U[sidx(i, j)] = t;
      A[sidx(i, j)] = t * dinv[j];
    }
    A[sidx(j, j)] = dj;  // D stored on the diagonal
  }
#pragma unroll
  for (int i = 0; i < 5; i++)
#pragma unroll
    for (int k = 0; k < i; k++) b[i] -= A[sidx(i, k)] * b[k];
#pragma unroll
  for (int i = 4; i >= 0; i--) {
    double t = b[i] * dinv[i];
#pragma unroll
    for (int k = i + 1; k < 5; k++) t -= A[sidx(k, i)] * b[k];
    b[i] = t;
  }
}

__device__ __forceinline__ void matvec5(const double* __restrict__ Lcm, const double v[5],
                                        double out[5]) {
  // SMatrix5x5 * EigenRep, LinearAlgebra.cpp:1120-1127 (column-major L)
#pragma unroll
  for (int i = 0; i < 5; i++) {
    double r = 0.0;
#pragma unroll
    for (int j = 0; j < 5; j++) r += Lcm[j * 5 + i] * v[j];
    out[i] = r;
  }
}

__device__ __forceinline__ double norm5(const double a[5]) { return sqrt(dot5(a, a)); }

// rate term: |r|^(n-1) * gdot0 (Solvers.cpp:175-181) -- n != 1 branch.
__device__ __forceinline__ double gamma_nm1(const StepConst& C, int k, double absr) {
  const int n = C.uniform_n ? C.nexp[0] : C.nexp[k];
  return n != 1 ? C.gdot0[k] * ipow(absr, n - 1) : C.gdot0[k];
}

// Taylor variant uses (n - 1 > 0) as its test (Solvers.cpp:66).
__device__ __forceinline__ double gamma_nm1_taylor(const StepConst& C, int k, double absr) {
  const int n = C.uniform_n ? C.nexp[0] : C.nexp[k];
  return n - 1 > 0 ? C.gdot0[k] * ipow(absr, n - 1) : C.gdot0[k];
}

// |r|^n * gdot0 for n > 0 (Solvers.cpp:267-270).
__device__ __forceinline__ double gamma_n(const StepConst& C, int k, double absr) {
  const int n = C.uniform_n ? C.nexp[0] : C.nexp[k];
  return n > 0 ? C.gdot0[k] * ipow(absr, n) : C.gdot0[k];
}

__device__ __forceinline__ int nexp_of(const StepConst& C, int k) {
  return C.uniform_n ? C.nexp[0] : C.nexp[k];
}

__device__ __forceinline__ double sgn(double v) {  // Solvers.h:23-25
  return (double)((0.0 < v) - (v < 0.0));
}

// ---------------------------------------------------------------------------
// Crystal-frame Newton solve.
//
// r_k = m_k . sigma with m_k = proj5(R m0_k R^T) equals m0_k . (Q sigma),
// Q_ab = <B_a, R^T B_b R> (any R: trace identities only), and
// m_k = Q^T m0_k.  In that frame the twelve Schmid vectors are constants
// with one zero each, m0_k,i = kSign[k][i] * w_i, w = (1/(2 sqrt 3), 1/2,
// 1/(2 sqrt 3), ...), so the resolved shear stresses, the stress sum and the
// Newton matrix are sums of signed terms.  The orientation is updated
// additively and drifts off SO(3) (MaterialGrid.cpp:870-871), so Q is not
// orthogonal: with P = Q^-1 the sample-frame system (L0^-1 + Q^T J_c Q) d = rhs
// is (P^T L0^-1 P + J_c) d_c = P^T rhs with d = P d_c.  The iterate is kept in
// both frames; the step norm and the back-track use the sample frame, as the
// reference does.
namespace cf {

// FCC {111}<110> Schmid vectors of BurgerVectorBasis.h:16-72, 122-132 in the
// deviatoric basis of LinearAlgebra.h:207-247, divided by w_i.
__device__ __constant__ static const signed char kSign[kSlips][5] = {
    {1, -1, 0, 1, 1},   {-1, -1, 1, 0, 1},  {-2, 0, 1, -1, 0}, {-1, 1, 0, -1, 1},
    {-1, -1, -1, 0, -1}, {-2, 0, -1, -1, 0}, {-1, 1, 0, 1, 1},  {-1, -1, 1, 0, -1},
    {-2, 0, 1, 1, 0},   {1, -1, 0, -1, 1},  {-1, -1, -1, 0, 1}, {-2, 0, -1, 1, 0}};
__device__ __forceinline__ constexpr double w(int i) { return i == 1 ? 0.5 : 0.28867513459481287; }

// proj5 of a symmetric 3x3 given by entries (00,11,22,12,02,01)
__device__ __forceinline__ void proj_sym(double c00, double c11, double c22, double c12, double c02, double c01,
                                         double v[5]) {
  const double s2 = 0.70710678118654752440, s6 = 0.40824829046386301637;
  v[0] = (c11 - c00) * s2;
  v[1] = (2.0 * c22 - c00 - c11) * s6;
  v[2] = 2.0 * c12 * s2;
  v[3] = 2.0 * c02 * s2;
  v[4] = 2.0 * c01 * s2;
}

// Q(M)_ab = <B_a, M^T B_b M> (row-major, Q[5a+b]) for a 3x3 M (row-major):
// column b is proj5(M^T B_b M); with rho_i the rows of M,
// M^T e_i e_j^T M = rho_i rho_j^T.
__device__ __forceinline__ void dev_map(const double M[9], double Q[25]) {
  const double s2 = 0.70710678118654752440, s6 = 0.40824829046386301637;
  const double* r0 = M;
  const double* r1 = M + 3;
  const double* r2 = M + 6;
  auto sq = [](const double* a, int k, int l) { return a[k] * a[l]; };
  auto mix = [](const double* a, const double* b, int k, int l) { return a[k] * b[l] + b[k] * a[l]; };
  double col[5];
  auto put = [&](int bcol) {
#pragma unroll
    for (int a = 0; a < 5; a++) Q[a * 5 + bcol] = col[a];
  };
  proj_sym((sq(r1, 0, 0) - sq(r0, 0, 0)) * s2, (sq(r1, 1, 1) - sq(r0, 1, 1)) * s2,
           (sq(r1, 2, 2) - sq(r0, 2, 2)) * s2, (sq(r1, 1, 2) - sq(r0, 1, 2)) * s2,
           (sq(r1, 0, 2) - sq(r0, 0, 2)) * s2, (sq(r1, 0, 1) - sq(r0, 0, 1)) * s2, col);
  put(0);
  proj_sym((2.0 * sq(r2, 0, 0) - sq(r0, 0, 0) - sq(r1, 0, 0)) * s6,
           (2.0 * sq(r2, 1, 1) - sq(r0, 1, 1) - sq(r1, 1, 1)) * s6,
           (2.0 * sq(r2, 2, 2) - sq(r0, 2, 2) - sq(r1, 2, 2)) * s6,
           (2.0 * sq(r2, 1, 2) - sq(r0, 1, 2) - sq(r1, 1, 2)) * s6,
           (2.0 * sq(r2, 0, 2) - sq(r0, 0, 2) - sq(r1, 0, 2)) * s6,
           (2.0 * sq(r2, 0, 1) - sq(r0, 0, 1) - sq(r1, 0, 1)) * s6, col);
  put(1);
  proj_sym(mix(r1, r2, 0, 0) * s2, mix(r1, r2, 1, 1) * s2, mix(r1, r2, 2, 2) * s2, mix(r1, r2, 1, 2) * s2,
           mix(r1, r2, 0, 2) * s2, mix(r1, r2, 0, 1) * s2, col);
  put(2);
  proj_sym(mix(r0, r2, 0, 0) * s2, mix(r0, r2, 1, 1) * s2, mix(r0, r2, 2, 2) * s2, mix(r0, r2, 1, 2) * s2,
           mix(r0, r2, 0, 2) * s2, mix(r0, r2, 0, 1) * s2, col);
  put(3);
  proj_sym(mix(r0, r1, 0, 0) * s2, mix(r0, r1, 1, 1) * s2, mix(r0, r1, 2, 2) * s2, mix(r0, r1, 1, 2) * s2,
           mix(r0, r1, 0, 2) * s2, mix(r0, r1, 0, 1) * s2, col);
  put(4);
}

// P = Q(R)^-1.  Q(R) maps X to dev(R^T X R) on deviatoric symmetric tensors;
// its inverse maps Y to R^-T Y R^-1 + alpha H, H = R^-T R^-1, with alpha
// making the result traceless: P = Q(R^-1) - h k^T / tr H, h = proj5(H),
// k = proj5(R^-1 R^-T).  For a rotation P = Q^T.
__device__ __forceinline__ void inverse_map(const double R[9], double P[25]) {
  // R^-1 = adj(R) / det(R)
  const double c00 = R[4] * R[8] - R[5] * R[7], c01 = R[5] * R[6] - R[3] * R[8], c02 = R[3] * R[7] - R[4] * R[6];
  const double det = R[0] * c00 + R[1] * c01 + R[2] * c02;
  const double id = rcp_nr(det);
  double Ri[9];
  Ri[0] = c00 * id;
  Ri[3] = c01 * id;
  Ri[6] = c02 * id;
  Ri[1] = (R[2] * R[7] - R[1] * R[8]) * id;
  Ri[4] = (R[0] * R[8] - R[2] * R[6]) * id;
  Ri[7] = (R[1] * R[6] - R[0] * R[7]) * id;
  Ri[2] = (R[1] * R[5] - R[2] * R[4]) * id;
  Ri[5] = (R[2] * R[3] - R[0] * R[5]) * id;
  Ri[8] = (R[0] * R[4] - R[1] * R[3]) * id;
  dev_map(Ri, P);
  // H = Ri^T Ri (columns of Ri), K = Ri Ri^T (rows of Ri)
  auto hcol = [&](int i, int j) { return Ri[i] * Ri[j] + Ri[3 + i] * Ri[3 + j] + Ri[6 + i] * Ri[6 + j]; };
  auto krow = [&](int i, int j) { return Ri[3 * i] * Ri[3 * j] + Ri[3 * i + 1] * Ri[3 * j + 1] + Ri[3 * i + 2] * Ri[3 * j + 2]; };
  double h[5], k[5];
  const double h00 = hcol(0, 0), h11 = hcol(1, 1), h22 = hcol(2, 2);
  proj_sym(h00, h11, h22, hcol(1, 2), hcol(0, 2), hcol(0, 1), h);
  proj_sym(krow(0, 0), krow(1, 1), krow(2, 2), krow(1, 2), krow(0, 2), krow(0, 1), k);
  const double itr = rcp_nr(h00 + h11 + h22);
#pragma unroll
  for (int a = 0; a < 5; a++) {
    const double ha = h[a] * itr;
#pragma unroll
    for (int bcol = 0; bcol < 5; bcol++) P[a * 5 + bcol] -= ha * k[bcol];
  }
}

// y = Q x  and  y = Q^T x
__device__ __forceinline__ void rot(const double Q[25], const double x[5], double y[5]) {
#pragma unroll
  for (int a = 0; a < 5; a++) {
    double t = 0.0;
#pragma unroll
    for (int b = 0; b < 5; b++) t += Q[a * 5 + b] * x[b];
    y[a] = t;
  }
}
__device__ __forceinline__ void rot_t(const double Q[25], const double x[5], double y[5]) {
#pragma unroll
  for (int b = 0; b < 5; b++) {
    double t = 0.0;
#pragma unroll
    for (int a = 0; a < 5; a++) t += Q[a * 5 + b] * x[a];
    y[b] = t;
  }
}

// P^T A P for A = W W^T (W lower triangular, packed by rows): M = W^T P, then
// M^T M; packed lower result (150 multiply-adds; 200 for P^T (A P)).
__device__ __forceinline__ void congruence_chol(const double P[25], const double* __restrict__ W, double out[15]) {
  double M[25];
#pragma unroll
  for (int i = 0; i < 5; i++)
#pragma unroll
    for (int c = 0; c < 5; c++) {
      double t = 0.0;
#pragma unroll
      for (int j = i; j < 5; j++) t += W[sidx(j, i)] * P[j * 5 + c];
      M[i * 5 + c] = t;
    }
#pragma unroll
  for (int a = 0; a < 5; a++)
#pragma unroll
    for (int c = 0; c <= a; c++) {
      double t = 0.0;
#pragma unroll
      for (int i = 0; i < 5; i++) t += M[i * 5 + a] * M[i * 5 + c];
      out[sidx(a, c)] = t;
    }
}

// m0_k . x, exact sums of signed (and doubled) terms of u = w * x
__device__ __forceinline__ double schmid_dot(int k, const double u[5]) {
  double d = -0.0;  // -0.0 + t == t, so the first term costs nothing
#pragma unroll
  for (int i = 0; i < 5; i++) {
    const int sg = kSign[k][i];
    if (sg == 1) d += u[i];
    if (sg == -1) d -= u[i];
    if (sg == 2) d += 2.0 * u[i];
    if (sg == -2) d -= 2.0 * u[i];
  }
  return d;
}

__device__ __forceinline__ void scaled(const double x[5], double u[5]) {
#pragma unroll
  for (int i = 0; i < 5; i++) u[i] = w(i) * x[i];
}

// x^E by a fixed squaring chain (E known at compile time)
template <int E>
__device__ __forceinline__ double ipow_c(double x) {
  if constexpr (E == 0) {
    return 1.0;
  } else if constexpr (E == 1) {
    return x;
  } else {
    const double h = ipow_c<E / 2>(x);
    return (E & 1) ? h * h * x : h * h;
  }
}

// gdot0 |r|^(n-1) (Solvers.cpp:176-180); NE > 0: every slip has exponent NE.
template <int NE>
__device__ __forceinline__ double gamma_nm1(const StepConst& C, int k, double absr) {
  if constexpr (NE > 0) {
    return NE != 1 ? C.gdot0[k] * ipow_c<NE - 1>(absr) : C.gdot0[k];
  } else {
    return vpfft::gamma_nm1(C, k, absr);
  }
}
template <int NE>
__device__ __forceinline__ double gamma_n(const StepConst& C, int k, double absr) {
  if constexpr (NE > 0) {
    return C.gdot0[k] * ipow_c<NE>(absr);
  } else {
    return vpfft::gamma_n(C, k, absr);
  }
}
template <int NE>
__device__ __forceinline__ double nexp(const StepConst& C, int k) {
  if constexpr (NE > 0) {
    return (double)NE;
  } else {
    return (double)nexp_of(C, k);
  }
}

// Fast Walsh-Hadamard transform of 4 values: the sum and the three signed
// sums (+,+,-,-), (+,-,-,+), (+,-,+,-).
__device__ __forceinline__ void had4(double x0, double x1, double x2, double x3, double h[4]) {
  const double p = x0 + x1, m = x0 - x1, q = x2 + x3, n = x2 - x3;
  h[0] = p + q;
  h[1] = p - q;
  h[2] = m - n;
  h[3] = m + n;
}

// S = sum_k kSign_k g_k and J = sum_k c_k kSign_k kSign_k^T (packed lower),
// grouped by the structure of kSign: the slips k = f, f + 3, f + 6, f + 9 of
// family f (= k mod 3) share their nonzero positions and differ only in
// signs that are the rows of a 4 x 4 Hadamard matrix.  So each family
// reduces to one 4-point Hadamard transform, and J takes ~35 additions
// instead of 120.  (Same values as the direct sums up to rounding.)
__device__ __forceinline__ void schmid_sums(const double g[kSlips], const double c[kSlips], double S[5],
                                            double J[15]) {
  double A[4], B[4], Cc[4], Ga[4], Gb[4], Gc[4];
  had4(c[0], c[3], c[6], c[9], A);
  had4(c[1], c[4], c[7], c[10], B);
  had4(c[2], c[5], c[8], c[11], Cc);
  had4(g[0], g[3], g[6], g[9], Ga);
  had4(g[1], g[4], g[7], g[10], Gb);
  had4(g[2], g[5], g[8], g[11], Gc);
  const double ab = A[0] + B[0];
  J[sidx(0, 0)] = __builtin_fma(4.0, Cc[0], ab);
  J[sidx(1, 0)] = B[0] - A[0];
  J[sidx(1, 1)] = ab;
  J[sidx(2, 0)] = __builtin_fma(-2.0, Cc[3], -B[3]);
  J[sidx(2, 1)] = -B[3];
  J[sidx(2, 2)] = B[0] + Cc[0];
  J[sidx(3, 0)] = __builtin_fma(2.0, Cc[1], A[1]);
  J[sidx(3, 1)] = -A[1];
  J[sidx(3, 2)] = -Cc[2];
  J[sidx(3, 3)] = A[0] + Cc[0];
  J[sidx(4, 0)] = A[2] - B[2];
  J[sidx(4, 1)] = -A[2] - B[2];
  J[sidx(4, 2)] = B[1];
  J[sidx(4, 3)] = A[3];
  J[sidx(4, 4)] = ab;
  S[0] = __builtin_fma(-2.0, Gc[0], Ga[2] - Gb[0]);
  S[1] = -Ga[2] - Gb[0];
  S[2] = Gb[3] + Gc[3];
  S[3] = Ga[3] - Gc[1];
  S[4] = Ga[0] + Gb[2];
}

// U (uniform slip systems): every slip system of the voxel has the same CRSS
// and the same reference slip rate, inv_tau[0] and C.gdot0[0].  Then 1/tau is
// folded into the scaled stress (5 products instead of 12) and gdot0 and the
// exponent into the sums' constants.

// u = w * x, times 1/tau when U
template <bool U>
__device__ __forceinline__ void scaled_u(const double x[5], const double inv_tau[kSlips], double u[5]) {
#pragma unroll
  for (int i = 0; i < 5; i++) u[i] = U ? (w(i) * inv_tau[0]) * x[i] : w(i) * x[i];
}

// max_k |m0_k . sc| / tau_k (MaterialGrid.h:550-566)
template <bool U = false>
__device__ __forceinline__ double max_rss(const double sc[5], const double inv_tau[kSlips]) {
  double u[5];
  scaled_u<U>(sc, inv_tau, u);
  double mx = 0.0;
#pragma unroll
  for (int k = 0; k < kSlips; k++) mx = fmax(mx, fabs(U ? schmid_dot(k, u) : schmid_dot(k, u) * inv_tau[k]));
  return mx;
}

// G = P^T P (packed lower): |P x|^2 = x^T G x for crystal-frame x.
__device__ __forceinline__ void gram(const double P[25], double G[15]) {
#pragma unroll
  for (int a = 0; a < 5; a++)
#pragma unroll
    for (int c = 0; c <= a; c++) {
      double t = 0.0;
#pragma unroll
      for (int i = 0; i < 5; i++) t += P[i * 5 + a] * P[i * 5 + c];
      G[sidx(a, c)] = t;
    }
}

// y = G x for a symmetric G (packed lower)
__device__ __forceinline__ void symv(const double G[15], const double x[5], double y[5]) {
#pragma unroll
  for (int i = 0; i < 5; i++) {
    double t = 0.0;
#pragma unroll
    for (int j = 0; j < 5; j++) t += G[sidx(i, j)] * x[j];
    y[i] = t;
  }
}

// x^T G x for a symmetric G (packed lower)
__device__ __forceinline__ double quad(const double G[15], const double x[5]) {
  double acc = 0.0;
#pragma unroll
  for (int i = 0; i < 5; i++) {
    double t = 0.0;
#pragma unroll
    for (int j = 0; j < i; j++) t += G[sidx(i, j)] * x[j];
    acc += x[i] * __builtin_fma(2.0, t, G[sidx(i, i)] * x[i]);
  }
  return acc;
}

// Solvers::SolveConstrainedConstitutiveEquations (Solvers.cpp:111-222) with
// the Newton matrix built, and the iterate kept, in the crystal frame of the
// voxel.  cc: the start point Q sigma (in), the solution (out; the caller maps
// it back, sigma = P cc = Q^T G cc); G = P^T P with P = Q^-1, so the sample-
// frame step and midpoint norms of the stopping rule are x^T G x; q = P^T
// (L0^-1 lambda + edot + d~); A0 = P^T L0^-1 P (packed lower).  A back-track
// moves cc towards the previous iterate: the map is linear, so that is the
// reference's sample-frame back-track.  Budget, back-track factor and stopping
// rule as in the reference; *err_out untouched when the budget runs out in
// the back-track.  Neither the orientation, nor P, nor a sample-frame iterate
// is live in the loop.
template <int NE, bool U = false>
__device__ __forceinline__ void constrained_nr(double cc[5], const double G[15], const double inv_tau[kSlips],
                                               const StepConst& C, const double q[5], const double A0[15],
                                               double eps, double max_rss, int max_it, double* err_out,
                                               int* steps_out) {
  double saved[5];
#pragma unroll
  for (int i = 0; i < 5; i++) saved[i] = cc[i];
  int rem = max_it;
  int steps = 0;
  double dn = 1.0, an = 1e-10;  // rel = 1e5 until a step is taken (Solvers.cpp:137)
  const double eps2 = eps * eps;
  // U: S and J below are formed from |r|^(n-1) alone; the residual takes
  // them times gdot0, the matrix times n gdot0 / tau
  const double cs = U ? C.gdot0[0] : 1.0, cj = U ? (double)NE * C.gdot0[0] * inv_tau[0] : 1.0;
  while (rem > 0) {
    double r[kSlips];
    bool ok;
    do {
      double u[5];
      scaled_u<U>(cc, inv_tau, u);
      double mx = 0.0;
#pragma unroll
      for (int k = 0; k < kSlips; k++) {
        r[k] = U ? schmid_dot(k, u) : schmid_dot(k, u) * inv_tau[k];
        mx = fmax(mx, fabs(r[k]));
      }
      ok = !(mx > max_rss);
      if (!ok) {
#pragma unroll
        for (int i = 0; i < 5; i++) cc[i] = saved[i] + (cc[i] - saved[i]) * 0.2;
      }
      rem--;
      if (rem < 0) {
        *steps_out = steps;
        return;  // cc is the returned state (Solvers.cpp:169)
      }
    } while (!ok);
    steps++;
    // S_c = sum_k m0_k gdot_k r_k ; J_c = sum_k n gdot_k / tau_k m0_k m0_k^T
    double gr[kSlips], ck[kSlips], S[5], J[15];
#pragma unroll
    for (int k = 0; k < kSlips; k++) {
      if constexpr (U) {
        const double p = NE != 1 ? ipow_c<NE - 1>(fabs(r[k])) : 1.0;
        gr[k] = p * r[k];
        ck[k] = p;
      } else {
        const double g = gamma_nm1<NE>(C, k, fabs(r[k]));
        gr[k] = g * r[k];
        ck[k] = nexp<NE>(C, k) * g * inv_tau[k];
      }
    }
    schmid_sums(gr, ck, S, J);
    // rhs_c = q - A0 cc - S_c ;  (A0 + J_c) d_c = rhs_c
    double rhs[5];
#pragma unroll
    for (int i = 0; i < 5; i++) {
      double Ac = 0.0;
#pragma unroll
      for (int j = 0; j < 5; j++) Ac += A0[i >= j ? sidx(i, j) : sidx(j, i)] * cc[j];
      rhs[i] = q[i] - Ac - (U ? w(i) * cs : w(i)) * S[i];
    }
#pragma unroll
    for (int i = 0; i < 5; i++)
#pragma unroll
      for (int j = 0; j <= i; j++)
        J[sidx(i, j)] = A0[sidx(i, j)] + (U ? (w(i) * w(j)) * cj : w(i) * w(j)) * J[sidx(i, j)];
    ldlt_solve5_inplace(J, rhs);
    // the sample-frame step P d_c and midpoint P (cc + d_c / 2), by their norms
    double mid[5];
#pragma unroll
    for (int i = 0; i < 5; i++) {
      saved[i] = cc[i];
      mid[i] = __builtin_fma(0.5, rhs[i], cc[i]);
      cc[i] += rhs[i];
    }
    dn = quad(G, rhs);
    an = quad(G, mid);
    // rel = |d| / |avg| < eps, compared squared; rel itself is formed once at exit
    if (dn < eps2 * an) break;
  }
  *err_out = sqrt(dn) / sqrt(an);
  *steps_out = steps;
}

// Solvers::ApplyConstitutiveEquations (Solvers.cpp:255-274), crystal frame.
template <int NE, bool U = false>
__device__ __forceinline__ void apply_constitutive(const double sc[5], const double inv_tau[kSlips],
                                                   const StepConst& C, double ec[5]) {
  double u[5], a[5];
  scaled_u<U>(sc, inv_tau, u);
#pragma unroll
  for (int i = 0; i < 5; i++) a[i] = -0.0;
#pragma unroll
  for (int k = 0; k < kSlips; k++) {
    const double r = U ? schmid_dot(k, u) : schmid_dot(k, u) * inv_tau[k];
    // the rate with the sign of r (copysign: one bit operation, where the
    // product by sgn(r) took a compare, a select, a conversion and a multiply;
    // only a zero rate's sign can differ, and it adds nothing)
    const double g = __builtin_copysign(U ? ipow_c<NE>(fabs(r)) : gamma_n<NE>(C, k, fabs(r)), r);
#pragma unroll
    for (int i = 0; i < 5; i++) {
      const int si = kSign[k][i];
      if (si == 1) a[i] += g;
      if (si == -1) a[i] -= g;
      if (si == 2) a[i] += 2.0 * g;
      if (si == -2) a[i] -= 2.0 * g;
    }
  }
#pragma unroll
  for (int i = 0; i < 5; i++) ec[i] = (U ? w(i) * C.gdot0[0] : w(i)) * a[i];
}

}  // namespace cf

// ---------------------------------------------------------------------------
// Bit-faithful prologue path.
//
// The Taylor initialisation (Solvers.cpp:19-101) starts from the Sachs guess,
// where the Newton matrix sum_k n gdot |r_k|^(n-1)/tau m_k m_k^T is dominated
// by one or two slip systems; for n = 20 its condition number reaches ~1e10
// and a 1-ulp change in any input moves the converged Taylor stress by up to
// ~1e-4 relative.  So this path repeats the oracle's floating-point
// operations one for one: Schmid tensors by 3x3 products and basis dot
// products (MaterialGrid.cpp:115-122), divisions by tau, the correctly
// rounded integer power, the same accumulation order, complete-pivoting LU
// (LinearAlgebra.cpp:1171-1187), no FMA contraction.
// ---------------------------------------------------------------------------
namespace faithful {

__device__ __forceinline__ void dd_mul(double ah, double al, double bh, double bl, double* rh, double* rl) {
#pragma clang fp contract(off)
  const double p = ah * bh;
  double e = __builtin_fma(ah, bh, -p);
  e += ah * bl + al * bh;
  const double s = p + e;
  *rl = e - (s - p);
  *rh = s;
}

// x / y, correctly rounded, from rinv = RN(1/y): q = RN(x rinv) is within an
// ulp of x / y, the remainder x - q y is exact in one fma, and q + rem rinv
// rounds to RN(x / y) (Markstein's theorem; normal range).  Three operations
// instead of a full division when one divisor serves many quotients.
__device__ __forceinline__ double div_cr(double x, double y, double rinv) {
  const double q = x * rinv;
  const double rem = __builtin_fma(-q, y, x);
  return __builtin_fma(rem, rinv, q);
}

// x^n (n >= 0) rounded once from a double-double binary power.
__device__ __forceinline__ double ipow_cr(double x, int n) {
#pragma clang fp contract(off)
  double rh = 1.0, rl = 0.0, bh = x, bl = 0.0;
  while (n > 0) {
    if (n & 1) dd_mul(rh, rl, bh, bl, &rh, &rl);
    n >>= 1;
    if (n) dd_mul(bh, bl, bh, bl, &bh, &bl);
  }
  return rh + rl;
}

__device__ __forceinline__ double dot5(const double* a, const double* b) {
#pragma clang fp contract(off)
  double r = 0.0;
#pragma unroll
  for (int i = 0; i < 5; i++) r += a[i] * b[i];
  return r;
}

__device__ __forceinline__ double norm5(const double* a) {
#pragma clang fp contract(off)
  return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2] + a[3] * a[3] + a[4] * a[4]);
}

// basis entry (a, e) of LinearAlgebra.h:207-247, e = 3 i + j
__device__ __forceinline__ double basis(int a, int e) {
  const double s2 = 1.0 / 1.41421356237309504880;  // 1/sqrt(2.0) as computed in double
  const double s6 = 1.0 / 2.44948974278317809820;
  const signed char pat[5][9] = {{-1, 0, 0, 0, 1, 0, 0, 0, 0}, {-1, 0, 0, 0, -1, 0, 0, 0, 2},
                                 {0, 0, 0, 0, 0, 1, 0, 1, 0},  {0, 0, 1, 0, 0, 0, 1, 0, 0},
                                 {0, 1, 0, 1, 0, 0, 0, 0, 0}};
  return (double)pat[a][e] * (a == 1 ? s6 : s2);
}

__device__ __forceinline__ void schmid(const double R[9], double m[kSlips][5]) {
#pragma clang fp contract(off)
  const double d = 2.0 * 2.44948974278317809820;  // 2 sqrt(6)
#pragma unroll
  for (int k = 0; k < kSlips; k++) {
    double m0[9], tmp[9], rot[9];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++)
        m0[3 * i + j] = ((double)(kB[k][i] * kN[k][j]) + (double)(kN[k][i] * kB[k][j])) / d;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++)
        tmp[3 * i + j] = R[3 * i + 0] * m0[0 * 3 + j] + R[3 * i + 1] * m0[1 * 3 + j] + R[3 * i + 2] * m0[2 * 3 + j];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++)  // (R m0) R^T
        rot[3 * i + j] = tmp[3 * i + 0] * R[3 * j + 0] + tmp[3 * i + 1] * R[3 * j + 1] + tmp[3 * i + 2] * R[3 * j + 2];
#pragma unroll
    for (int a = 0; a < 5; a++) {
      double r = 0.0;
#pragma unroll
      for (int e = 0; e < 9; e++) r += rot[e] * basis(a, e);
      m[k][a] = r;
    }
  }
}

// Where a voxel's 12 Schmid tensors are read from: m(k, i) is component i
// of slip k.  SchmidRegs: an array in registers.  SchmidSplit: slips below KR
// in registers, the others in this wave's LDS rows (lane-major, so a read is
// one conflict-free 64-lane access); the values are the same doubles either
// way, only fewer registers are live through the Taylor solve.
struct SchmidRegs {
  const double (*m)[5];
  __device__ __forceinline__ double operator()(int k, int i) const { return m[k][i]; }
};
template <int KR>
struct SchmidSplit {
  double reg[KR > 0 ? KR : 1][5];
  const double* lds;  // [(k - KR) * 5 + i][64]
  int lane;
  __device__ __forceinline__ double operator()(int k, int i) const {
    return k < KR ? reg[k < KR ? k : 0][i] : lds[((k - KR) * 5 + i) * 64 + lane];
  }
};

// m_k . s in dot5's order
template <typename MS>
__device__ __forceinline__ double mdot(const MS& m, int k, const double* s) {
#pragma clang fp contract(off)
  double r = 0.0;
#pragma unroll
  for (int i = 0; i < 5; i++) r += m(k, i) * s[i];
  return r;
}

template <typename MS>
__device__ __forceinline__ double max_rss_m(const double s[5], const MS& m, const double tau[kSlips]) {
#pragma clang fp contract(off)
  double mx = 0.0;
#pragma unroll
  for (int k = 0; k < kSlips; k++) {
    const double v = fabs(mdot(m, k, s) / tau[k]);
    mx = v > mx ? v : mx;
  }
  return mx;
}
__device__ __forceinline__ double max_rss(const double s[5], const double m[kSlips][5], const double tau[kSlips]) {
  return max_rss_m(s, SchmidRegs{m}, tau);
}

template <typename MS>
__device__ __forceinline__ void sachs_m(const MS& m, const double tau[kSlips], const double edot[5], double out[5]) {
#pragma clang fp contract(off)
  double mx = 0.0;
#pragma unroll
  for (int k = 0; k < kSlips; k++) {
    const double v = fabs(mdot(m, k, edot) / tau[k]);
    mx = v > mx ? v : mx;
  }
#pragma unroll
  for (int i = 0; i < 5; i++) out[i] = edot[i] / mx;
}
__device__ __forceinline__ void sachs(const double m[kSlips][5], const double tau[kSlips], const double edot[5],
                                      double out[5]) {
  sachs_m(SchmidRegs{m}, tau, edot, out);
}

// Complete-pivoting LU solve; A column-major.  Pivot scan column-major with a
// strict '>' (first maximum wins); swaps done with selects so every index is
// static.  Unknowns past the rank of an exactly singular matrix are 0.
__device__ __forceinline__ void lu_fullpiv(const double Ac[25], const double bi[5], double x[5]) {
#pragma clang fp contract(off)
  double A[5][5], b[5];
  int cp[5];
#pragma unroll
  for (int i = 0; i < 5; i++) {
#pragma unroll
    for (int j = 0; j < 5; j++) A[i][j] = Ac[j * 5 + i];
    b[i] = bi[i];
    cp[i] = i;
  }
  int rank = 5;
  bool stop = false;
#pragma unroll
  for (int k = 0; k < 5; k++) {
    double best = -1.0;
    int pr = k, pc = k;
#pragma unroll
    for (int j = k; j < 5; j++)
#pragma unroll
      for (int i = k; i < 5; i++) {
        const double v = fabs(A[i][j]);
        const bool gt = v > best;
        best = gt ? v : best;
        pr = gt ? i : pr;
        pc = gt ? j : pc;
      }
    if (!stop && best == 0.0) {
      rank = k;
      stop = true;
    }
    if (!stop) {
#pragma unroll
      for (int i = k + 1; i < 5; i++) {
        const bool sw = pr == i;
#pragma unroll
        for (int j = 0; j < 5; j++) {
          const double t = A[k][j];
          A[k][j] = sw ? A[i][j] : t;
          A[i][j] = sw ? t : A[i][j];
        }
        const double t = b[k];
        b[k] = sw ? b[i] : t;
        b[i] = sw ? t : b[i];
      }
#pragma unroll
      for (int j = k + 1; j < 5; j++) {
        const bool sw = pc == j;
#pragma unroll
        for (int i = 0; i < 5; i++) {
          const double t = A[i][k];
          A[i][k] = sw ? A[i][j] : t;
          A[i][j] = sw ? t : A[i][j];
        }
        const int t = cp[k];
        cp[k] = sw ? cp[j] : t;
        cp[j] = sw ? t : cp[j];
      }
      const double rkk = 1.0 / A[k][k];  // the multipliers below are quotients by A[k][k] (div_cr)
#pragma unroll
      for (int i = k + 1; i < 5; i++) {
        const double f = div_cr(A[i][k], A[k][k], rkk);
        A[i][k] = f;
#pragma unroll
        for (int j = k + 1; j < 5; j++) A[i][j] -= f * A[k][j];
        b[i] -= f * b[k];
      }
    }
  }
  double y[5];
#pragma unroll
  for (int i = 4; i >= 0; i--) {
    double s = b[i];
#pragma unroll
    for (int j = i + 1; j < 5; j++)
      if (j < rank) s -= A[i][j] * y[j];
    y[i] = i < rank ? s / A[i][i] : 0.0;
  }
#pragma unroll
  for (int c = 0; c < 5; c++) {
    double v = 0.0;
#pragma unroll
    for (int i = 0; i < 5; i++) v = cp[i] == c ? y[i] : v;
    x[c] = v;
  }
}

// Inverse of a 5x5 (column-major in and out) by Gauss-Jordan elimination with
// partial pivoting, the oracle's stand-in for Eigen's inverse()
// (MaterialGrid.cpp:248, 632), operation for operation: the pivot row is the
// first maximum of |column k| at or below row k, the pivot row is scaled by
// the reciprocal of the pivot, every other row with a nonzero multiplier is
// reduced.  Columns of the left block at or left of the pivot are not read
// again, so they are not kept.  Returns false on an exactly zero pivot
// (out is then not meaningful).
__device__ __forceinline__ bool inverse_gj5(const double Ac[25], double out[25]) {
#pragma clang fp contract(off)
  double L[5][5], Q[5][5];  // left block, right block (row i, column j)
#pragma unroll
  for (int i = 0; i < 5; i++)
#pragma unroll
    for (int j = 0; j < 5; j++) {
      L[i][j] = Ac[j * 5 + i];
      Q[i][j] = i == j ? 1.0 : 0.0;
    }
  bool ok = true;
#pragma unroll
  for (int k = 0; k < 5; k++) {
    int p = k;
    double best = fabs(L[k][k]);
#pragma unroll
    for (int i = k + 1; i < 5; i++) {
      const double v = fabs(L[i][k]);
      const bool gt = v > best;
      best = gt ? v : best;
      p = gt ? i : p;
    }
    ok = ok && best != 0.0;
#pragma unroll
    for (int i = k + 1; i < 5; i++) {
      const bool sw = p == i;
#pragma unroll
      for (int j = k; j < 5; j++) {
        const double t = L[k][j];
        L[k][j] = sw ? L[i][j] : t;
        L[i][j] = sw ? t : L[i][j];
      }
#pragma unroll
      for (int j = 0; j < 5; j++) {
        const double t = Q[k][j];
        Q[k][j] = sw ? Q[i][j] : t;
        Q[i][j] = sw ? t : Q[i][j];
      }
    }
    const double inv = 1.0 / L[k][k];
#pragma unroll
    for (int j = k + 1; j < 5; j++) L[k][j] *= inv;
#pragma unroll
    for (int j = 0; j < 5; j++) Q[k][j] *= inv;
#pragma unroll
    for (int i = 0; i < 5; i++) {
      if (i == k) continue;
      const double f = L[i][k];
      const bool nz = f != 0.0;
#pragma unroll
      for (int j = k + 1; j < 5; j++) L[i][j] = nz ? L[i][j] - f * L[k][j] : L[i][j];
#pragma unroll
      for (int j = 0; j < 5; j++) Q[i][j] = nz ? Q[i][j] - f * Q[k][j] : Q[i][j];
    }
  }
#pragma unroll
  for (int i = 0; i < 5; i++)
#pragma unroll
    for (int j = 0; j < 5; j++) out[j * 5 + i] = Q[i][j];
  return ok;
}

// The voxel's secant modulus M = sum_k m_k m_k^T gdot0_k |r_k|^(n_k - 1) / tau_k
// of BuildLinearRefMediumStiffness (MaterialGrid.cpp:228-243) at the Taylor
// stress s, column-major, in the oracle's operation order (quotients by
// tau, the correctly rounded power).  Exactly symmetric: m_i m_j == m_j m_i.
__device__ __forceinline__ void secant_modulus(const double m[kSlips][5], const double tau[kSlips], const double s[5],
                                               const StepConst& C, double M[25]) {
#pragma clang fp contract(off)
#pragma unroll
  for (int e = 0; e < 25; e++) M[e] = 0.0;
#pragma unroll
  for (int k = 0; k < kSlips; k++) {
    const double rt = 1.0 / tau[k];
    const double rss = div_cr(dot5(m[k], s), tau[k], rt);
    const int n = C.uniform_n ? C.nexp[0] : C.nexp[k];
    const double g = (n - 1 > 0) ? C.gdot0[k] * ipow_cr(fabs(rss), n - 1) : C.gdot0[k];
#pragma unroll
    for (int j = 0; j < 5; j++)
#pragma unroll
      for (int i = j; i < 5; i++) M[j * 5 + i] += div_cr(m[k][i] * m[k][j] * g, tau[k], rt);
  }
#pragma unroll
  for (int j = 0; j < 5; j++)
#pragma unroll
    for (int i = 0; i < j; i++) M[j * 5 + i] = M[i * 5 + j];
}

// Solvers::SolveConstitutiveEquations (Solvers.cpp:19-101), operation for operation.
template <typename MS>
__device__ __forceinline__ void taylor_nr_m(const double init[5], const MS& m, const double tau[kSlips],
                                          const StepConst& C, double eps, int max_it, double out[5],
                                          int* steps_out = nullptr) {
#pragma clang fp contract(off)
  double cur[5], nw[5], saved[5], rtau[kSlips];
#pragma unroll
  for (int i = 0; i < 5; i++) { cur[i] = init[i]; nw[i] = init[i]; saved[i] = 0.0; }
#pragma unroll
  for (int k = 0; k < kSlips; k++) rtau[k] = 1.0 / tau[k];  // quotients by tau via div_cr
  int rem = max_it;
  int steps = 0;
  const double max_rss = C.imposed_max_rss;
  while (rem > 0) {
    // the Schmid tensors held in LDS are read again each step rather than
    // kept in registers across the loop
    __asm__ volatile("" ::: "memory");
    double rss[kSlips];
    bool ok;
    do {
      ok = true;
#pragma unroll
      for (int k = 0; k < kSlips; k++) {
        double r = div_cr(mdot(m, k, cur), tau[k], rtau[k]);
        if (fabs(r) < 1e-10) r = 0.0;
        if (fabs(r) > max_rss) ok = false;
        rss[k] = r;
      }
      if (!ok) {
#pragma unroll
        for (int i = 0; i < 5; i++) cur[i] = saved[i] + (cur[i] - saved[i]) * 0.2;
      }
      rem--;
      if (rem < 0) {
#pragma unroll
        for (int i = 0; i < 5; i++) out[i] = nw[i];
        if (steps_out) *steps_out = steps;
        return;
      }
    } while (!ok);
    double res[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double J[25];
#pragma unroll
    for (int e = 0; e < 25; e++) J[e] = 0.0;
#pragma unroll
    for (int k = 0; k < kSlips; k++) {
      const int n = C.uniform_n ? C.nexp[0] : C.nexp[k];
      const double g = (n - 1 > 0) ? C.gdot0[k] * ipow_cr(fabs(rss[k]), n - 1) : C.gdot0[k];
#pragma unroll
      for (int i = 0; i < 5; i++) res[i] += m(k, i) * g * rss[k];
      // J(i,j) and J(j,i) receive bit-identical increments (m_i m_j == m_j m_i),
      // so the lower triangle is computed and mirrored
#pragma unroll
      for (int j = 0; j < 5; j++)
#pragma unroll
        for (int i = j; i < 5; i++) {
          const double t = m(k, i) * m(k, j) * (double)n * g;
          J[j * 5 + i] -= div_cr(t, tau[k], rtau[k]);
        }
    }
#pragma unroll
    for (int j = 0; j < 5; j++)
#pragma unroll
      for (int i = 0; i < j; i++) J[j * 5 + i] = J[i * 5 + j];
#pragma unroll
    for (int i = 0; i < 5; i++) res[i] = res[i] - C.edot[i];
#pragma unroll
    for (int i = 0; i < 5; i++) saved[i] = cur[i];
    double delta[5];
    lu_fullpiv(J, res, delta);
    steps++;
    double sum[5];
#pragma unroll
    for (int i = 0; i < 5; i++) {
      nw[i] = cur[i] + delta[i];
      sum[i] = nw[i] + cur[i];
    }
    const double rel = 2.0 * norm5(delta) / norm5(sum);
#pragma unroll
    for (int i = 0; i < 5; i++) cur[i] = nw[i];
    if (rel < eps) break;
  }
#pragma unroll
  for (int i = 0; i < 5; i++) out[i] = nw[i];
  if (steps_out) *steps_out = steps;
}

__device__ __forceinline__ void taylor_nr(const double init[5], const double m[kSlips][5], const double tau[kSlips],
                                          const StepConst& C, double eps, int max_it, double out[5],
                                          int* steps_out = nullptr) {
  taylor_nr_m(init, SchmidRegs{m}, tau, C, eps, max_it, out, steps_out);
}

// Solvers::SolveConstrainedConstitutiveEquations (Solvers.cpp:111-222) in the
// reference's own form, for Schmid tensors given in the sample frame: the
// Newton matrix J' = -I - L0 J (:202-204) solved by the complete-pivoting LU,
// operation for operation as the oracle restates it.  C supplies gdot0, the
// exponents, edot and L0 (column-major).  *err_out is left untouched when the
// budget runs out in the admissibility back-track (:166-170).
__device__ __forceinline__ void constrained_literal(const double init[5], const double m[kSlips][5],
                                                    const double tau[kSlips], const StepConst& C, const double lam[5],
                                                    const double dtl[5], double eps, double max_rss, int max_it,
                                                    double out[5], double* err_out) {
#pragma clang fp contract(off)
  double cur[5], saved[5], nw[5] = {-1.0, -3.0, -7.0, -9.0, -11.0};
  double rel = 1e5;
#pragma unroll
  for (int i = 0; i < 5; i++) { cur[i] = init[i]; saved[i] = init[i]; }
  int rem = max_it;
  while (rem > 0) {
    double rss[kSlips];
    bool ok;
    do {
      ok = true;
#pragma unroll
      for (int k = 0; k < kSlips; k++) {
        rss[k] = dot5(m[k], cur) / tau[k];
        if (fabs(rss[k]) > max_rss) ok = false;
      }
      if (!ok) {
#pragma unroll
        for (int i = 0; i < 5; i++) cur[i] = saved[i] + (cur[i] - saved[i]) * 0.2;
      }
      rem--;
      if (rem < 0) {
#pragma unroll
        for (int i = 0; i < 5; i++) out[i] = cur[i];
        return;
      }
    } while (!ok);
    double S[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, J[25];
#pragma unroll
    for (int e = 0; e < 25; e++) J[e] = 0.0;
#pragma unroll
    for (int k = 0; k < kSlips; k++) {
      const int n = C.uniform_n ? C.nexp[0] : C.nexp[k];
      const double g = n != 1 ? C.gdot0[k] * ipow_cr(fabs(rss[k]), n - 1) : C.gdot0[k];
#pragma unroll
      for (int i = 0; i < 5; i++) S[i] += m[k][i] * g * rss[k];
#pragma unroll
      for (int j = 0; j < 5; j++)
#pragma unroll
        for (int i = 0; i < 5; i++) J[j * 5 + i] += m[k][i] * m[k][j] * (double)n * g / tau[k];
    }
    double Jp[25];
#pragma unroll
    for (int j = 0; j < 5; j++)
#pragma unroll
      for (int i = 0; i < 5; i++) {
        double r = 0.0;
#pragma unroll
        for (int k = 0; k < 5; k++) r += C.L0[k * 5 + i] * J[j * 5 + k];
        Jp[j * 5 + i] = (i == j ? -1.0 : -0.0) - r;
      }
    double t[5], res[5];
#pragma unroll
    for (int i = 0; i < 5; i++) t[i] = S[i] - C.edot[i] - dtl[i];
#pragma unroll
    for (int i = 0; i < 5; i++) {
      double r = 0.0;
#pragma unroll
      for (int j = 0; j < 5; j++) r += C.L0[j * 5 + i] * t[j];
      res[i] = cur[i] - lam[i] + r;
    }
#pragma unroll
    for (int i = 0; i < 5; i++) saved[i] = cur[i];
    double delta[5], diff[5], avg[5];
    lu_fullpiv(Jp, res, delta);
#pragma unroll
    for (int i = 0; i < 5; i++) {
      nw[i] = cur[i] + delta[i];
      diff[i] = nw[i] - saved[i];
      avg[i] = (nw[i] + saved[i]) * 0.5;
    }
    rel = norm5(diff) / norm5(avg);
#pragma unroll
    for (int i = 0; i < 5; i++) cur[i] = nw[i];
    if (rel < eps) break;
  }
  *err_out = rel;
#pragma unroll
  for (int i = 0; i < 5; i++) out[i] = nw[i];
}

// gdot0 |r|^n sgn(r) of one slip system for the rate laws of Solvers.cpp:255-373
// (n > 0; gdot0 alone otherwise), r = m . s / tau, with the correctly rounded power.
__device__ __forceinline__ double slip_rate(const StepConst& C, int k, double rss) {
#pragma clang fp contract(off)
  const int n = C.uniform_n ? C.nexp[0] : C.nexp[k];
  return n > 0 ? C.gdot0[k] * ipow_cr(fabs(rss), n) : C.gdot0[k];
}

}  // namespace faithful

}  // namespace vpfft
