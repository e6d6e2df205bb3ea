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

// In-place variant: A's packed lower triangle is overwritten by L (unit
// diagonal implied), the solution replaces b.  Same arithmetic as ldlt_solve5.
__device__ __forceinline__ void ldlt_solve5_inplace(double A[15], double b[5]) {
  double dinv[5];
#pragma unroll
  for (int j = 0; j < 5; j++) {
    double d = A[sidx(j, j)];
#pragma unroll
    for (int k = 0; k < j; k++) d -= A[sidx(j, k)] * A[sidx(j, k)] * A[sidx(k, k)];
    const double scale = fabs(A[sidx(j, j)]);
    const bool ok = d > 1e-300 && d > 1e-14 * scale;
    const double dj = ok ? d : 0.0;
    dinv[j] = ok ? 1.0 / d : 0.0;
#pragma unroll
    for (int i = j + 1; i < 5; i++) {
      double t = A[sidx(i, j)];
#pragma unroll
      for (int k = 0; k < j; k++) t -= A[sidx(i, k)] * A[sidx(j, k)] * A[sidx(k, k)];
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

// Solvers::SolveConstitutiveEquations (Taylor NR), Solvers.cpp:19-101.
__device__ __forceinline__ void taylor_nr(const double init[5], const double m[kSlips][5],
                                          const double inv_tau[kSlips], const StepConst& C,
                                          double eps, int max_it, double out[5]) {
  double cur[5], nw[5], saved[5];
#pragma unroll
  for (int i = 0; i < 5; i++) { cur[i] = init[i]; nw[i] = init[i]; saved[i] = 0.0; }
  int rem = max_it;
  const double max_rss = C.imposed_max_rss;
  while (rem > 0) {
    double rss[kSlips];
    bool ok;
    do {
      ok = true;
#pragma unroll
      for (int k = 0; k < kSlips; k++) {
        double r = dot5(m[k], cur) * inv_tau[k];
        r = fabs(r) < 1e-10 ? 0.0 : r;
        ok = ok && !(fabs(r) > max_rss);
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
        return;
      }
    } while (!ok);
    double res[5] = {0, 0, 0, 0, 0};
    double A[15];
#pragma unroll
    for (int e = 0; e < 15; e++) A[e] = 0.0;
#pragma unroll
    for (int k = 0; k < kSlips; k++) {
      const double g = gamma_nm1_taylor(C, k, fabs(rss[k]));
      const double gr = g * rss[k];
      const double ck = (double)nexp_of(C, k) * g * inv_tau[k];
#pragma unroll
      for (int i = 0; i < 5; i++) {
        res[i] += m[k][i] * gr;
#pragma unroll
        for (int j = 0; j <= i; j++) A[sidx(i, j)] += ck * m[k][i] * m[k][j];
      }
    }
    // J = -A ; delta = J^-1 (res - edot) = A^-1 (edot - res)
    double rhs[5], delta[5];
#pragma unroll
    for (int i = 0; i < 5; i++) rhs[i] = C.edot[i] - res[i];
    ldlt_solve5(A, rhs, delta);
    double sum[5];
#pragma unroll
    for (int i = 0; i < 5; i++) {
      saved[i] = cur[i];
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
}

// Solvers::SolveConstrainedConstitutiveEquations, Solvers.cpp:111-222.
// q = L0^-1 lambda + edot + dtilde (constant over the solve).  *err_out is
// left untouched when the budget runs out inside the back-track (:166-170).
// Register-lean form: resolved shear stresses are recomputed where needed
// rather than kept (12 doubles), the iterate doubles as the returned value,
// and the 5x5 system is factored in place.
__device__ __forceinline__ void constrained_nr(double cur[5], const double m[kSlips][5],
                                               const double inv_tau[kSlips], const StepConst& C,
                                               const double q[5], double eps, double max_rss,
                                               int max_it, double* err_out, int* steps_out) {
  double saved[5];
  double rel = 1e5;
#pragma unroll
  for (int i = 0; i < 5; i++) saved[i] = cur[i];
  int rem = max_it;
  int steps = 0;
  while (rem > 0) {
    bool ok;
    do {
      double mx = 0.0;
#pragma unroll
      for (int k = 0; k < kSlips; k++) mx = fmax(mx, fabs(dot5(m[k], cur) * inv_tau[k]));
      ok = !(mx > max_rss);
      if (!ok) {
#pragma unroll
        for (int i = 0; i < 5; i++) cur[i] = saved[i] + (cur[i] - saved[i]) * 0.2;
      }
      rem--;
      if (rem < 0) {
        *steps_out = steps;
        return;  // cur is the returned state (Solvers.cpp:169)
      }
    } while (!ok);
    steps++;
    double S[5] = {0, 0, 0, 0, 0};
    double A[15];
#pragma unroll
    for (int i = 0, e = 0; i < 5; i++)
#pragma unroll
      for (int j = 0; j <= i; j++, e++) A[e] = C.L0inv[j * 5 + i];
#pragma unroll
    for (int k = 0; k < kSlips; k++) {
      const double r = dot5(m[k], cur) * inv_tau[k];
      const double g = gamma_nm1(C, k, fabs(r));
      const double gr = g * r;
      const double ck = (double)nexp_of(C, k) * g * inv_tau[k];
#pragma unroll
      for (int i = 0; i < 5; i++) {
        S[i] += m[k][i] * gr;
        const double cm = ck * m[k][i];
#pragma unroll
        for (int j = 0; j <= i; j++) A[sidx(i, j)] += cm * m[k][j];
      }
    }
    // L0^-1 Res = L0^-1 cur + S - q  ;  delta = -(L0^-1 + J)^-1 L0^-1 Res
    double rhs[5];
#pragma unroll
    for (int i = 0; i < 5; i++) {
      double Lc = 0.0;
#pragma unroll
      for (int j = 0; j < 5; j++) Lc += C.L0inv[j * 5 + i] * cur[j];
      rhs[i] = q[i] - Lc - S[i];
    }
    ldlt_solve5_inplace(A, rhs);
    double dn = 0.0, an = 0.0;
#pragma unroll
    for (int i = 0; i < 5; i++) {
      saved[i] = cur[i];
      const double nw = cur[i] + rhs[i];
      const double diff = nw - saved[i];
      const double avg = (nw + saved[i]) * 0.5;
      dn += diff * diff;
      an += avg * avg;
      cur[i] = nw;
    }
    rel = sqrt(dn) / sqrt(an);
    if (rel < eps) break;
  }
  *err_out = rel;
  *steps_out = steps;
}

// Solvers::ApplyConstitutiveEquations, Solvers.cpp:255-274.
__device__ __forceinline__ void apply_constitutive(const double s[5], const double m[kSlips][5],
                                                   const double inv_tau[kSlips],
                                                   const StepConst& C, double e[5]) {
#pragma unroll
  for (int i = 0; i < 5; i++) e[i] = 0.0;
#pragma unroll
  for (int k = 0; k < kSlips; k++) {
    const double r = dot5(m[k], s) * inv_tau[k];
    const double g = gamma_n(C, k, fabs(r)) * sgn(r);
#pragma unroll
    for (int i = 0; i < 5; i++) e[i] += m[k][i] * g;
  }
}

// max_k |m_k . s / tau_k|  (MaterialGrid.h:550-566)
__device__ __forceinline__ double max_rss(const double s[5], const double m[kSlips][5],
                                          const double inv_tau[kSlips]) {
  double mx = 0.0;
#pragma unroll
  for (int k = 0; k < kSlips; k++) mx = fmax(mx, fabs(dot5(m[k], s) * inv_tau[k]));
  return mx;
}

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

__device__ __forceinline__ double max_rss(const double s[5], const double m[kSlips][5], const double tau[kSlips]) {
#pragma clang fp contract(off)
  double mx = 0.0;
#pragma unroll
  for (int k = 0; k < kSlips; k++) {
    const double v = fabs(dot5(m[k], s) / tau[k]);
    mx = v > mx ? v : mx;
  }
  return mx;
}

__device__ __forceinline__ void sachs(const double m[kSlips][5], const double tau[kSlips], const double edot[5],
                                      double out[5]) {
#pragma clang fp contract(off)
  double mx = 0.0;
#pragma unroll
  for (int k = 0; k < kSlips; k++) {
    const double v = fabs(dot5(edot, m[k]) / tau[k]);
    mx = v > mx ? v : mx;
  }
#pragma unroll
  for (int i = 0; i < 5; i++) out[i] = edot[i] / mx;
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
#pragma unroll
      for (int i = k + 1; i < 5; i++) {
        const double f = A[i][k] / A[k][k];
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

// Solvers::SolveConstitutiveEquations (Solvers.cpp:19-101), operation for operation.
__device__ __forceinline__ void taylor_nr(const double init[5], const double m[kSlips][5], const double tau[kSlips],
                                          const StepConst& C, double eps, int max_it, double out[5]) {
#pragma clang fp contract(off)
  double cur[5], nw[5], saved[5];
#pragma unroll
  for (int i = 0; i < 5; i++) { cur[i] = init[i]; nw[i] = init[i]; saved[i] = 0.0; }
  int rem = max_it;
  const double max_rss = C.imposed_max_rss;
  while (rem > 0) {
    double rss[kSlips];
    bool ok;
    do {
      ok = true;
#pragma unroll
      for (int k = 0; k < kSlips; k++) {
        double r = dot5(m[k], cur) / tau[k];
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
      for (int i = 0; i < 5; i++) res[i] += m[k][i] * g * rss[k];
      // J(i,j) and J(j,i) receive bit-identical increments (m_i m_j == m_j m_i),
      // so the lower triangle is computed and mirrored; x / 1.0 == x exactly,
      // so the division is skipped when tau is 1
      const bool unit_tau = tau[k] == 1.0;
#pragma unroll
      for (int j = 0; j < 5; j++)
#pragma unroll
        for (int i = j; i < 5; i++) {
          const double t = m[k][i] * m[k][j] * (double)n * g;
          J[j * 5 + i] -= unit_tau ? t : t / tau[k];
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
}

}  // namespace faithful

}  // namespace vpfft
