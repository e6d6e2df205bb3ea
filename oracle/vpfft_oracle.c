/*
 * vpfft_oracle.c -- CPU FP64 restatement of exmatex/VPFFT's strain step
 * (MaterialGrid::RunSingleStrainStep and everything it calls).
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker the GPU library is compared
 * against and the CPU baseline timer.  The product library never links it.
 *
 * Each function cites the reference file:line it restates.  Where the
 * reference uses a third-party routine that is absent here the restatement
 * says what replaces it:
 *   - Eigen fullPivLu().solve (LinearAlgebra.cpp:1171-1187) -> Gaussian
 *     elimination with complete pivoting (vo_lu_solve_fullpiv);
 *   - Eigen inverse() on 4x4 / 5x5 (MaterialGrid.cpp:248,604,632) ->
 *     Gauss-Jordan with partial pivoting (vo_inverse);
 *   - FFTW3 unnormalised C2C, forward sign -1 (MaterialGrid.cpp:356-375,
 *     639,746) -> recursive mixed-radix Cooley-Tukey (vo_fft1d/vo_fft3d).
 *   - std::pow(double, int) (Solvers.cpp:67,178,268,296,335; MaterialGrid.cpp:237)
 *     -> vo_ipow, the exact power rounded once (double-double binary
 *     powering).  glibc's pow is within 0.52 ulp of it and differs in about
 *     0.1% of arguments; a defined rounding lets the GPU reproduce the
 *     ill-conditioned Taylor initialisation bit for bit.
 * These change results only at rounding level (~1e-16 relative).
 *
 * Parity unpinned: nothing can reproduce the reference binary here (Eigen
 * and FFTW-MPI are absent and the reference ships no fixtures), so
 * vo_strain_step's iteration count and per-voxel Newton trajectories on a
 * heterogeneous polycrystal are pinned only by the consistency of its own
 * pieces (each checked against closed forms, invariants and numpy) and by
 * its special cases (homogeneous single crystal, iteration cap, degenerate
 * loading); the exact roundings of Eigen's LU / inverse and of FFTW are
 * unpinned for the same reason.
 *
 * Per-voxel work runs in parallel with OpenMP, but every reduction the
 * reference performs in voxel order (running mean of L0, the per-iteration
 * sums) is done sequentially in that same order, so results do not depend on
 * the thread count.
 */
#include "vpfft_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int vo_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Tables                                                                    */
/* ------------------------------------------------------------------------ */

/* Deviatoric basis, Src/LinearAlgebra.h:207-247.  Entries are
 * (matrix entry) * (1/sqrt(c)), the product order of operator*(Float,
 * SMatrix3x3) (LinearAlgebra.cpp:459-462, 271-288). */
void vo_eigen_basis(double B[6 * 9]) {
  static const double pat[6][9] = {
      {-1, 0, 0, 0, 1, 0, 0, 0, 0},
      {-1, 0, 0, 0, -1, 0, 0, 0, 2},
      {0, 0, 0, 0, 0, 1, 0, 1, 0},
      {0, 0, 1, 0, 0, 0, 1, 0, 0},
      {0, 1, 0, 1, 0, 0, 0, 0, 0},
      {1, 0, 0, 0, 1, 0, 0, 0, 1}};
  const double f[6] = {1.0 / sqrt(2.0), 1.0 / sqrt(6.0), 1.0 / sqrt(2.0),
                       1.0 / sqrt(2.0), 1.0 / sqrt(2.0), 1.0 / sqrt(3.0)};
  for (int a = 0; a < 6; a++)
    for (int e = 0; e < 9; e++) B[a * 9 + e] = pat[a][e] * f[a];
}

/* FCC {111}<110> slip systems in the reference's order,
 * Src/BurgerVectorBasis.h:16-72. */
static const double kBurgers[12][3] = {
    {0, 1, 1}, {1, 0, 1}, {1, -1, 0}, {0, 1, -1}, {1, 0, 1}, {1, 1, 0},
    {0, 1, 1}, {1, 0, -1}, {1, 1, 0}, {0, 1, -1}, {1, 0, -1}, {1, -1, 0}};
static const double kNormals[12][3] = {
    {1, 1, -1}, {1, 1, -1}, {1, 1, -1}, {1, -1, -1}, {1, -1, -1}, {1, -1, -1},
    {1, -1, 1}, {1, -1, 1}, {1, -1, 1}, {1, 1, 1},   {1, 1, 1},   {1, 1, 1}};

/* (b (x) n +/- n (x) b) / (2 sqrt 6), Src/BurgerVectorBasis.h:122-150. */
void vo_schmid_basis(double sym[12 * 9], double anti[12 * 9]) {
  const double d = 2.0 * sqrt(6.0);
  for (int s = 0; s < 12; s++)
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        double m1 = kBurgers[s][i] * kNormals[s][j];
        double m2 = kNormals[s][i] * kBurgers[s][j];
        if (sym) sym[s * 9 + i * 3 + j] = (m1 + m2) / d;
        if (anti) anti[s * 9 + i * 3 + j] = (m1 - m2) / d;
      }
}

typedef struct {
  double B[6 * 9];
  double m0[12 * 9];
  double a0[12 * 9];
} tables_t;

static const tables_t *tables(void) {
  static tables_t t;
  static int init = 0;
  if (!init) {
#ifdef _OPENMP
#pragma omp critical(vo_tables)
#endif
    {
      if (!init) {
        vo_eigen_basis(t.B);
        vo_schmid_basis(t.m0, t.a0);
        init = 1;
      }
    }
  }
  return &t;
}

/* EigenRep(const SMatrix3x3&): v_a = <M, B_a>, LinearAlgebra.h:307-312,
 * InnerProduct LinearAlgebra.cpp:812-820. */
void vo_eigenrep_from_matrix(const double M[9], double v[5]) {
  const double *B = tables()->B;
  for (int a = 0; a < 5; a++) {
    double r = 0;
    for (int e = 0; e < 9; e++) r += M[e] * B[a * 9 + e];
    v[a] = r;
  }
}

/* EigenRep::ToMatrixRep, LinearAlgebra.h:352-360. */
void vo_matrix_from_eigenrep(const double v[5], double M[9]) {
  const double *B = tables()->B;
  for (int e = 0; e < 9; e++) M[e] = 0;
  for (int a = 0; a < 5; a++)
    for (int e = 0; e < 9; e++) M[e] += B[a * 9 + e] * v[a];
}

static void mat3_mul(const double A[9], const double B[9], double C[9]) {
  /* SMatrix3x3::operator*, LinearAlgebra.cpp:419-439 */
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      C[i * 3 + j] = A[i * 3 + 0] * B[0 * 3 + j] + A[i * 3 + 1] * B[1 * 3 + j] +
                     A[i * 3 + 2] * B[2 * 3 + j];
}

static void mat3_transpose(const double A[9], double T[9]) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) T[j * 3 + i] = A[i * 3 + j];
}

/* m_k = EigenRep(R * m0_k * R^T), MaterialGrid.cpp:107-124. */
void vo_update_schmid(const double R[9], double m[60]) {
  const tables_t *t = tables();
  double RT[9], tmp[9], rot[9];
  mat3_transpose(R, RT);
  for (int s = 0; s < 12; s++) {
    mat3_mul(R, t->m0 + s * 9, tmp);
    mat3_mul(tmp, RT, rot);
    vo_eigenrep_from_matrix(rot, m + s * 5);
  }
}

/* vo_update_schmid for n voxels (R n x 9, m n x 60), MaterialGrid.cpp:107-124. */
void vo_update_schmid_all(long n, const double *R, double *m) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (long v = 0; v < n; v++) vo_update_schmid(R + 9 * v, m + 60 * v);
}

/* x^n for integer n >= 0, correctly rounded: binary powering in double-double
 * (products split exactly with fma), one final rounding. */
static inline void dd_mul(double ah, double al, double bh, double bl, double *rh, double *rl) {
  double p = ah * bh;
  double e = fma(ah, bh, -p);
  e += ah * bl + al * bh;
  double s = p + e;
  *rl = e - (s - p);
  *rh = s;
}

double vo_ipow(double x, int n) {
  double rh = 1.0, rl = 0.0, bh = x, bl = 0.0;
  while (n > 0) {
    if (n & 1) dd_mul(rh, rl, bh, bl, &rh, &rl);
    n >>= 1;
    if (n) dd_mul(bh, bl, bh, bl, &bh, &bl);
  }
  return rh + rl;
}

static inline double dot5(const double *a, const double *b) {
  /* InnerProduct(EigenRep, EigenRep), LinearAlgebra.cpp:800-807 */
  double r = 0;
  for (int i = 0; i < 5; i++) r += a[i] * b[i];
  return r;
}

static inline double norm5(const double *a) {
  /* EigenRep::Norm, LinearAlgebra.h:513-520 */
  return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2] + a[3] * a[3] + a[4] * a[4]);
}

/* ------------------------------------------------------------------------ */
/* Small dense linear algebra                                                */
/* ------------------------------------------------------------------------ */

/* Solve A x = b, A 5x5 column-major, by Gaussian elimination with complete
 * pivoting: the algorithm of Eigen's FullPivLU used at
 * LinearAlgebra.cpp:1171-1187.  When the remaining block is exactly zero the
 * unknowns past the rank are set to 0 (FullPivLU::solve's behaviour for a
 * rank-deficient matrix).  Returns the rank. */
int vo_lu_solve_fullpiv(const double A_[25], const double b_[5], double x[5]) {
  double A[5][5], b[5];
  int colperm[5];
  for (int i = 0; i < 5; i++) {
    for (int j = 0; j < 5; j++) A[i][j] = A_[j * 5 + i];
    b[i] = b_[i];
    colperm[i] = i;
  }
  int rank = 5;
  for (int k = 0; k < 5; k++) {
    int pr = k, pc = k;
    double best = -1.0;
    for (int j = k; j < 5; j++)     /* Eigen scans column-major */
      for (int i = k; i < 5; i++) {
        double v = fabs(A[i][j]);
        if (v > best) { best = v; pr = i; pc = j; }
      }
    if (best == 0.0) { rank = k; break; }
    if (pr != k) {
      for (int j = 0; j < 5; j++) { double t = A[k][j]; A[k][j] = A[pr][j]; A[pr][j] = t; }
      double t = b[k]; b[k] = b[pr]; b[pr] = t;
    }
    if (pc != k) {
      for (int i = 0; i < 5; i++) { double t = A[i][k]; A[i][k] = A[i][pc]; A[i][pc] = t; }
      int t = colperm[k]; colperm[k] = colperm[pc]; colperm[pc] = t;
    }
    for (int i = k + 1; i < 5; i++) {
      double f = A[i][k] / A[k][k];
      A[i][k] = f;
      for (int j = k + 1; j < 5; j++) A[i][j] -= f * A[k][j];
      b[i] -= f * b[k];
    }
  }
  double y[5] = {0, 0, 0, 0, 0};
  for (int i = rank - 1; i >= 0; i--) {
    double s = b[i];
    for (int j = i + 1; j < rank; j++) s -= A[i][j] * y[j];
    y[i] = s / A[i][i];
  }
  for (int i = 0; i < 5; i++) x[colperm[i]] = y[i];
  return rank;
}

/* In-place inverse of an n x n column-major matrix by Gauss-Jordan with
 * partial pivoting (stands in for Eigen's inverse(), MaterialGrid.cpp:248,
 * 604, 632).  Returns 0, or -1 when singular. */
int vo_inverse(int n, double *A) {
  double M[8][16];
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      M[i][j] = A[j * n + i];
      M[i][n + j] = (i == j) ? 1.0 : 0.0;
    }
  for (int k = 0; k < n; k++) {
    int p = k;
    for (int i = k + 1; i < n; i++)
      if (fabs(M[i][k]) > fabs(M[p][k])) p = i;
    if (M[p][k] == 0.0) return -1;
    if (p != k)
      for (int j = 0; j < 2 * n; j++) { double t = M[k][j]; M[k][j] = M[p][j]; M[p][j] = t; }
    double inv = 1.0 / M[k][k];
    for (int j = 0; j < 2 * n; j++) M[k][j] *= inv;
    for (int i = 0; i < n; i++) {
      if (i == k) continue;
      double f = M[i][k];
      if (f == 0.0) continue;
      for (int j = 0; j < 2 * n; j++) M[i][j] -= f * M[k][j];
    }
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) A[j * n + i] = M[i][n + j];
  return 0;
}

/* SMatrix5x5 * EigenRep, LinearAlgebra.cpp:1120-1127. */
static inline void mat5_vec(const double L[25], const double v[5], double out[5]) {
  for (int i = 0; i < 5; i++) {
    double r = 0;
    for (int j = 0; j < 5; j++) r += L[j * 5 + i] * v[j];
    out[i] = r;
  }
}

/* ------------------------------------------------------------------------ */
/* Constitutive solvers, Src/Solvers.cpp                                     */
/* ------------------------------------------------------------------------ */

/* MaterialGrid::GetMaxRSS, MaterialGrid.h:550-566 (the exit on < 1e-10 is
 * left to the caller). */
double vo_get_max_rss(const double s[5], const double m[60], const double tau[12]) {
  double mx = 0;
  for (int k = 0; k < 12; k++) {
    double rss = dot5(m + 5 * k, s);
    double v = fabs(rss / tau[k]);
    mx = v > mx ? v : mx;
  }
  return mx;
}

/* Solvers::SolveSachEquation, Solvers.cpp:230-243. */
void vo_sachs(const double m[60], const double tau[12], const double edot[5], double out[5]) {
  double mx = 0;
  for (int k = 0; k < 12; k++) {
    double rss = dot5(edot, m + 5 * k) / tau[k];
    double v = fabs(rss);
    mx = v > mx ? v : mx;
  }
  for (int i = 0; i < 5; i++) out[i] = edot[i] / mx;
}

/* Solvers::SolveConstitutiveEquations (Taylor NR), Solvers.cpp:19-101. */
void vo_taylor_nr(const double init[5], const double m[60], const double tau[12],
                  const double gdot0[12], const int nexp[12], const double edot[5],
                  double eps, double max_rss, int max_it, double out[5]) {
  double cur[5], nw[5], saved[5] = {0, 0, 0, 0, 0};
  memcpy(cur, init, sizeof cur);
  memcpy(nw, init, sizeof nw);
  int rem = max_it;
  while (rem > 0) {
    double rss[12];
    int ok;
    do {
      ok = 1;
      for (int k = 0; k < 12; k++) {
        rss[k] = dot5(m + 5 * k, cur) / tau[k];
        if (fabs(rss[k]) < 1e-10) rss[k] = 0;
        if (fabs(rss[k]) > max_rss) ok = 0;
      }
      if (!ok)
        for (int i = 0; i < 5; i++) cur[i] = saved[i] + (cur[i] - saved[i]) * 0.2;
      rem--;
      if (rem < 0) { memcpy(out, nw, sizeof nw); return; }
    } while (!ok);
    double g[12];
    for (int k = 0; k < 12; k++)
      g[k] = (nexp[k] - 1 > 0) ? gdot0[k] * vo_ipow(fabs(rss[k]), nexp[k] - 1) : gdot0[k];
    double res[5] = {0, 0, 0, 0, 0}, J[25];
    memset(J, 0, sizeof J);
    for (int k = 0; k < 12; k++) {
      const double *mk = m + 5 * k;
      for (int i = 0; i < 5; i++) res[i] += mk[i] * g[k] * rss[k];
      for (int j = 0; j < 5; j++)
        for (int i = 0; i < 5; i++)
          J[j * 5 + i] -= mk[i] * mk[j] * (double)nexp[k] * g[k] / tau[k];
    }
    for (int i = 0; i < 5; i++) res[i] = res[i] - edot[i];
    memcpy(saved, cur, sizeof cur);
    double delta[5];
    vo_lu_solve_fullpiv(J, res, delta);
    double sum[5];
    for (int i = 0; i < 5; i++) { nw[i] = cur[i] + delta[i]; sum[i] = nw[i] + cur[i]; }
    double rel = 2.0 * norm5(delta) / norm5(sum);
    memcpy(cur, nw, sizeof cur);
    if (rel < eps) break;
  }
  memcpy(out, nw, sizeof nw);
}

/* Solvers::SolveConstrainedConstitutiveEquations, Solvers.cpp:111-222.
 * *err_out is left untouched when the step budget runs out inside the
 * admissibility back-track (Solvers.cpp:166-170). */
void vo_constrained_nr(const double init[5], const double m[60], const double tau[12],
                       const double gdot0[12], const int nexp[12], const double edot[5],
                       const double lambda[5], const double dtilde[5], const double L0[25],
                       double eps, double max_rss, int max_it, double out[5], double *err_out) {
  double cur[5], saved[5];
  double nw[5] = {-1, -3, -7, -9, -11};
  double rel = 1e5;
  memcpy(cur, init, sizeof cur);
  memcpy(saved, init, sizeof saved);
  int rem = max_it;
  while (rem > 0) {
    double rss[12];
    int ok;
    do {
      ok = 1;
      for (int k = 0; k < 12; k++) {
        rss[k] = dot5(m + 5 * k, cur) / tau[k];
        if (fabs(rss[k]) > max_rss) ok = 0;
      }
      if (!ok)
        for (int i = 0; i < 5; i++) cur[i] = saved[i] + (cur[i] - saved[i]) * 0.2;
      rem--;
      if (rem < 0) { memcpy(out, cur, sizeof cur); return; }
    } while (!ok);
    double g[12];
    for (int k = 0; k < 12; k++)
      g[k] = (nexp[k] != 1) ? gdot0[k] * vo_ipow(fabs(rss[k]), nexp[k] - 1) : gdot0[k];
    double S[5] = {0, 0, 0, 0, 0}, J[25];
    memset(J, 0, sizeof J);
    for (int k = 0; k < 12; k++) {
      const double *mk = m + 5 * k;
      for (int i = 0; i < 5; i++) S[i] += mk[i] * g[k] * rss[k];
      for (int j = 0; j < 5; j++)
        for (int i = 0; i < 5; i++)
          J[j * 5 + i] += mk[i] * mk[j] * (double)nexp[k] * g[k] / tau[k];
    }
    /* J' = -I - L0 * J  (Solvers.cpp:202-204, SMatrix5x5::operator*
     * LinearAlgebra.cpp:1106-1115) */
    double Jp[25];
    for (int j = 0; j < 5; j++)
      for (int i = 0; i < 5; i++) {
        double r = 0;
        for (int k = 0; k < 5; k++) r += L0[k * 5 + i] * J[j * 5 + k];
        Jp[j * 5 + i] = ((i == j) ? -1.0 : -0.0) - r;
      }
    double t[5], Lt[5], res[5];
    for (int i = 0; i < 5; i++) t[i] = S[i] - edot[i] - dtilde[i];
    mat5_vec(L0, t, Lt);
    for (int i = 0; i < 5; i++) res[i] = cur[i] - lambda[i] + Lt[i];
    memcpy(saved, cur, sizeof cur);
    double delta[5];
    vo_lu_solve_fullpiv(Jp, res, delta);
    double diff[5], avg[5];
    for (int i = 0; i < 5; i++) {
      nw[i] = cur[i] + delta[i];
      diff[i] = nw[i] - saved[i];
      avg[i] = (nw[i] + saved[i]) * 0.5;
    }
    rel = norm5(diff) / norm5(avg);
    memcpy(cur, nw, sizeof cur);
    if (rel < eps) break;
  }
  *err_out = rel;
  memcpy(out, nw, sizeof nw);
}

static inline int sgn(double v) { return (0.0 < v) - (v < 0.0); }  /* Solvers.h:23-25 */

/* Solvers::ApplyConstitutiveEquations, Solvers.cpp:255-274. */
void vo_apply_constitutive(const double s[5], const double m[60], const double tau[12],
                           const double gdot0[12], const int nexp[12], double out[5]) {
  double e[5] = {0, 0, 0, 0, 0};
  for (int k = 0; k < 12; k++) {
    double rss = dot5(m + 5 * k, s) / tau[k];
    double g = (nexp[k] > 0) ? gdot0[k] * vo_ipow(fabs(rss), nexp[k]) : gdot0[k];
    double sg = (double)sgn(rss);
    for (int i = 0; i < 5; i++) e[i] += m[5 * k + i] * g * sg;
  }
  memcpy(out, e, sizeof e);
}

/* MaterialGrid::ShearRate + Solvers::CalculateShearRate,
 * MaterialGrid.cpp:1098-1119, Solvers.cpp:282-302. */
void vo_shear_rate(const double s[5], const double m[60], const double R[9],
                   const double tau[12], const double gdot0[12], const int nexp[12],
                   double W[9]) {
  const tables_t *t = tables();
  double RT[9], tmp[9], rot[9];
  mat3_transpose(R, RT);
  for (int e = 0; e < 9; e++) W[e] = 0;
  for (int k = 0; k < 12; k++) {
    mat3_mul(R, t->a0 + 9 * k, tmp);
    mat3_mul(tmp, RT, rot);
    double rss = dot5(m + 5 * k, s) / tau[k];
    double g = (nexp[k] > 0) ? gdot0[k] * vo_ipow(fabs(rss), nexp[k]) : gdot0[k];
    double sg = (double)sgn(rss);
    for (int e = 0; e < 9; e++) W[e] += rot[e] * g * sg;
  }
}

/* Solvers::CalculateHardening (Voce, expression as written),
 * Solvers.cpp:314-373. */
void vo_hardening(double *acc_shear, double dt, const double s[5], const double m[60],
                  const double crss_in[12], const vo_material *mat, double crss_out[12]) {
  double g[12], local = 0;
  for (int k = 0; k < 12; k++) {
    double rss = dot5(m + 5 * k, s) / crss_in[k];
    g[k] = (mat->nexp[k] > 0) ? mat->gdot0[k] * vo_ipow(fabs(rss), mat->nexp[k])
                              : mat->gdot0[k];
    local += fabs(g[k]) * dt;
  }
  double A = *acc_shear;
  for (int i = 0; i < 12; i++) {
    double dtau = 0;
    for (int j = 0; j < 12; j++) dtau += mat->hmat[i * 12 + j] * fabs(g[j]) * dt;
    double tiny = 1e-4 * mat->tau0[i];
    double voce = 0;
    if (fabs(mat->theta0[i]) > tiny) {
      voce = mat->theta1[i] * local;
      if (fabs(mat->tau1[i]) > tiny) {
        double fr = fabs(mat->theta0[i] / mat->tau1[i]);
        double e_ini = exp(-A * fr);
        double e_del = exp(-local * fr);
        voce = voce - (fr * mat->tau1[i] * mat->theta1[i]) / fr * e_ini * (e_del - 1.0) -
               mat->theta1[i] / fr * e_ini *
                   (e_del * ((A + local) * fr + 1.0) - (A * fr + 1.0));
      }
    }
    crss_out[i] = crss_in[i] + dtau * voce / local;
  }
  *acc_shear = A + local;
}

/* ------------------------------------------------------------------------ */
/* Green operator, MaterialGrid.cpp:583-615 + LinearAlgebra.h:541-599 +       */
/* Symmetrize(Tensor4Rank) LinearAlgebra.cpp:868-879                          */
/* ------------------------------------------------------------------------ */

#define T4(i, j, k, l) ((i) * 27 + (j) * 9 + (k) * 3 + (l))

/* SMatrix5x5::ToTensor4RankRep, LinearAlgebra.h:583-599. */
static void to_tensor4(const double L[25], double T[81]) {
  const double *B = tables()->B;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++)
        for (int l = 0; l < 3; l++) {
          double r = 0;
          for (int a = 0; a < 5; a++)
            for (int b = 0; b < 5; b++)
              r += L[b * 5 + a] * B[a * 9 + i * 3 + j] * B[b * 9 + k * 3 + l];
          T[T4(i, j, k, l)] = r;
        }
}

/* SMatrix5x5(const Tensor4Rank&), LinearAlgebra.h:541-554. */
static void from_tensor4(const double T[81], double S[25]) {
  const double *B = tables()->B;
  for (int a = 0; a < 5; a++)
    for (int b = 0; b < 5; b++) {
      double r = 0;
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
          for (int k = 0; k < 3; k++)
            for (int l = 0; l < 3; l++)
              r += T[T4(i, j, k, l)] * B[a * 9 + i * 3 + j] * B[b * 9 + k * 3 + l];
      S[b * 5 + a] = r;
    }
}

static void symmetrize4(const double T[81], double S[81]) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++)
        for (int l = 0; l < 3; l++)
          S[T4(i, j, k, l)] = (T[T4(i, j, k, l)] + T[T4(j, i, k, l)] + T[T4(i, j, l, k)] +
                               T[T4(j, i, l, k)]) / 4.0;
}

/* SymGreensOperator = SMatrix5x5(Symmetrize(ConstructGreensOperator(L0, k)))
 * for a non-special frequency with wave vector k (MaterialGrid.cpp:715-717). */
void vo_green_sym5(const double L0[25], const double k[3], double S5[25]) {
  double L[81], A[16], G[81], Gs[81];
  to_tensor4(L0, L);
  memset(A, 0, sizeof A);          /* Eigen column-major 4x4: A(i,k) = A[k*4+i] */
  for (int kk = 0; kk < 3; kk++)
    for (int l = 0; l < 3; l++)
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) A[kk * 4 + i] += k[l] * k[j] * L[T4(i, j, kk, l)];
  for (int i = 0; i < 3; i++) { A[i * 4 + 3] = k[i]; A[3 * 4 + i] = k[i]; }
  A[15] = 0;
  vo_inverse(4, A);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int kk = 0; kk < 3; kk++)
        for (int l = 0; l < 3; l++) G[T4(i, j, kk, l)] = -k[j] * k[l] * A[kk * 4 + i];
  symmetrize4(G, Gs);
  from_tensor4(Gs, S5);
}

/* Special planes: GreensOperator4 = L_Inv.ToTensor4RankRep() with
 * L_Inv = (-L0)^-1 (MaterialGrid.cpp:628-632, 710-717). */
void vo_special_sym5(const double L0[25], double S5[25]) {
  double Li[25], T[81], Ts[81];
  for (int i = 0; i < 25; i++) Li[i] = -L0[i];
  vo_inverse(5, Li);
  to_tensor4(Li, T);
  symmetrize4(T, Ts);
  from_tensor4(Ts, S5);
}

/* ------------------------------------------------------------------------ */
/* FFT (stands in for FFTW's unnormalised C2C; forward sign = -1)            */
/* ------------------------------------------------------------------------ */

static void fft_rec(const double *in, int istride, double *out, int n, int sign,
                    const double *tw, int N, int twstep) {
  if (n == 1) { out[0] = in[0]; out[1] = in[1]; return; }
  int p = 0;
  if (n % 4 == 0) p = 4;
  else if (n % 2 == 0) p = 2;
  else {
    for (int f = 3; f * f <= n; f += 2)
      if (n % f == 0) { p = f; break; }
    if (!p) p = n;
  }
  int m = n / p;
  for (int r = 0; r < p; r++)
    fft_rec(in + 2 * r * istride, istride * p, out + 2 * r * m, m, sign, tw, N, twstep * p);
  /* combine: X[k' + q m] = sum_r W_n^{r (k' + q m)} Y_r[k'] */
  double tmp[2 * 64];
  double *t = tmp;
  double *heap = NULL;
  if (p > 64) { heap = (double *)malloc(sizeof(double) * 2 * p); t = heap; }
  for (int kp = 0; kp < m; kp++) {
    for (int q = 0; q < p; q++) {
      int k = kp + q * m;
      double sr = 0, si = 0;
      for (int r = 0; r < p; r++) {
        long idx = ((long)r * k % n) * twstep % N;
        double wr = tw[2 * idx], wi = sign * tw[2 * idx + 1];
        double yr = out[2 * (r * m + kp)], yi = out[2 * (r * m + kp) + 1];
        sr += wr * yr - wi * yi;
        si += wr * yi + wi * yr;
      }
      t[2 * q] = sr;
      t[2 * q + 1] = si;
    }
    for (int q = 0; q < p; q++) {
      out[2 * (kp + q * m)] = t[2 * q];
      out[2 * (kp + q * m) + 1] = t[2 * q + 1];
    }
  }
  free(heap);
}

/* twiddle table: tw[t] = (cos(2 pi t / N), sin(2 pi t / N)) */
static double *make_tw(int N) {
  double *tw = (double *)malloc(sizeof(double) * 2 * N);
  for (int t = 0; t < N; t++) {
    double a = 2.0 * M_PI * (double)t / (double)N;
    tw[2 * t] = cos(a);
    tw[2 * t + 1] = -sin(a);   /* exp(-i a): forward kernel; sign flips it */
  }
  return tw;
}

/* Unnormalised DFT of length n: X[k] = sum_j x[j] exp(sign * 2 pi i j k / n)
 * with sign = -1 forward, +1 backward (FFTW's convention). */
void vo_fft1d(double *data, int n, int sign) {
  double *tw = make_tw(n);
  double *out = (double *)malloc(sizeof(double) * 2 * n);
  /* tw holds exp(-i a); multiply the imaginary part by -sign */
  for (int t = 0; t < n; t++) tw[2 * t + 1] = -tw[2 * t + 1];
  fft_rec(data, 1, out, n, sign, tw, n, 1);
  memcpy(data, out, sizeof(double) * 2 * n);
  free(out);
  free(tw);
}

static void fft_axis(double *data, int n0, int n1, int n2, int axis, int sign) {
  int len = axis == 0 ? n0 : (axis == 1 ? n1 : n2);
  long stride = axis == 0 ? (long)n1 * n2 : (axis == 1 ? n2 : 1);
  long count = (long)n0 * n1 * n2 / len;
  double *tw = make_tw(len);
  for (int t = 0; t < len; t++) tw[2 * t + 1] = -tw[2 * t + 1];
#ifdef _OPENMP
#pragma omp parallel
#endif
  {
    double *buf = (double *)malloc(sizeof(double) * 2 * len);
    double *out = (double *)malloc(sizeof(double) * 2 * len);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (long p = 0; p < count; p++) {
      long base;
      if (axis == 2) base = p * len;
      else if (axis == 1) base = (p / n2) * (long)n1 * n2 + (p % n2);
      else base = p;
      for (int j = 0; j < len; j++) {
        buf[2 * j] = data[2 * (base + j * stride)];
        buf[2 * j + 1] = data[2 * (base + j * stride) + 1];
      }
      fft_rec(buf, 1, out, len, sign, tw, len, 1);
      for (int j = 0; j < len; j++) {
        data[2 * (base + j * stride)] = out[2 * j];
        data[2 * (base + j * stride) + 1] = out[2 * j + 1];
      }
    }
    free(buf);
    free(out);
  }
  free(tw);
}

/* 3-D unnormalised C2C DFT of an interleaved complex array, row-major
 * (x, y, z), z fastest -- fftw_mpi_plan_dft_3d's layout
 * (MaterialGrid.cpp:356-375). */
void vo_fft3d(double *data, int nx, int ny, int nz, int sign) {
  fft_axis(data, nx, ny, nz, 2, sign);
  fft_axis(data, nx, ny, nz, 1, sign);
  fft_axis(data, nx, ny, nz, 0, sign);
}

/* The symmetrised 5x5 Green operator at frequency index v of the C2C grid
 * (MaterialGrid.cpp:682-731): 0 for the DC term (:733-740), the special-plane
 * operator Sspec for an index N/2 or N/2+1 along an even axis (:710-713),
 * else vo_green_sym5 at the normalised wave vector.  Returns 0 (DC) or 1. */
static int green_at(const vo_material *mat, const double L0[25], const double Sspec[25], long v, double S5[25]) {
  const int nx = mat->nx, ny = mat->ny, nz = mat->nz;
  const int cx = nx / 2, cy = ny / 2, cz = nz / 2;
  const int ex = nx % 2 == 0, ey = ny % 2 == 0, ez = nz % 2 == 0;
  int ix = (int)(v / ((long)ny * nz)), iy = (int)((v / nz) % ny), iz = (int)(v % nz);
  if (ix == 0 && iy == 0 && iz == 0) return 0;
  if ((ex && (ix == cx || ix == cx + 1)) || (ey && (iy == cy || iy == cy + 1)) ||
      (ez && (iz == cz || iz == cz + 1))) {
    memcpy(S5, Sspec, sizeof(double) * 25);
  } else {
    double k[3] = {(double)ix, (double)iy, (double)iz};
    if (ix > cx) k[0] -= nx;
    if (iy > cy) k[1] -= ny;
    if (iz > cz) k[2] -= nz;
    k[0] /= (nx * mat->lx);
    k[1] /= (ny * mat->ly);
    k[2] /= (nz * mat->lz);
    /* SVector3::Normalize with its float sqrtf, LinearAlgebra.cpp:545-559 */
    double len = (double)sqrtf((float)(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]));
    k[0] /= len; k[1] /= len; k[2] /= len;
    vo_green_sym5(L0, k, S5);
  }
  return 1;
}

/* The operator of every frequency for one L0 and one set of voxel lengths
 * (N x 25, DC rows zero), or NULL when it does not fit in memory.  Within a
 * strain step the reference rebuilds exactly these values every iteration. */
static double *green_table(const vo_material *mat, const double L0[25]) {
  const long N = (long)mat->nx * mat->ny * mat->nz;
  double *T = (double *)malloc(sizeof(double) * 25 * N);
  if (!T) return NULL;
  double Sspec[25];
  vo_special_sym5(L0, Sspec);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64)
#endif
  for (long v = 0; v < N; v++)
    if (!green_at(mat, L0, Sspec, v, T + 25 * v)) memset(T + 25 * v, 0, sizeof(double) * 25);
  return T;
}

/* StrainRateEvolutionWithFFT with the operator from table (NULL: computed). */
static void srev(const vo_material *mat, const double L0[25], const double *table, const double *phi,
                 double *dtilde) {
  const int nx = mat->nx, ny = mat->ny, nz = mat->nz;
  const long N = (long)nx * ny * nz;
  double *P[5], *D[5];
  for (int c = 0; c < 5; c++) {
    P[c] = (double *)malloc(sizeof(double) * 2 * N);
    D[c] = (double *)malloc(sizeof(double) * 2 * N);
    for (long v = 0; v < N; v++) { P[c][2 * v] = phi[c * N + v]; P[c][2 * v + 1] = 0; }
    vo_fft3d(P[c], nx, ny, nz, -1);
  }
  double Sspec[25];
  if (!table) vo_special_sym5(L0, Sspec);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64)
#endif
  for (long v = 0; v < N; v++) {
    double S5l[25];
    const double *S5 = S5l;
    if (table) {
      S5 = table + 25 * v;
      if (v == 0) { for (int c = 0; c < 5; c++) { D[c][0] = 0; D[c][1] = 0; } continue; }
    } else if (!green_at(mat, L0, Sspec, v, S5l)) {
      for (int c = 0; c < 5; c++) { D[c][0] = 0; D[c][1] = 0; }
      continue;
    }
    for (int i = 0; i < 5; i++) {
      double re = 0, im = 0;
      for (int j = 0; j < 5; j++) {
        re -= S5[j * 5 + i] * P[j][2 * v];
        im -= S5[j * 5 + i] * P[j][2 * v + 1];
      }
      D[i][2 * v] = re;
      D[i][2 * v + 1] = im;
    }
  }
  const double norm = 1.0 / (double)(nx * ny * nz);
  for (int c = 0; c < 5; c++) {
    vo_fft3d(D[c], nx, ny, nz, +1);
    for (long v = 0; v < N; v++) dtilde[c * N + v] = D[c][2 * v] * norm;
    free(P[c]);
    free(D[c]);
  }
}

/* MaterialGrid::StrainRateEvolutionWithFFT, MaterialGrid.cpp:624-761.
 * phi: 5 real components, each N long (component-major).  dtilde receives
 * Re(IFFT(-S5 * FFT(phi))) / N, the only part the reference reads
 * (MaterialGrid.h:354-363). */
void vo_strain_rate_evolution(const vo_material *mat, const double L0[25],
                              const double *phi, double *dtilde) {
  srev(mat, L0, NULL, phi, dtilde);
}

/* ------------------------------------------------------------------------ */
/* Grid-level phases                                                         */
/* ------------------------------------------------------------------------ */

static long nvox(const vo_material *mat) { return (long)mat->nx * mat->ny * mat->nz; }

/* MaterialGrid::InitializeComputation, MaterialGrid.cpp:153-210.
 * Returns -1 when a voxel's max RSS is below 1e-10 (where the reference
 * calls exit). */
int vo_initialize(const vo_material *mat, const double *m_all, const double *crss,
                  const double edot[5], double *sigma, double *lambda, double *phi) {
  const long N = nvox(mat);
  int nmin = mat->nexp[0];
  for (int k = 1; k < 12; k++) nmin = mat->nexp[k] < nmin ? mat->nexp[k] : nmin;
  double imposed = 2.0 * pow(dot5(edot, edot) / mat->gdot0[0], 1.0 / (double)(float)nmin);
  imposed = imposed > 2.0 ? imposed : 2.0;
  int bad = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 256) reduction(| : bad)
#endif
  for (long v = 0; v < N; v++) {
    const double *m = m_all + 60 * v;
    const double *tau = crss + 12 * v;
    double guess[5];
    vo_sachs(m, tau, edot, guess);
    double mrss = vo_get_max_rss(guess, m, tau);
    if (mrss > 1.1)
      for (int i = 0; i < 5; i++) guess[i] /= mrss;
    if (mrss < 1e-10) { bad = 1; continue; }
    double s[5];
    vo_taylor_nr(guess, m, tau, mat->gdot0, mat->nexp, edot, 5e-4, imposed, 1000, s);
    for (int i = 0; i < 5; i++) {
      sigma[5 * v + i] = s[i];
      lambda[5 * v + i] = s[i];
      phi[i * N + v] = s[i];
    }
  }
  return bad ? -1 : 0;
}

/* MaterialGrid::BuildLinearRefMediumStiffness, MaterialGrid.cpp:219-282
 * (single rank: the slab weighting multiplies and divides by Nx).
 * Returns 0, or -1 when a voxel's secant modulus is exactly singular: the
 * reference's Eigen inverse (:248) has no answer there, and averaging the
 * un-inverted modulus would hide it.  L0 is then left untouched. */
int vo_build_L0(const vo_material *mat, const double *m_all, const double *crss,
                const double *sigma, double L0[25]) {
  const long N = nvox(mat);
  double *inv = (double *)malloc(sizeof(double) * 25 * N);
  int bad = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) reduction(| : bad)
#endif
  for (long v = 0; v < N; v++) {
    const double *m = m_all + 60 * v;
    const double *tau = crss + 12 * v;
    double M[25];
    memset(M, 0, sizeof M);
    for (int k = 0; k < 12; k++) {
      double rss = dot5(m + 5 * k, sigma + 5 * v) / tau[k];
      double g = (mat->nexp[k] - 1 > 0)
                     ? mat->gdot0[k] * vo_ipow(fabs(rss), mat->nexp[k] - 1)
                     : mat->gdot0[k];
      for (int j = 0; j < 5; j++)
        for (int i = 0; i < 5; i++) M[j * 5 + i] += m[5 * k + i] * m[5 * k + j] * g / tau[k];
    }
    if (vo_inverse(5, M) != 0) bad = 1;
    memcpy(inv + 25 * v, M, sizeof M);
  }
  if (bad) {
    free(inv);
    return -1;
  }
  double A[25];
  memset(A, 0, sizeof A);
  for (long v = 0; v < N; v++)           /* SMatrix5x5::MovingAverage, LinearAlgebra.cpp:981-988 */
    for (int e = 0; e < 25; e++) A[e] += (inv[25 * v + e] - A[e]) / (double)(v + 1);
  free(inv);
  for (int e = 0; e < 25; e++) {
    double w = A[e] * (double)mat->nx;    /* MaterialGrid.cpp:257-271 */
    A[e] = w / (double)mat->nx;
    if (fabs(A[e]) < 1e-10) A[e] = 0;     /* NormalizeValues, LinearAlgebra.h:605-613 */
    L0[e] = A[e];
  }
  return 0;
}

/* Per-voxel body of the fixed-point loop, MaterialGrid.cpp:946-1003.
 * sums12 = {sum sigma (5), sum edot (5), sum |sigma-lambda|, sum |edot-E-d|},
 * accumulated in voxel order. */
int vo_voxel_update(const vo_material *mat, const double *m_all, const double *crss,
                    const double edot[5], const double L0[25], const double *dtilde,
                    double *sigma, double *lambda, double *phi, double *sums12,
                    double *max_nr_error) {
  const long N = nvox(mat);
  double *er = (double *)malloc(sizeof(double) * 8 * N);  /* edot(5), |ds|, |de|, nr err */
  int bad = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 256) reduction(| : bad)
#endif
  for (long v = 0; v < N; v++) {
    const double *m = m_all + 60 * v;
    const double *tau = crss + 12 * v;
    double d[5], s[5], lam[5], e[5];
    for (int i = 0; i < 5; i++) { d[i] = dtilde[i * N + v]; s[i] = sigma[5 * v + i]; lam[i] = lambda[5 * v + i]; }
    double mrss = vo_get_max_rss(s, m, tau);
    if (mrss < 1e-10) { bad = 1; continue; }
    for (int i = 0; i < 5; i++) s[i] /= mrss;
    double nr_err = 0;
    vo_constrained_nr(s, m, tau, mat->gdot0, mat->nexp, edot, lam, d, L0, 5e-4, 1e9, 100, s, &nr_err);
    vo_apply_constitutive(s, m, tau, mat->gdot0, mat->nexp, e);
    double Ld[5], t[5], Lt[5];
    mat5_vec(L0, d, Ld);
    for (int i = 0; i < 5; i++) phi[i * N + v] = s[i] - Ld[i];        /* :982-984 */
    for (int i = 0; i < 5; i++) t[i] = edot[i] + d[i] - e[i];          /* :1177-1181 */
    mat5_vec(L0, t, Lt);
    for (int i = 0; i < 5; i++) lam[i] += Lt[i];                       /* :1183 */
    double se[5], ee[5];
    for (int i = 0; i < 5; i++) {
      ee[i] = e[i] - edot[i] - d[i];
      se[i] = s[i] - lam[i];
      sigma[5 * v + i] = s[i];
      lambda[5 * v + i] = lam[i];
      er[8 * v + i] = e[i];
    }
    er[8 * v + 5] = sqrt(dot5(se, se));
    er[8 * v + 6] = sqrt(dot5(ee, ee));
    er[8 * v + 7] = nr_err;
  }
  if (!bad) {
    double S[12];
    memset(S, 0, sizeof S);
    double mx = 0;
    for (long v = 0; v < N; v++) {
      for (int i = 0; i < 5; i++) { S[i] += sigma[5 * v + i]; S[5 + i] += er[8 * v + i]; }
      S[10] += er[8 * v + 5];
      S[11] += er[8 * v + 6];
      mx = er[8 * v + 7] > mx ? er[8 * v + 7] : mx;
    }
    memcpy(sums12, S, sizeof S);
    if (max_nr_error) *max_nr_error = mx;
  }
  free(er);
  return bad ? -1 : 0;
}

/* Epilogue, MaterialGrid.cpp:1074-1079: OrientationEvolutionWithFFT
 * (:769-876), UpdateVoxelLength (:1190-1196), UpdateHardeningCRSS
 * (:1127-1157).
 *
 * The FFT inside OrientationEvolutionWithFFT contributes exactly zero:
 * its result is read back as an EigenRep whose ToMatrixRep() is an exactly
 * symmetric matrix (each entry is the same sum of the same products as its
 * transpose, LinearAlgebra.h:352-360), and AntiSymmetrize() of an exactly
 * symmetric matrix is exactly 0 (LinearAlgebra.h:83-90).  So the update is
 * R += dt * ShearRate with the Schmid tensors of the old R, followed by
 * UpdateSchmidtTensors on the new R. */
void vo_epilogue(vo_material *mat, double *R, double *crss, double *acc_shear,
                 const double *sigma, const double edot[5], double dt) {
  const long N = nvox(mat);
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (long v = 0; v < N; v++) {
    double m[60], W[9], mn[60], cn[12];
    vo_update_schmid(R + 9 * v, m);
    vo_shear_rate(sigma + 5 * v, m, R + 9 * v, crss + 12 * v, mat->gdot0, mat->nexp, W);
    for (int e = 0; e < 9; e++) R[9 * v + e] += W[e] * dt;
    vo_update_schmid(R + 9 * v, mn);
    vo_hardening(acc_shear + v, dt, sigma + 5 * v, mn, crss + 12 * v, mat, cn);
    memcpy(crss + 12 * v, cn, sizeof cn);
  }
  double E[9];
  vo_matrix_from_eigenrep(edot, E);
  mat->lx *= (1.0 + E[0] * dt);
  mat->ly *= (1.0 + E[4] * dt);
  mat->lz *= (1.0 + E[8] * dt);
}

/* MaterialGrid::RunSingleStrainStep, MaterialGrid.cpp:897-1084. */
int vo_strain_step(vo_material *mat, double *R, double *crss, double *acc_shear,
                   double *sigma, double *lambda, double *phi, double *dtilde,
                   const double edot[5], int max_iter, double dt, double tol,
                   int do_epilogue, vo_step_info *info) {
  const long N = nvox(mat);
  double *m_all = (double *)malloc(sizeof(double) * 60 * N);
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (long v = 0; v < N; v++) vo_update_schmid(R + 9 * v, m_all + 60 * v);
  memset(info, 0, sizeof *info);
  if (vo_initialize(mat, m_all, crss, edot, sigma, lambda, phi) != 0) {
    info->status = -1;
    free(m_all);
    return -1;
  }
  double L0[25];
  if (vo_build_L0(mat, m_all, crss, sigma, L0) != 0) {
    info->status = -2;
    free(m_all);
    return -2;
  }
  memcpy(info->L0, L0, sizeof L0);
  int steps = 0;
  double es = 0, ee = 0, avg[5] = {0, 0, 0, 0, 0}, avge[5] = {0, 0, 0, 0, 0};
  const double nd = (double)(mat->nx * mat->ny * mat->nz);
  double *table = green_table(mat, L0);   /* same values every iteration; NULL: recomputed */
  do {
    srev(mat, L0, table, phi, dtilde);
    double S[12], mx = 0;
    if (vo_voxel_update(mat, m_all, crss, edot, L0, dtilde, sigma, lambda, phi, S, &mx) != 0) {
      info->status = -1;
      free(m_all);
      free(table);
      return -1;
    }
    es = S[10] / nd;
    ee = S[11] / nd;
    for (int i = 0; i < 5; i++) { avg[i] = S[i] / nd; avge[i] = S[5 + i] / nd; }
    es /= sqrt(dot5(avg, avg));
    ee /= sqrt(dot5(edot, edot));
    info->max_nr_error = mx;
    steps++;
  } while (steps < max_iter && (es > tol || ee > tol));
  info->iterations = steps;
  info->err_stress = es;
  info->err_strain_rate = ee;
  memcpy(info->avg_stress, avg, sizeof avg);
  memcpy(info->avg_strain_rate, avge, sizeof avge);
  free(m_all);
  free(table);
  if (do_epilogue) vo_epilogue(mat, R, crss, acc_shear, sigma, edot, dt);
  return 0;
}
