/*
 * vpfft_oracle.h -- CPU FP64 restatement of exmatex/VPFFT's strain step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the GPU
 * library is compared against; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  Nothing in vpfft_amd/ links or
 * calls it.
 *
 * Layouts follow the reference's host types (Src/LinearAlgebra.h):
 *   EigenRep    = double[5]                      (LinearAlgebra.h:298-300)
 *   SMatrix3x3  = double[3][3], row-major        (LinearAlgebra.h:114)
 *   SMatrix5x5  = double[25], column-major m[j*5+i]  (LinearAlgebra.h:574-578)
 *   voxel index = x*Ny*Nz + y*Nz + z             (MaterialGrid.h:121-124)
 * Per-voxel fields are arrays of those structs (AoS), the FFT-side fields
 * (polarization, strain-rate variation) are one real array per component
 * (SoA), as the reference keeps one FFTW buffer per component
 * (MaterialGrid.h:320-321).
 */
#ifndef VPFFT_ORACLE_H
#define VPFFT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define VO_NSLIP 12

typedef struct {
  int nx, ny, nz;
  double lx, ly, lz;            /* voxel lengths, MaterialGrid.h:258-260 */
  double gdot0[VO_NSLIP];       /* GammaDotBase, MaterialGrid.h:271 */
  int    nexp[VO_NSLIP];        /* RateSensitivity, MaterialGrid.h:272 */
  double tau0[VO_NSLIP], tau1[VO_NSLIP];       /* Voce, MaterialGrid.h:279-282 */
  double theta0[VO_NSLIP], theta1[VO_NSLIP];
  double hmat[VO_NSLIP * VO_NSLIP];            /* HardeningMatrix[i][j] row-major */
} vo_material;

typedef struct {
  int iterations;
  int status;                   /* 0 ok, -1 degenerate (max RSS < 1e-10), -2 singular secant modulus */
  double err_stress, err_strain_rate;
  double avg_stress[5], avg_strain_rate[5];
  double L0[25];
  double max_nr_error;
} vo_step_info;

/* ---- crystal + basis tables ---- */
void vo_eigen_basis(double B[6 * 9]);
void vo_schmid_basis(double sym[VO_NSLIP * 9], double anti[VO_NSLIP * 9]);
void vo_eigenrep_from_matrix(const double M[9], double v[5]);
void vo_matrix_from_eigenrep(const double v[5], double M[9]);
void vo_update_schmid(const double R[9], double m[VO_NSLIP * 5]);
void vo_update_schmid_all(long n, const double *R, double *m);

/* ---- small dense linear algebra ---- */
double vo_ipow(double x, int n);
int vo_lu_solve_fullpiv(const double A[25], const double b[5], double x[5]);
int vo_inverse(int n, double *A);   /* column-major n x n, in place */

/* ---- per-voxel constitutive solvers (Src/Solvers.cpp) ---- */
double vo_get_max_rss(const double s[5], const double m[60], const double tau[12]);
void vo_sachs(const double m[60], const double tau[12], const double edot[5], double out[5]);
void vo_taylor_nr(const double init[5], const double m[60], const double tau[12],
                  const double gdot0[12], const int nexp[12], const double edot[5],
                  double eps, double max_rss, int max_it, double out[5]);
void vo_constrained_nr(const double init[5], const double m[60], const double tau[12],
                       const double gdot0[12], const int nexp[12], const double edot[5],
                       const double lambda[5], const double dtilde[5], const double L0[25],
                       double eps, double max_rss, int max_it, double out[5], double *err_out);
void vo_apply_constitutive(const double s[5], const double m[60], const double tau[12],
                           const double gdot0[12], const int nexp[12], double out[5]);
void vo_shear_rate(const double s[5], const double m[60], const double R[9],
                   const double tau[12], const double gdot0[12], const int nexp[12],
                   double W[9]);
void vo_hardening(double *acc_shear, double dt, const double s[5], const double m[60],
                  const double crss_in[12], const vo_material *mat, double crss_out[12]);

/* ---- Green operator + FFT (Src/MaterialGrid.cpp:583-761) ---- */
void vo_green_sym5(const double L0[25], const double k[3], double S5[25]);
void vo_special_sym5(const double L0[25], double S5[25]);
void vo_fft3d(double *data, int nx, int ny, int nz, int sign);
void vo_fft1d(double *data, int n, int sign);
void vo_strain_rate_evolution(const vo_material *mat, const double L0[25],
                              const double *phi, double *dtilde);

/* ---- grid-level phases ---- */
int vo_initialize(const vo_material *mat, const double *m_all, const double *crss,
                  const double edot[5], double *sigma, double *lambda, double *phi);
/* 0, or -1 when a voxel's secant modulus is exactly singular (L0 untouched) */
int vo_build_L0(const vo_material *mat, const double *m_all, const double *crss,
                const double *sigma, double L0[25]);
int vo_voxel_update(const vo_material *mat, const double *m_all, const double *crss,
                    const double edot[5], const double L0[25], const double *dtilde,
                    double *sigma, double *lambda, double *phi, double *sums12,
                    double *max_nr_error);
void vo_epilogue(vo_material *mat, double *R, double *crss, double *acc_shear,
                 const double *sigma, const double edot[5], double dt);
int vo_strain_step(vo_material *mat, double *R, double *crss, double *acc_shear,
                   double *sigma, double *lambda, double *phi, double *dtilde,
                   const double edot[5], int max_iter, double dt, double tol,
                   int do_epilogue, vo_step_info *info);

int vo_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
