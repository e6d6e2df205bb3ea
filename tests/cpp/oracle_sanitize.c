/* oracle_sanitize.c -- drives every phase of the CPU oracle
 * (oracle/vpfft_oracle.c) on small, ragged inputs, for a build with
 * AddressSanitizer and UndefinedBehaviorSanitizer (tests/test_oracle_sanitize.py):
 * two strain steps with the epilogue on a 10 x 6 x 7 polycrystal (odd and
 * even axes, so both special-plane branches and the mixed-radix FFT run),
 * the per-voxel solvers, the Green operator and the FFT on a 15 x 9 x 7
 * grid.  Exits 0 when every result is finite. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "vpfft_oracle.h"

static unsigned long long state = 12345;
static double u01(void) {
  state = state * 6364136223846793005ULL + 1442695040888963407ULL;
  return (double)(state >> 11) / 9007199254740992.0;
}

static void rotation(double R[9]) {
  double q[4], s = 0;
  for (int i = 0; i < 4; i++) {
    q[i] = u01() - 0.5;
    s += q[i] * q[i];
  }
  s = sqrt(s);
  const double w = q[0] / s, x = q[1] / s, y = q[2] / s, z = q[3] / s;
  const double r[9] = {1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
                       2 * (x * y + z * w),     1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
                       2 * (x * z - y * w),     2 * (y * z + x * w),     1 - 2 * (x * x + y * y)};
  memcpy(R, r, sizeof r);
}

static int finite_all(const double* a, long n, const char* what) {
  for (long i = 0; i < n; i++)
    if (!isfinite(a[i])) {
      fprintf(stderr, "%s[%ld] not finite\n", what, i);
      return 0;
    }
  return 1;
}

int main(void) {
  vo_material mat;
  memset(&mat, 0, sizeof mat);
  mat.nx = 10, mat.ny = 6, mat.nz = 7;
  mat.lx = mat.ly = mat.lz = 1.0;
  for (int k = 0; k < VO_NSLIP; k++) {
    mat.gdot0[k] = 1.0;
    mat.nexp[k] = 10;
    mat.tau0[k] = mat.tau1[k] = 1.0;
    mat.theta0[k] = 20.0;
    mat.theta1[k] = 2.0;
    for (int j = 0; j < VO_NSLIP; j++) mat.hmat[k * VO_NSLIP + j] = 1.0;
  }
  const long n = (long)mat.nx * mat.ny * mat.nz;
  double* R = malloc(sizeof(double) * 9 * n);
  double* crss = malloc(sizeof(double) * 12 * n);
  double* acc = calloc(n, sizeof(double));
  double* sigma = malloc(sizeof(double) * 5 * n);
  double* lambda = malloc(sizeof(double) * 5 * n);
  double* phi = malloc(sizeof(double) * 5 * n);
  double* dtilde = malloc(sizeof(double) * 5 * n);
  double grains[4][9];
  for (int g = 0; g < 4; g++) rotation(grains[g]);
  for (long v = 0; v < n; v++) {
    memcpy(R + 9 * v, grains[(v / 13) % 4], sizeof(double) * 9);
    for (int k = 0; k < 12; k++) crss[12 * v + k] = 1.0;
  }
  const double edot[5] = {-1.0606601717798212, -0.61237243569579458, 0, 0, 0};
  int ok = 1;
  for (int step = 0; step < 2 && ok; step++) {
    vo_step_info info;
    const int rc = vo_strain_step(&mat, R, crss, acc, sigma, lambda, phi, dtilde, edot, 30, 0.005, 1e-5, 1, &info);
    printf("step %d rc %d iterations %d avg %.17g\n", step, rc, info.iterations, info.avg_stress[0]);
    ok = rc == 0 && finite_all(sigma, 5 * n, "sigma") && finite_all(R, 9 * n, "R") &&
         finite_all(crss, 12 * n, "crss");
  }
  /* the per-voxel solvers on one voxel */
  double m[60], tau[12], s0[5], s1[5], e5[5], W[9], tau2[12], A = 0.1, err = 7.0;
  vo_update_schmid(R, m);
  for (int k = 0; k < 12; k++) tau[k] = 1.0 + 0.01 * k;
  vo_sachs(m, tau, edot, s0);
  vo_taylor_nr(s0, m, tau, mat.gdot0, mat.nexp, edot, 5e-4, 2.0, 1000, s1);
  double L0[25] = {0};
  for (int i = 0; i < 5; i++) L0[6 * i] = 2.0;
  const double dt5[5] = {0.01, 0, -0.01, 0.02, 0};
  vo_constrained_nr(s1, m, tau, mat.gdot0, mat.nexp, edot, s1, dt5, L0, 5e-4, 1e9, 100, s0, &err);
  vo_apply_constitutive(s0, m, tau, mat.gdot0, mat.nexp, e5);
  vo_shear_rate(s0, m, R, tau, mat.gdot0, mat.nexp, W);
  vo_hardening(&A, 0.005, s0, m, tau, &mat, tau2);
  ok = ok && finite_all(s0, 5, "constrained") && finite_all(e5, 5, "apply") && finite_all(W, 9, "W") &&
       finite_all(tau2, 12, "hardening");
  /* Green operator and a ragged FFT round trip */
  const double k3[3] = {0.3, -0.2, 0.7};
  double S5[25];
  vo_green_sym5(L0, k3, S5);
  vo_special_sym5(L0, S5 + 0);
  const int fx = 15, fy = 9, fz = 7;
  double* f = malloc(sizeof(double) * 2 * fx * fy * fz);
  for (int i = 0; i < 2 * fx * fy * fz; i++) f[i] = u01();
  const double f0 = f[5];
  vo_fft3d(f, fx, fy, fz, -1);
  vo_fft3d(f, fx, fy, fz, 1);
  ok = ok && fabs(f[5] / (fx * fy * fz) - f0) < 1e-12;
  printf("ok %d\n", ok);
  free(R), free(crss), free(acc), free(sigma), free(lambda), free(phi), free(dtilde), free(f);
  return ok ? 0 : 1;
}
