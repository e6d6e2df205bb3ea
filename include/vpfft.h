/*
 * vpfft.h -- C ABI of the MI355X VPFFT strain-step engine (libvpfft.so).
 *
 * Drop-in boundary for the hot path of exmatex/VPFFT: the strain step of
 * VPFFT::DataStructures::MaterialGrid (Src/MaterialGrid.h:59-577) and the
 * per-voxel solvers of VPFFT::Solvers (Src/Solvers.h:51-153).  Plain C types
 * only: pointers, sizes, doubles.  Host arrays are borrowed for the duration
 * of a call and copied; the grid handle owns all device memory.
 *
 * Host array layouts are the reference's (Src/LinearAlgebra.h):
 *   orientation  N x 9  row-major SMatrix3x3 per voxel     (LinearAlgebra.h:114)
 *   stress, ...  N x 5  EigenRep per voxel                 (LinearAlgebra.h:298-300)
 *   crss         N x 12 per voxel, slip order of BurgerVectorBasis.h:16-72
 *   5x5 matrices 25 doubles column-major m[j*5+i]          (LinearAlgebra.h:574-575)
 *   voxel index  x*Ny*Nz + y*Nz + z                        (MaterialGrid.h:121-124)
 *   polarization / strain-rate variation: 5 x N, one real array per
 *   component (MaterialGrid.h:320-321, real parts only).
 *
 * Errors: every entry point returns 0 on success or a negative VPFFT_E_*
 * code; vpfft_last_error() gives a message (thread-local).  The reference
 * exits the process with status 0 on these conditions (Src/Error.h:25-32,
 * MaterialGrid.h:560-564).
 *
 * Threading: calls on one handle must be serialised by the caller; distinct
 * handles may be used from distinct host threads.  One HIP stream per handle.
 */
#ifndef VPFFT_H
#define VPFFT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VPFFT_NSLIP 12

enum {
  VPFFT_OK = 0,
  VPFFT_E_INVALID = -1,     /* bad argument */
  VPFFT_E_NOMEM = -2,       /* device allocation failed */
  VPFFT_E_DEVICE = -3,      /* HIP runtime error */
  VPFFT_E_DEGENERATE = -4,  /* a voxel's max resolved shear stress < 1e-10 (MaterialGrid.h:560) */
  VPFFT_E_NONFINITE = -5,   /* NaN/Inf in the solution */
  VPFFT_E_UNSUPPORTED = -6, /* grid edge with a prime factor other than 2,3,5,7 */
  VPFFT_E_STATE = -7        /* call out of order (e.g. step before upload) */
};

enum { VPFFT_FP64 = 0, VPFFT_MIXED = 1 };

/* Field identifiers for vpfft_grid_download / vpfft_grid_upload_field. */
enum {
  VPFFT_FIELD_ORIENTATION = 0,   /* N x 9  (OrientationField, MaterialGrid.h:289) */
  VPFFT_FIELD_CRSS = 1,          /* N x 12 (LocalCRSS, MaterialGrid.h:270) */
  VPFFT_FIELD_STRESS = 2,        /* N x 5  (StressField, MaterialGrid.h:290) */
  VPFFT_FIELD_LAGRANGE = 3,      /* N x 5  (LagrangeMultiplierField, MaterialGrid.h:291) */
  VPFFT_FIELD_POLARIZATION = 4,  /* 5 x N  real part of _PolarizationField after the last iteration */
  VPFFT_FIELD_STRAIN_RATE_VARIATION = 5, /* 5 x N  d~ of the last iteration (DisplacementVariation real part) */
  VPFFT_FIELD_ACC_SHEAR = 6,     /* N      (AccumulatedShear, MaterialGrid.h:293) */
  VPFFT_FIELD_SCHMID = 7,        /* N x 60 (LocalSchmidtTensor, MaterialGrid.h:292), download only */
  VPFFT_FIELD_STRAIN_RATE = 8    /* N x 5  Solvers::ApplyConstitutiveEquations (Solvers.cpp:255-274) of the
                                    current stress, orientation and CRSS at every voxel, download only */
};

typedef struct {
  int nx, ny, nz;               /* MaterialGrid(int,int,int), MaterialGrid.h:72-78 */
  int precision;                /* VPFFT_FP64 or VPFFT_MIXED (FP32 FFT + Green, FP64 Newton) */
} vpfft_grid_desc;

/* Material constants.  The reference hard-codes them in InitializeMemory
 * (MaterialGrid.cpp:322-336, 390); here they are inputs. */
typedef struct {
  double gamma_dot_base[VPFFT_NSLIP];   /* GammaDotBase */
  int rate_sensitivity[VPFFT_NSLIP];    /* RateSensitivity (integer exponent n) */
  double tau0[VPFFT_NSLIP], tau1[VPFFT_NSLIP];       /* Voce hardening */
  double theta0[VPFFT_NSLIP], theta1[VPFFT_NSLIP];
  double hardening[VPFFT_NSLIP * VPFFT_NSLIP];      /* HardeningMatrix[i][j], row-major */
} vpfft_material;

typedef struct {
  int iterations;          /* fixed-point iterations taken (StepsTaken, MaterialGrid.cpp:1055) */
  int converged;           /* 1 if both errors <= tolerance at exit */
  int status;              /* same as the return code */
  int precision;
  double err_stress, err_strain_rate;          /* MaterialGrid.cpp:1052-1053 */
  double avg_stress[5], avg_strain_rate[5];    /* last iteration's averages */
  double L0[25];                               /* reference medium, column-major */
  double max_nr_error;                         /* max per-voxel NR error of the last iteration */
  double nr_steps_last;                        /* Newton steps summed over voxels, last iteration */
  double nr_steps_total;                       /* ... summed over all iterations */
  double ms_prologue, ms_loop, ms_epilogue;    /* device time (HIP events) */
  double max_nr_steps;                         /* largest per-voxel Newton step count, last iteration */
  double nr_wave_steps_total;                  /* sum over iterations and 64-voxel waves of 64 x the wave's
                                                  largest step count: nr_steps_total / this = the Newton
                                                  loop's lane efficiency */
  double taylor_steps_total;                   /* Newton steps of the Taylor initialisation, summed over voxels */
  double err_ratio_5;                          /* max(err_stress, err_strain_rate) of the last iteration
                                                  over the same five iterations earlier (1 with fewer):
                                                  near 1 = the fixed point has stalled */
} vpfft_step_info;

/* ---- grid lifetime: MaterialGrid(int,int,int) + InitializeMemory (MaterialGrid.cpp:312-394) */
int vpfft_grid_create(const vpfft_grid_desc* desc, const vpfft_material* mat, void** out_grid);
/* no reference counterpart (the reference never frees; Cleanup is never called, MaterialGrid.cpp:287-305) */
int vpfft_grid_destroy(void* grid);

/* MaterialGrid::SetVoxelLength, MaterialGrid.h:97-102 */
int vpfft_grid_set_voxel_length(void* grid, double lx, double ly, double lz);
int vpfft_grid_get_voxel_length(void* grid, double out_lxyz[3]);

/* precision of the next strain steps (VPFFT_FP64 / VPFFT_MIXED) */
int vpfft_grid_set_precision(void* grid, int precision);

/* MaterialGrid::SendDataAndInitialize(SMatrix3x3*, EigenRep*), MaterialGrid.cpp:425-534,
 * plus the CRSS the reference fills with a constant (MaterialGrid.cpp:390).
 * orientation: N x 9 (required); crss: N x 12 (NULL: every entry = crss_value);
 * stress: N x 5 (NULL allowed; the reference's value is overwritten by
 * InitializeComputation before use, MaterialGrid.cpp:192-202).
 * Accumulated shear is reset to 0. */
int vpfft_grid_upload(void* grid, const double* orientation, const double* crss, double crss_value,
                      const double* stress);

/* Accessors by field id (MaterialGrid.h:129-185, 338-424): copy into / out of host arrays. */
int vpfft_grid_upload_field(void* grid, int field, const double* host);
int vpfft_grid_download(void* grid, int field, double* host);

/* MaterialGrid::RunSingleStrainStep(const EigenRep&, int, Float, Float) -> EigenRep,
 * MaterialGrid.h:221-223, MaterialGrid.cpp:897-1084.
 * edot: macroscopic strain rate (EigenRep); avg_stress (out, may be NULL):
 * the returned <sigma>; info (out, may be NULL).  A step that reaches
 * max_iter without meeting tol returns VPFFT_OK with info->converged = 0
 * (the reference is silent about it).  The fixed-point loop runs as a HIP
 * graph of VPFFT_GRAPH_ITERS iterations (environment, read at grid creation;
 * default 4, 0 = launch every kernel from the host) replayed until the
 * device-side convergence test stops it; both forms give the same bits. */
int vpfft_run_strain_step(void* grid, const double edot[5], int max_iter, double dt, double tol,
                          double avg_stress[5], vpfft_step_info* info);

/* MaterialGrid::RunVPFFT, MaterialGrid.cpp:1202-1228: num_steps strain steps;
 * avg_stress_out: num_steps x 5 (may be NULL); infos: num_steps (may be NULL).
 * stress_strain_path: when non-NULL, writes the reference's Stress_Strain.txt
 * lines (step, E(5), <sigma>(5), |E|, |<sigma>|). */
int vpfft_run_vpfft(void* grid, const double edot[5], int max_iter, int num_steps, double dt, double tol,
                    double* avg_stress_out, vpfft_step_info* infos, const char* stress_strain_path);

/* MaterialGrid::StrainRateEvolutionWithFFT(const SMatrix5x5&), MaterialGrid.cpp:624-761:
 * d~ = Re IFFT(-S5(k) FFT(phi)) / N with the given L0 (column-major 25).
 * phi_in and dtilde_out are 5 x N host arrays.  Uses the grid's buffers
 * (overwrites the polarization field). */
int vpfft_strain_rate_evolution(void* grid, const double L0[25], const double* phi_in, double* dtilde_out);

/* The same transform without a grid and without allocating device memory
 * (no reference counterpart; for callers that cannot allocate -- code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * d_phi and d_dtilde are device arrays, 5 x N doubles each, N = nx*ny*nz
 * (precision VPFFT_MIXED transforms in FP32 like a mixed grid).  The per-call
 * constants, twiddles and half spectrum live in a static device scratch array
 * of 32 MiB (each GPU the library runs on holds one until the process ends):
 * a grid whose parts do not fit (larger than about 64^3) is refused with
 * VPFFT_E_INVALID.  Runs on `stream` (NULL: the default stream) and returns
 * when the stream has finished it.  Calls on one GPU share its scratch array:
 * calls that overlap on one GPU (several host threads or streams) race. */
int vpfft_strain_rate_evolution_noalloc(int nx, int ny, int nz, const double voxel_length[3], int precision,
                                        const double L0[25], const double* d_phi, double* d_dtilde, void* stream);

/* Solvers::SolveConstrainedConstitutiveEquations over a batch of n voxels
 * (Solvers.h:90-102, Solvers.cpp:111-222).  Per voxel (AoS): orientation
 * n x 9 (Schmid tensors are R m0 R^T, MaterialGrid.cpp:121), crss n x 12,
 * stress_in / lambda / dtilde n x 5; L0 column-major 25; edot 5.
 * stress_out n x 5; nr_error n (in: value kept when the step budget runs
 * out, Solvers.cpp:166-170; out: final relative error); nr_steps n (may be NULL). */
int vpfft_solve_constrained_batch(long n, const vpfft_material* mat, const double* orientation, const double* crss,
                                  const double L0[25], const double edot[5], const double* stress_in,
                                  const double* lambda, const double* dtilde, double* stress_out,
                                  double* nr_error, int* nr_steps);

/* Solvers::SolveConstitutiveEquations (Taylor NR, Solvers.h:55-65, Solvers.cpp:19-101)
 * over a batch, with the reference's InitializeComputation constants
 * (epsilon 5e-4, 1000 steps, the imposed maximum resolved stress of
 * MaterialGrid.cpp:162-166). */
int vpfft_solve_taylor_batch(long n, const vpfft_material* mat, const double* orientation, const double* crss,
                             const double edot[5], const double* stress_in, double* stress_out);

/* VPFFT::Solvers (Solvers.h:51-153) over n independent voxels, on the GPU,
 * with the arguments the reference's per-voxel functions take.  Host
 * buffers, copied in and out; the call returns when the results are back.
 * Per voxel (AoS): schmid n x 12 x 5 -- the SchmidtTensors vector, sample
 * frame; crss n x 12.  One strain rate edot (EigenRep) serves the whole
 * batch, as one macroscopic strain rate serves every voxel of a grid.  mat
 * gives GammaDotBase and RateSensitivity (and the Voce parameters and the
 * hardening matrix for vpfft_solvers_hardening).  Arithmetic follows the
 * reference operation for operation; x^n with an integer n is the double
 * rounded once from the exact power (std::pow's glibc result agrees but for
 * rare halfway cases), and exp is the device library's.  n = 0 does nothing. */
/* SolveConstitutiveEquations (Solvers.h:51-59, Solvers.cpp:19-101). */
int vpfft_solvers_taylor(long n, const vpfft_material* mat, const double* schmid, const double* crss,
                         const double edot[5], const double* stress_in, double eps, double max_rss, int max_it,
                         double* stress_out);
/* SolveConstrainedConstitutiveEquations (Solvers.h:67-79, Solvers.cpp:111-222):
 * lambda, dtilde, stress_in n x 5; L0 the HomogeonousReference, column-major
 * 25, used as given.  nr_error n: NR_Error_Out, in and out -- left as it came
 * in when the step budget runs out inside the admissibility back-track
 * (Solvers.cpp:163-170). */
int vpfft_solvers_constrained(long n, const vpfft_material* mat, const double* schmid, const double* crss,
                              const double edot[5], const double* lambda, const double* dtilde, const double L0[25],
                              const double* stress_in, double eps, double max_rss, int max_it, double* stress_out,
                              double* nr_error);
/* SolveSachEquation (Solvers.h:88-90, Solvers.cpp:230-243). */
int vpfft_solvers_sachs(long n, const double* schmid, const double* crss, const double edot[5], double* out);
/* ApplyConstitutiveEquations (Solvers.h:106-110, Solvers.cpp:255-274): stress n x 5 -> strain rate n x 5. */
int vpfft_solvers_apply(long n, const vpfft_material* mat, const double* schmid, const double* crss,
                        const double* stress, double* strain_rate_out);
/* CalculateShearRate (Solvers.h:120-125, Solvers.cpp:282-302): anti n x 12 x 9,
 * the AntiSchmidtTensors (SMatrix3x3, row-major) in the frame of schmid;
 * shear_rate_out n x 9 (SMatrix3x3). */
int vpfft_solvers_shear_rate(long n, const vpfft_material* mat, const double* schmid, const double* anti,
                             const double* crss, const double* stress, double* shear_rate_out);
/* CalculateHardening (Solvers.h:136-147, Solvers.cpp:314-373): accumulated_shear
 * n, in and out (AccumulatedShear); crss_out n x 12 (the updated CRSS). */
int vpfft_solvers_hardening(long n, const vpfft_material* mat, const double* schmid, const double* crss,
                            const double* stress, double dt, double* accumulated_shear, double* crss_out);

/* Deterministic microstructure (no reference counterpart; the reference's
 * SetRandomMaterialGrid, UnitTester.cpp:174-231, uses rand()): periodic
 * Voronoi tessellation of num_grains seeded uniform centres in [0,N)^3
 * (minimum-image distance from voxel centres, ties to the lowest grain id),
 * each grain a Haar-uniform orientation (phi1, phi2 ~ U[0,2pi),
 * cos Phi ~ U[-1,1]) turned into the active ZXZ matrix of
 * SMatrix3x3::BuildActiveEulerMatrix (LinearAlgebra.cpp:161-181) in FP64.
 * Random stream: SplitMix64(seed).  orientation_out: N x 9 (may be NULL);
 * grain_out: N (may be NULL); euler_out: num_grains x 3 (may be NULL). */
int vpfft_generate_voronoi(int nx, int ny, int nz, int num_grains, uint64_t seed, double* orientation_out,
                           int32_t* grain_out, double* euler_out);

/* Device bytes vpfft_grid_create allocates for desc (no reference
 * counterpart): what an admission policy charges against an HBM budget. */
int vpfft_grid_bytes(const vpfft_grid_desc* desc, uint64_t* bytes_out);
/* Device-side copy of the state a strain step changes -- orientation, CRSS,
 * accumulated shear, stress, Lagrange multiplier, voxel lengths -- and its
 * restoration (no reference counterpart): lets a caller rerun a step, e.g. in
 * FP64 after a mixed-precision step that did not converge.  The copy is
 * allocated on first use (32 doubles per voxel). */
int vpfft_grid_save_state(void* grid);
int vpfft_grid_restore_state(void* grid);

/* FFT block shape (pencils per block) of one pass (0 = z, 1 = y, 2 = x) for
 * one precision of a grid; 0 restores the default (no reference
 * counterpart).  Changes how pencils are grouped into blocks, not the
 * arithmetic. */
int vpfft_grid_set_block_shape(void* grid, int precision, int pass, int block);

/* ---- batched strain steps (no reference counterpart: the reference runs one
 * grid per MPI job).  A batch advances several grids -- independent RVEs of
 * any sizes, materials and precisions -- through one RunSingleStrainStep each,
 * together: every pass of the fixed-point loop is one launch per group of
 * grids sharing a kernel variant, and each grid keeps its own reference
 * medium, convergence test and iteration count.  A grid's result is the same
 * bits whether it runs alone (vpfft_run_strain_step excepted: its reductions
 * are grouped differently) or in any batch.  Grids must be uploaded; a grid
 * may be in one batch call at a time; calls on one batch are serialised by
 * the caller. */
int vpfft_batch_create(void** out_batch);
int vpfft_batch_destroy(void* batch);
/* The FFT schedule autotuner.  level 1: the block shape of each (precision,
 * pass, grid shape) is timed over the candidates once, on the batch's own
 * data, and the fastest kept (the arithmetic does not depend on it, so the
 * bits do not change); level 2: block shapes and the radix order of the
 * pass's axis (the stage sequence, e.g. 96 = 8.4.3 or 4.8.3 or 3.4.8),
 * which changes the rounding of the transforms (results then agree with the
 * default schedule to rounding, not bit for bit); 0: fixed defaults.
 * Changing the level forgets earlier choices.  No reference counterpart (the
 * reference's FFTW plans are made once, MaterialGrid.cpp:341-372). */
int vpfft_batch_set_tuning(void* batch, int level);
/* The block shape (pencils per block) a batch uses for grid's pass
 * (0 = z, 1 = y, 2 = x); a negative value -n is the register form of an
 * n-point axis (one thread per pencil, the whole pencil in registers: the
 * default for the y pass of 16-, 24- and 32-point axes, and of 48- and
 * 64-point ones in the mixed form). */
int vpfft_batch_block_shape(void* batch, const void* grid, int pass, int* block_out);
/* The radix stages a batch uses for grid's pass (0 = z, 1 = y, 2 = x), in
 * the order they run: radices[0..*nstages) (at most 12). */
int vpfft_batch_radix_plan(void* batch, const void* grid, int pass, int* radices, int* nstages);
/* One strain step of count grids.  edot: count x 5; dt: count time steps
 * (dt <= 0: no orientation / hardening update for that grid, as in
 * vpfft_run_strain_step); avg_stress_out: count x 5 (may be NULL); infos:
 * count (may be NULL).  Per-grid errors (degenerate voxels, non-finite
 * results) are reported in infos[j].status; the return value is the first
 * of them. */
int vpfft_batch_run_strain_step(void* batch, void* const* grids, int count, const double* edot, int max_iter,
                                const double* dt, double tol, double* avg_stress_out, vpfft_step_info* infos);

/* Per-kernel device time (no reference counterpart; the reference has no
 * timers, README.make:40-43).  When enabled, every kernel launch of a strain
 * step is bracketed by HIP events on the grid's stream and its elapsed time
 * summed per kernel.  Enabling resets the totals and launches the loop's
 * kernels from the host instead of replaying a graph (same results). */
int vpfft_grid_set_profiling(void* grid, int enable);
/* Copies up to `capacity` totals (ms) and launch counts; returns the number of kernels. */
int vpfft_grid_kernel_times(void* grid, double* total_ms, long* launches, int capacity);
/* Name of kernel `id` ("" when out of range). */
const char* vpfft_kernel_name(int id);

/* Message of the last failing call on this thread. */
const char* vpfft_last_error(void);

/* Library version string. */
const char* vpfft_version(void);

#ifdef __cplusplus
}
#endif
#endif
