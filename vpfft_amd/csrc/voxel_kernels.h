// voxel_kernels.h -- launchers for the pointwise per-voxel kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "fft_green.h"

namespace vpfft {

// Fixed launch shape of every pointwise kernel: the voxel->thread assignment
// and therefore the order of every reduction is independent of timing.
constexpr int kPointThreads = 256;
constexpr int kNrWaves = 2;        // waves per SIMD the Newton kernel is compiled for
constexpr int kPartialStride = 24;  // 15 sums + 2 maxima per block, padded

// The Newton kernel runs one voxel per lane in one-wave blocks: no barrier,
// so a wave retires when its own lanes are done instead of waiting for the
// slowest wave of a larger block.
constexpr int kNrThreads = 64;
__host__ __device__ inline int nr_blocks(long n) { return (int)((n + kNrThreads - 1) / kNrThreads); }
// blocks of a one-voxel-per-thread launch of kPointThreads-thread blocks
__host__ __device__ inline int pt_blocks(long n) { return (int)((n + kPointThreads - 1) / kPointThreads); }
// partial-sum buffer rows needed by launch_nr (+ its folded rows)
inline long nr_partial_rows(long n) { return nr_blocks(n) + (nr_blocks(n) + 255) / 256; }
// partial rows one Newton launch writes: one per wave, a wave taking
// row_runs runs of 64 consecutive voxels (a z row of 256) or, for
// row_runs <= 1, one run
inline int nr_rows(long n, int row_runs) { return row_runs > 1 ? (int)(n / (64L * row_runs)) : nr_blocks(n); }

inline int point_blocks(long n) {
  long b = (n + kPointThreads - 1) / kPointThreads;
  if (b > 4096) b = 4096;
  if (b < 1) b = 1;
  return (int)b;
}

// Per-iteration record written by the device (all reductions done on device).
struct IterRecord {
  double sums[12];      // sum sigma(5), sum edot(5), sum|sigma-lambda|, sum|edot-E-d|
  double avg_stress[5];
  double avg_strain_rate[5];
  double err_stress, err_strain_rate;
  double max_nr_error;
  double nr_steps;      // total Newton steps taken over all voxels
  double degenerate;    // voxels whose max resolved shear stress was < 1e-10
  double max_nr_steps;  // largest Newton step count of any voxel
  double nr_wave_steps; // sum over waves of 64 x the wave's largest step count
  int done;
  int iteration;
};

struct LoopState {
  int iter;  // iterations completed in this strain step
  int done;  // set by the convergence test (MaterialGrid.cpp:1066-1068)
};

struct PrologueRecord {
  double L0[25];
  double L0inv[25];     // column-major
  double degenerate;
  double taylor_steps;
};

// The first non-finite value a kernel of the current strain step produced or
// read (flag 0: none).  Written once per step by whichever lane gets there
// first; reset at the start of every step.  The epilogue does nothing once it
// is set, so a failed step leaves the orientation and CRSS as they were.
struct HealthRecord {
  int flag;
  int kernel;   // HealthKernel
  int field;    // HealthField
  int iter;     // fixed-point iteration (1-based) for the loop kernels, 0 outside
  long voxel;   // x*Ny*Nz + y*Nz + z
  double value;
};
enum HealthKernel { HK_NONE = 0, HK_PROLOGUE = 1, HK_L0 = 2, HK_NR = 3, HK_EPILOGUE = 4 };
enum HealthField {
  HF_NONE = 0,
  HF_TAYLOR_STRESS = 1,    // the Taylor initialisation's stress
  HF_SECANT_INVERSE = 2,   // the inverse of the voxel's secant modulus (singular or non-finite)
  HF_L0 = 3,               // the reference medium or its inverse
  HF_DTILDE = 4,           // the strain-rate variation the FFT passes produced
  HF_STRESS = 5,           // the constrained Newton solve's stress
  HF_STRAIN_RATE = 6,      // the constitutive strain rate of that stress
  HF_LAMBDA = 7,           // the updated Lagrange multiplier
  HF_ORIENTATION = 8,      // the epilogue's orientation update
  HF_CRSS = 9              // the epilogue's hardening
};
const char* health_kernel_name(int k);
const char* health_field_name(int f);

// uniform_crss: every slip system of a voxel has the same CRSS (crss[0][v] is read)
hipError_t launch_prologue(int prec, bool uniform_crss, long n, const double* R, const double* crss,
                           const StepConst* C, void* field, double* sigma, double* lambda, double* partial,
                           HealthRecord* health, hipStream_t st);
hipError_t launch_L0_finalize(long n, int nblocks, const double* partial, StepConst* C, GreenConst<double>* g64,
                              GreenConst<float>* g32, PrologueRecord* rec, HealthRecord* health, hipStream_t st);
hipError_t launch_health_reset(HealthRecord* health, hipStream_t st);
// derive L0^-1 and the Green constants from C->L0 (already set)
hipError_t launch_derive_from_L0(StepConst* C, GreenConst<double>* g64, GreenConst<float>* g32, hipStream_t st);
// waves: 1 or 2 waves per SIMD (register budget 512 or 256 per lane);
// uniform_tau: every slip system of a voxel has the same CRSS (crss[0][v] is read)
// and the same reference slip rate (C.gdot0[0] is read)
// ls: the loop state (kernels return at entry once ls->done is set; may be
// null outside a loop); health: where the first non-finite value is recorded;
// row_runs: 0, each wave takes a 4 x 16 tile; 1, 64 consecutive voxels (a
// run of the Nz = 512 fused kernel); 4, a z row of 256 voxels with one
// partial row (the Nz = 256 fused kernel's; n a multiple of 256)
hipError_t launch_nr(int prec, int nexp_uniform, bool uniform_tau, long n, const void* dtil, void* phi, double* sigma,
                     double* lambda, const double* R, const double* crss, const StepConst* C, double* partial,
                     const LoopState* ls, HealthRecord* health, int row_runs, hipStream_t st);
// row_runs: as the loop's Newton launch (4 after launch_zfused at Nz = 256, 1 at Nz = 512)
hipError_t launch_iter_finalize(long n, int row_runs, const double* partial, const StepConst* C, double tol,
                                int max_iter, LoopState* ls, IterRecord* rec, hipStream_t st);
hipError_t launch_loop_reset(LoopState* ls, hipStream_t st);
hipError_t launch_epilogue(long n, double* R, double* crss, double* acc, const double* sigma, const StepConst* C,
                           HealthRecord* health, hipStream_t st);
hipError_t launch_aos_to_soa(const double* aos, double* soa, long count, int width, long soa_stride, long offset,
                             hipStream_t st);
hipError_t launch_soa_to_aos(const double* soa, double* aos, long count, int width, long soa_stride, long offset,
                             hipStream_t st);
hipError_t launch_convert(int prec_dst, void* dst, int prec_src, const void* src, long count, hipStream_t st);

// Per-voxel solver batches (kernel-level access to the Solvers:: functions).
hipError_t launch_constrained_batch(long n, const double* R, const double* crss, const StepConst* C,
                                    const double* s_in, const double* lam, const double* dt, double* s_out,
                                    double* err, int* steps, hipStream_t st);
hipError_t launch_taylor_batch(long n, const double* R, const double* crss, const StepConst* C, const double* s_in,
                               double* s_out, hipStream_t st);

// The fused inverse-z / per-voxel body / forward-z kernel of one iteration
// (Nz = 256 or 512), and the rebuild of phi = sigma - L0 d~ after a fused
// loop.  The fused kernel reads the half spectrum spec_in and writes
// spec_out (different buffers), fills the same partial rows as launch_nr and
// folds them the same way.
bool zfused_eligible(const FftGeom& g);
hipError_t launch_zfused(int prec, int nexp_uniform, bool uniform_tau, long n, const void* spec_in, void* spec_out,
                         const FftGeom& g, const void* tw, double scale, double* sigma, double* lambda,
                         const double* R, const double* crss, const StepConst* C, double* partial,
                         const LoopState* ls, HealthRecord* health, hipStream_t st);
hipError_t launch_phi_rebuild(int prec, long n, const void* dtil, void* phi, const double* sigma, const StepConst* C,
                              hipStream_t st);

// VPFFT::Solvers over n voxels given by their sample-frame Schmid tensors
// (n x 60) and CRSS (n x 12); C: material, strain rate, L0.  a5/b5/c5 are
// n x 5 inputs whose meaning depends on op (see k_solvers).
enum SolverOp {
  SOLVER_TAYLOR = 0, SOLVER_CONSTRAINED = 1, SOLVER_SACHS = 2, SOLVER_APPLY = 3, SOLVER_SHEAR_RATE = 4,
  SOLVER_HARDENING = 5
};
hipError_t launch_solvers(int op, long n, const double* schmid, const double* anti, const double* crss,
                          const StepConst* C, const double* a5, const double* b5, const double* c5, double eps,
                          double max_rss, int max_it, double dt, double* acc, double* err, double* out,
                          hipStream_t st);

// One grid of a batched pointwise launch.  dtil / phi are in the group's
// transform precision; partial holds nr_blocks(n) rows of kPartialStride
// (and of 18 for the prologue).
struct VoxJob {
  long n;
  const void* dtil;
  void* phi;
  double* sigma;
  double* lambda;
  double* R;
  double* crss;
  double* acc;
  StepConst* C;
  GreenConst<double>* g64;
  GreenConst<float>* g32;
  double* partial;
  LoopState* ls;
  IterRecord* rec;
  PrologueRecord* prol;
  HealthRecord* health;
};
// A batched launch: blocks [off[m], off[m+1]) belong to jobs[ids[m]]; the
// one-block-per-job kernels use ids[0..count).
struct VoxGroup {
  const VoxJob* jobs;
  const int* ids;
  const long* off;
  int count;
};
hipError_t launch_prologue_group(int prec, const VoxGroup& grp, long blocks, hipStream_t st);
hipError_t launch_L0_finalize_group(const VoxGroup& grp, hipStream_t st);
hipError_t launch_loop_reset_group(const VoxGroup& grp, hipStream_t st);
hipError_t launch_nr_group(int prec, int nexp_uniform, bool uniform_tau, const VoxGroup& grp, long blocks,
                           hipStream_t st);
hipError_t launch_iter_finalize_group(const VoxGroup& grp, double tol, int max_iter, hipStream_t st);
hipError_t launch_epilogue_group(const VoxGroup& grp, long blocks, hipStream_t st);

}  // namespace vpfft
