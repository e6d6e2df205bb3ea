// voxel_kernels.hip -- pointwise per-voxel kernels of one VPFFT strain step.
//
//   k_prologue   : UpdateSchmidtTensors + InitializeComputation +
//                  the per-voxel half of BuildLinearRefMediumStiffness
//                  (MaterialGrid.cpp:107-124, 153-210, 219-250)
//   k_L0_finalize: L0 = mean of secant-moduli inverses, NormalizeValues
//                  (MaterialGrid.cpp:252-281), plus the per-step constants
//                  derived from L0 (its inverse and the Green coefficients)
//   k_nr         : the per-voxel body of the fixed-point loop
//                  (MaterialGrid.cpp:946-1003) with its sums
//   k_iter_final : the convergence block (MaterialGrid.cpp:1009-1068)
//   k_epilogue   : orientation update + hardening (MaterialGrid.cpp:769-876,
//                  1098-1157)
//
// Layout in HBM: structure of arrays, component-major: R[9][N], crss[12][N],
// sigma[5][N], lambda[5][N], field[5][N] (polarization in, strain-rate
// variation out).  One voxel per lane, grid-stride with a fixed grid so
// every reduction has a fixed order.
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "fft_green.h"
#include "fft_regs.h"
#include "voxel_kernels.h"

namespace vpfft {

template <typename T> __device__ __forceinline__ double ld(const T* p) { return (double)*p; }

const char* health_kernel_name(int k) {
  switch (k) {
    case HK_PROLOGUE: return "prologue";
    case HK_L0: return "reference medium";
    case HK_NR: return "Newton kernel";
    case HK_EPILOGUE: return "epilogue";
    default: return "none";
  }
}
const char* health_field_name(int f) {
  switch (f) {
    case HF_TAYLOR_STRESS: return "Taylor stress";
    case HF_SECANT_INVERSE: return "inverse secant modulus";
    case HF_L0: return "L0 or its inverse";
    case HF_DTILDE: return "strain-rate variation (FFT output)";
    case HF_STRESS: return "stress";
    case HF_STRAIN_RATE: return "constitutive strain rate";
    case HF_LAMBDA: return "Lagrange multiplier";
    case HF_ORIENTATION: return "orientation";
    case HF_CRSS: return "CRSS";
    default: return "none";
  }
}

// Record the first non-finite value of the step (cold path: one lane of the
// grid wins the flag and fills the record).
__device__ __forceinline__ void report_health(HealthRecord* h, int kernel, int field, int iter, long voxel,
                                              double value) {
  if (h && atomicCAS(&h->flag, 0, 1) == 0) {
    h->kernel = kernel;
    h->field = field;
    h->iter = iter;
    h->voxel = voxel;
    h->value = value;
  }
}

// the sum of five values is finite iff all are (barring overflow of the sum,
// which is reported as non-finite too)
__device__ __forceinline__ double sum5(const double a[5]) { return a[0] + a[1] + a[2] + a[3] + a[4]; }

// Reduce W values over the block in a fixed order; thread 0 gets the result.
// W sums then M maxima go to out[0..W+M).
template <int W, int M = 1>
__device__ __forceinline__ void block_sum(double (&v)[W], double (&vmax)[M], double* out) {
  __shared__ double sh[kPointThreads / 64][W + M];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int w = 0; w < W; w++) v[w] += __shfl_xor(v[w], off);
#pragma unroll
    for (int w = 0; w < M; w++) vmax[w] = fmax(vmax[w], __shfl_xor(vmax[w], off));
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int w = 0; w < W; w++) sh[wave][w] = v[w];
#pragma unroll
    for (int w = 0; w < M; w++) sh[wave][W + w] = vmax[w];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 0; w < W; w++) {
      double acc = 0.0;
      for (int q = 0; q < kPointThreads / 64; q++) acc += sh[q][w];
      out[w] = acc;
    }
    for (int w = 0; w < M; w++) {
      double mx = 0.0;
      for (int q = 0; q < kPointThreads / 64; q++) mx = fmax(mx, sh[q][W + w]);
      out[W + w] = mx;
    }
  }
}

// A double moved across lanes by one DPP pattern (two 32-bit moves); lanes
// of rows outside ROWMASK read 0.
template <int CTRL, int ROWMASK = 0xf>
__device__ __forceinline__ double dpp64(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROWMASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWMASK, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// One step of a pairwise exchange inside rows of 16 lanes: the lane and its
// partner (the DPP pattern CTRL pairs them, and HIGH tells them apart) each
// keep one half of the H pairs (lo[j], hi[j]) and add the partner's copy of
// that half.  Both lanes of a pair add the same two numbers.
template <int CTRL, int H>
__device__ __forceinline__ void pair_halve(const double (&lo)[H], const double (&hi)[H], bool high,
                                           double (&out)[H]) {
#pragma unroll
  for (int j = 0; j < H; j++) {
    const double send = high ? lo[j] : hi[j];
    const double keep = high ? hi[j] : lo[j];
    out[j] = keep + dpp64<CTRL>(send);
  }
}

// x of this lane's row plus x of the paired row (rows 0-1, 2-3), then of the
// other half of the wave, by lane swaps; every lane gets the total over its
// lane index mod 16.  ADD selects + (else max).
template <bool ADD>
__device__ __forceinline__ double cross_rows(double x) {
  auto op = [](double a, double b) { return ADD ? a + b : fmax(a, b); };
  long long b = __double_as_longlong(x);
  auto lo = __builtin_amdgcn_permlane16_swap((unsigned)b, (unsigned)b, false, false);
  auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(b >> 32), (unsigned)(b >> 32), false, false);
  x = op(__longlong_as_double(((long long)hi[0] << 32) | lo[0]),
         __longlong_as_double(((long long)hi[1] << 32) | lo[1]));
  b = __double_as_longlong(x);
  lo = __builtin_amdgcn_permlane32_swap((unsigned)b, (unsigned)b, false, false);
  hi = __builtin_amdgcn_permlane32_swap((unsigned)(b >> 32), (unsigned)(b >> 32), false, false);
  return op(__longlong_as_double(((long long)hi[0] << 32) | lo[0]),
            __longlong_as_double(((long long)hi[1] << 32) | lo[1]));
}

// Maximum over one wave; every lane gets it.
__device__ __forceinline__ double wave_max_all(double x) {
  x = fmax(x, dpp64<0xB1>(x));   // quad_perm [1,0,3,2]
  x = fmax(x, dpp64<0x1B>(x));   // quad_perm [3,2,1,0]
  x = fmax(x, dpp64<0x141>(x));  // row_half_mirror
  x = fmax(x, dpp64<0x140>(x));  // row_mirror
  return cross_rows<false>(x);
}

// The per-wave row of the Newton kernel's partial sums: 14 sums of acc (slot
// 14 of acc is not read), then the Newton steps the wave ran for its lanes
// (wsteps; if negative, 64 x the wave's largest step count: one run of 64
// voxels), the largest NR error and the largest step count; lanes 0..16
// write out[0..16].
// The sums are a reduce-scatter in a fixed order: inside each row of 16 lanes
// the pairs (row mirror, half-row mirror, quad reversal, quad swap) each
// exchange half of what they hold, so a lane ends with one value summed over
// its row, and lane swaps add the four rows.  ~120 instructions for the 14
// sums against ~300 for one full reduction per value.
__device__ __forceinline__ void wave_partial_row(const double (&acc)[15], double max_err, double max_steps,
                                                 double* out, double wsteps = -1.0) {
  const int lane = threadIdx.x & 63;
  double lo8[8], hi8[8], s8[8];
#pragma unroll
  for (int j = 0; j < 8; j++) {
    lo8[j] = acc[j];
    hi8[j] = j + 8 < 14 ? acc[j + 8] : 0.0;
  }
  pair_halve<0x140>(lo8, hi8, lane & 8, s8);  // row_mirror: lanes l, 15 - l
  double lo4[4], hi4[4], s4[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    lo4[j] = s8[j];
    hi4[j] = s8[j + 4];
  }
  pair_halve<0x141>(lo4, hi4, lane & 4, s4);  // row_half_mirror: l, l ^ 7 within 8
  double lo2[2] = {s4[0], s4[1]}, hi2[2] = {s4[2], s4[3]}, s2[2];
  pair_halve<0x1B>(lo2, hi2, lane & 2, s2);  // quad_perm [3,2,1,0]: l, l ^ 3
  double lo1[1] = {s2[0]}, hi1[1] = {s2[1]}, s1[1];
  pair_halve<0xB1>(lo1, hi1, lane & 1, s1);  // quad_perm [1,0,3,2]: l, l ^ 1
  const double sum = cross_rows<true>(s1[0]);  // value lane % 16, over the wave
  const double me = wave_max_all(max_err), ms = wave_max_all(max_steps);
  const double val = lane < 14 ? sum : lane == 14 ? (wsteps < 0.0 ? 64.0 * ms : wsteps) : lane == 15 ? me : ms;
  if (lane < 17) out[lane] = val;
}

__device__ __forceinline__ void load_R(const double* __restrict__ R, long n, long v, double r[9]) {
#pragma unroll
  for (int e = 0; e < 9; e++) r[e] = R[e * n + v];
}
__device__ __forceinline__ void load_inv_tau(const double* __restrict__ crss, long n, long v, double it[kSlips]) {
#pragma unroll
  for (int k = 0; k < kSlips; k++) it[k] = 1.0 / crss[k * n + v];
}

// Voxel of thread tid in block blk of a one-voxel-per-thread launch.  When Ny % 4 == 0 and Nz % 16 == 0
// each wave takes a 4 (y) x 16 (z) tile of one x-plane instead of 64
// consecutive z: 16-lane runs still read whole 128-byte lines, and a compact
// tile spans fewer grains (256^3 with 4096 Voronoi grains: 2.9 on average
// against 6.6 for a 64-voxel line), so its lanes' Newton step counts (Taylor in
// the prologue, constrained in the loop) differ less and the wave idles less
// on its slowest lane.  Otherwise a block takes
// THREADS consecutive voxels.  Threads past the grid get v >= N.
template <int THREADS>
__device__ __forceinline__ long tile_voxel(long blk, int tid, int ny, int nz) {
  if ((ny & 3) || (nz & 15)) return blk * THREADS + tid;
  const long tile = blk * (THREADS / 64) + __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  const long ztn = nz >> 4, ytn = ny >> 2;
  const long zt = tile % ztn, rest = tile / ztn;
  const long yt = rest % ytn, x = rest / ytn;
  return (x * ny + yt * 4 + (lane >> 4)) * nz + zt * 16 + (lane & 15);
}

// The prologue's mapping: a wave takes a 4 x 4 x 4 cube when every edge is a
// multiple of 4 (1.96 grains per wave on average at 256^3 / 4096 grains),
// else the tile_voxel mapping.  Its loads are 32-byte runs, which the
// prologue, bound by its Taylor solves and not by HBM, can afford.
__device__ __forceinline__ long cube_voxel(long blk, int tid, int nx, int ny, int nz) {
  if ((nx & 3) || (ny & 3) || (nz & 3)) return tile_voxel<kPointThreads>(blk, tid, ny, nz);
  const long tile = blk * (kPointThreads / 64) + __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  const long ztn = nz >> 2, ytn = ny >> 2;
  const long zt = tile % ztn, rest = tile / ztn;
  const long yt = rest % ytn, xt = rest / ytn;
  return ((xt * 4 + (lane >> 4)) * ny + yt * 4 + ((lane >> 2) & 3)) * nz + zt * 4 + (lane & 3);
}

// ---------------------------------------------------------------------------
// InitializeComputation + the secant moduli of BuildLinearRefMediumStiffness
// for voxel v; adds M_loc^-1 (15), the degenerate count (slot 15) and the
// Taylor Newton steps (slot 16) to acc.
// InitializeComputation for voxel v (MaterialGrid.cpp:153-210): sigma,
// lambda and the polarization get the Taylor stress; the degenerate count and
// the Taylor Newton steps go to acc[0], acc[1].
// The prologue holds slips 0..kPrologueRegSlips-1 of the voxel's Schmid
// tensors in registers and the other seven in LDS (17.5 KB per wave, two
// 256-thread blocks per CU), so the Taylor solve runs at two waves per SIMD.
constexpr int kPrologueRegSlips = 5;
constexpr int kPrologueLdsDoubles = (kSlips - kPrologueRegSlips) * 5 * 64;  // per wave

template <typename T, bool UNI>
__device__ __forceinline__ void prologue_voxel(long v, long n, const double* __restrict__ R,
                                               const double* __restrict__ crss, const StepConst& C,
                                               T* __restrict__ field, double* __restrict__ sigma,
                                               double* __restrict__ lambda, double (&acc)[2],
                                               HealthRecord* __restrict__ health, double* mlds) {
  double tau[kSlips];
  if (UNI) {  // every slip system of the voxel has the same CRSS
    const double t = crss[v];
#pragma unroll
    for (int k = 0; k < kSlips; k++) tau[k] = t;
  } else {
#pragma unroll
    for (int k = 0; k < kSlips; k++) tau[k] = crss[k * n + v];
  }
  // bit-faithful: Schmid tensors, Sachs guess (Solvers.cpp:230-243), max RSS
  // rescale (MaterialGrid.cpp:175-190), Taylor NR (Solvers.cpp:19-101)
  const int lane = threadIdx.x & 63;
  faithful::SchmidSplit<kPrologueRegSlips> m;
  {
    double r[9], mm[kSlips][5];
    load_R(R, n, v, r);
    faithful::schmid(r, mm);
#pragma unroll
    for (int k = 0; k < kSlips; k++)
#pragma unroll
      for (int i = 0; i < 5; i++) {
        if (k < kPrologueRegSlips) m.reg[k < kPrologueRegSlips ? k : 0][i] = mm[k][i];
        else mlds[((k - kPrologueRegSlips) * 5 + i) * 64 + lane] = mm[k][i];
      }
  }
  m.lds = mlds;
  m.lane = lane;
  double guess[5];
  faithful::sachs_m(m, tau, C.edot, guess);
  const double mrss = faithful::max_rss_m(guess, m, tau);
  if (mrss > 1.1) {
#pragma unroll
    for (int i = 0; i < 5; i++) guess[i] /= mrss;
  }
  if (!(mrss >= 1e-10)) acc[0] += 1.0;  // degenerate (the reference exits here)
  double s[5];
  int tsteps = 0;
  faithful::taylor_nr_m(guess, m, tau, C, 5e-4, 1000, s, &tsteps);
  acc[1] += (double)tsteps;
#pragma unroll
  for (int i = 0; i < 5; i++) {
    sigma[i * n + v] = s[i];
    lambda[i * n + v] = s[i];
    field[i * n + v] = (T)s[i];
  }
  if (!__builtin_isfinite(sum5(s))) report_health(health, HK_PROLOGUE, HF_TAYLOR_STRESS, 0, v, sum5(s));
}

// The per-voxel half of BuildLinearRefMediumStiffness (MaterialGrid.cpp:
// 219-250) for voxel v at its Taylor stress: the secant modulus and its
// inverse as the oracle computes them (the inverses of nearly singular moduli
// carry most of L0's weight, so nothing is cut off or regularised); the
// inverse's lower triangle is added to acc.  A separate kernel from the
// Taylor solve, which needs most of the register file.
__device__ __forceinline__ void secant_voxel(long v, long n, const double* __restrict__ R,
                                             const double* __restrict__ crss, const double* __restrict__ sigma,
                                             const StepConst& C, double (&acc)[15],
                                             HealthRecord* __restrict__ health) {
  double r[9], tau[kSlips], m[kSlips][5], s[5];
  load_R(R, n, v, r);
#pragma unroll
  for (int k = 0; k < kSlips; k++) tau[k] = crss[k * n + v];
#pragma unroll
  for (int i = 0; i < 5; i++) s[i] = sigma[i * n + v];
  faithful::schmid(r, m);
  double M[25], Mi[25];
  faithful::secant_modulus(m, tau, s, C, M);
  const bool ok = faithful::inverse_gj5(M, Mi);
  double chk = 0.0;
#pragma unroll
  for (int e = 0; e < 25; e++) chk += Mi[e];
  if (!ok || !__builtin_isfinite(chk)) {
    report_health(health, HK_PROLOGUE, HF_SECANT_INVERSE, 0, v, ok ? chk : 0.0);
  } else {
#pragma unroll
    for (int i = 0; i < 5; i++)
#pragma unroll
      for (int j = 0; j <= i; j++) acc[sidx(i, j)] += Mi[j * 5 + i];
  }
}

// Prologue partial rows are 18 wide: the Taylor kernel's block b fills slots
// 15 (degenerate voxels) and 16 (Taylor Newton steps) of row b, the secant
// kernel's block b slots 0..14 (inverse secant moduli, packed lower) of row
// nb + b, where nb is the number of blocks of either launch.
template <typename T, bool UNI>
__global__ void __launch_bounds__(kPointThreads) __attribute__((amdgpu_waves_per_eu(2, 2)))
    k_prologue(long n, const double* __restrict__ R, const double* __restrict__ crss,
               const StepConst* __restrict__ Cp, T* __restrict__ field, double* __restrict__ sigma,
               double* __restrict__ lambda, double* __restrict__ partial, HealthRecord* __restrict__ health) {
  __shared__ double mlds[kPointThreads / 64][kPrologueLdsDoubles];
  const StepConst& C = *Cp;
  double acc[2] = {0.0, 0.0};
  double mx[1] = {0.0};
  // linear index i walks the blocks of cube_voxel (a permutation of the voxels)
  for (long i = (long)blockIdx.x * kPointThreads + threadIdx.x; i < (long)pt_blocks(n) * kPointThreads;
       i += (long)gridDim.x * kPointThreads) {
    const long v = cube_voxel(i / kPointThreads, threadIdx.x, C.nx, C.ny, C.nz);
    if (v < n) prologue_voxel<T, UNI>(v, n, R, crss, C, field, sigma, lambda, acc, health, mlds[threadIdx.x >> 6]);
  }
  block_sum<2>(acc, mx, partial + (long)blockIdx.x * 18 + 15);
}

__global__ void __launch_bounds__(kPointThreads) k_secant(long n, const double* __restrict__ R,
                                                          const double* __restrict__ crss,
                                                          const double* __restrict__ sigma,
                                                          const StepConst* __restrict__ Cp,
                                                          double* __restrict__ partial,
                                                          HealthRecord* __restrict__ health) {
  const StepConst& C = *Cp;
  double acc[15];
#pragma unroll
  for (int w = 0; w < 15; w++) acc[w] = 0.0;
  double mx[1] = {0.0};
  for (long i = (long)blockIdx.x * kPointThreads + threadIdx.x; i < (long)pt_blocks(n) * kPointThreads;
       i += (long)gridDim.x * kPointThreads) {
    const long v = cube_voxel(i / kPointThreads, threadIdx.x, C.nx, C.ny, C.nz);
    if (v < n) secant_voxel(v, n, R, crss, sigma, C, acc, health);
  }
  block_sum<15>(acc, mx, partial + ((long)gridDim.x + blockIdx.x) * 18);
}

// Constants derived from L0 (already in C): L0^-1 and the Green-operator
// coefficients.  Returns false when L0 is singular or L0^-1 not finite.
__device__ bool derive_constants(StepConst* C, GreenConst<double>* g64, GreenConst<float>* g32) {
  // L0^-1 by the oracle's Gauss-Jordan (the special-plane operator is
  // (-L0)^-1, MaterialGrid.cpp:628-632), symmetrised
  double Li[25];
  bool ok = faithful::inverse_gj5(C->L0, Li);
  for (int i = 0; i < 5; i++)
    for (int j = 0; j <= i; j++) {
      const double v = Li[j * 5 + i];
      ok = ok && __builtin_isfinite(v);
      C->L0inv[j * 5 + i] = v;
      C->L0inv[i * 5 + j] = v;
    }
  // its Cholesky factor, for the Newton kernel's P^T L0^-1 P (a medium that is
  // not positive definite fails the step on the host)
  for (int j = 0; j < 5; j++) {
    double d = C->L0inv[j * 5 + j];
    for (int k = 0; k < j; k++) d -= C->L0inv_w[sidx(j, k)] * C->L0inv_w[sidx(j, k)];
    ok = ok && d > 0.0;
    const double wjj = d > 0.0 ? sqrt(d) : 0.0;
    C->L0inv_w[sidx(j, j)] = wjj;
    for (int i = j + 1; i < 5; i++) {
      double t = C->L0inv[j * 5 + i];
      for (int k = 0; k < j; k++) t -= C->L0inv_w[sidx(i, k)] * C->L0inv_w[sidx(j, k)];
      C->L0inv_w[sidx(i, j)] = wjj > 0.0 ? t / wjj : 0.0;
    }
  }
  // L_ijkl = sum_ab L0(a,b) B_a,ij B_b,kl ; K_ik = sum_jl L_ijkl k_j k_l
  const double s2 = 0.70710678118654752440, s6 = 0.40824829046386301637;
  double Bm[5][3][3];
  for (int a = 0; a < 5; a++)
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) Bm[a][i][j] = 0.0;
  Bm[0][0][0] = -s2; Bm[0][1][1] = s2;
  Bm[1][0][0] = -s6; Bm[1][1][1] = -s6; Bm[1][2][2] = 2.0 * s6;
  Bm[2][1][2] = s2; Bm[2][2][1] = s2;
  Bm[3][0][2] = s2; Bm[3][2][0] = s2;
  Bm[4][0][1] = s2; Bm[4][1][0] = s2;
  const int ri[6] = {0, 1, 2, 1, 0, 0}, rk[6] = {0, 1, 2, 2, 2, 1};
  const int qj[6] = {0, 1, 2, 1, 0, 0}, ql[6] = {0, 1, 2, 2, 2, 1};
  for (int r = 0; r < 6; r++)
    for (int q = 0; q < 6; q++) {
      const int i = ri[r], k = rk[r], j = qj[q], l = ql[q];
      double c = 0.0;
      for (int a = 0; a < 5; a++)
        for (int b = 0; b < 5; b++) {
          const double L = C->L0[b * 5 + a];
          c += L * Bm[a][i][j] * Bm[b][k][l];
          if (j != l) c += L * Bm[a][i][l] * Bm[b][k][j];
        }
      g64->Kc[r][q] = c;
      g32->Kc[r][q] = (float)c;
    }
  for (int e = 0; e < 25; e++) {
    g64->L0inv[e] = C->L0inv[e];
    g32->L0inv[e] = (float)C->L0inv[e];
  }
  for (int a = 0; a < 3; a++) {
    g64->kscale[a] = C->kscale[a];
    g32->kscale[a] = (float)C->kscale[a];
  }
  g64->nx = g32->nx = C->nx;
  g64->ny = g32->ny = C->ny;
  g64->nz = g32->nz = C->nz;
  return ok;
}

// L0 from nblocks prologue partial rows (18 wide); one block.
__device__ __forceinline__ void l0_finalize_body(long n, int nblocks, const double* __restrict__ partial,
                                                 StepConst* C, GreenConst<double>* g64, GreenConst<float>* g32,
                                                 PrologueRecord* rec, HealthRecord* health) {
  __shared__ double sh[256][18];
  double acc[18];
  for (int w = 0; w < 18; w++) acc[w] = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
    for (int w = 0; w < 15; w++) acc[w] += partial[((long)nblocks + b) * 18 + w];
    for (int w = 15; w < 17; w++) acc[w] += partial[(long)b * 18 + w];
  }
  for (int w = 0; w < 18; w++) sh[threadIdx.x][w] = acc[w];
  __syncthreads();
  for (int stride = 128; stride > 0; stride >>= 1) {
    if (threadIdx.x < stride)
      for (int w = 0; w < 18; w++) sh[threadIdx.x][w] += sh[threadIdx.x + stride][w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double inv_n = 1.0 / (double)n;
    for (int i = 0; i < 5; i++)
      for (int j = 0; j <= i; j++) {
        double v = sh[0][sidx(i, j)] * inv_n;
        if (fabs(v) < 1e-10) v = 0.0;  // NormalizeValues, LinearAlgebra.h:605-613
        C->L0[j * 5 + i] = v;
        C->L0[i * 5 + j] = v;
      }
    bool ok = derive_constants(C, g64, g32);
    for (int e = 0; e < 25; e++) {
      ok = ok && __builtin_isfinite(C->L0[e]);
      rec->L0[e] = C->L0[e];
      rec->L0inv[e] = C->L0inv[e];
    }
    if (!ok) report_health(health, HK_L0, HF_L0, 0, -1, C->L0[0]);
    rec->degenerate = sh[0][15];
    rec->taylor_steps = sh[0][16];
  }
}


__global__ void __launch_bounds__(256) k_L0_finalize(long n, int nblocks, const double* __restrict__ partial,
                                                     StepConst* C, GreenConst<double>* g64, GreenConst<float>* g32,
                                                     PrologueRecord* rec, HealthRecord* health) {
  l0_finalize_body(n, nblocks, partial, C, g64, g32, rec, health);
}

__global__ void k_health_reset(HealthRecord* h) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    h->flag = 0;
    h->kernel = HK_NONE;
    h->field = HF_NONE;
    h->iter = 0;
    h->voxel = -1;
    h->value = 0.0;
  }
}

hipError_t launch_health_reset(HealthRecord* h, hipStream_t st) {
  hipLaunchKernelGGL(k_health_reset, dim3(1), dim3(64), 0, st, h);
  return hipGetLastError();
}

__global__ void k_loop_reset(LoopState* ls) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    ls->iter = 0;
    ls->done = 0;
  }
}

hipError_t launch_loop_reset(LoopState* ls, hipStream_t st) {
  hipLaunchKernelGGL(k_loop_reset, dim3(1), dim3(64), 0, st, ls);
  return hipGetLastError();
}

__global__ void k_derive(StepConst* C, GreenConst<double>* g64, GreenConst<float>* g32) {
  if (threadIdx.x == 0 && blockIdx.x == 0) (void)derive_constants(C, g64, g32);
}

// values alike in every lane of the block, moved to scalar registers
__device__ __forceinline__ int vuni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ long vuni(long x) {
  const unsigned long u = (unsigned long)x;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
  return (long)(((unsigned long)hi << 32) | lo);
}
template <typename P> __device__ __forceinline__ P* vuni(P* p) { return (P*)vuni((long)p); }

__device__ __forceinline__ int vox_member(const VoxGroup& grp, long b, long* local) {
  int lo = 0, hi = grp.count;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (vuni(grp.off[mid]) <= b) lo = mid;
    else hi = mid;
  }
  *local = b - vuni(grp.off[lo]);
  return vuni(grp.ids[lo]);
}

// A uniform pointer from a job table, made directly in the global address
// space from its integer value.  A generic pointer cast to global and back is
// folded back to generic, and its loads become flat loads (and read-only
// constant structs land in vector registers instead of scalar ones).
template <typename P> __device__ __forceinline__ P* vuni_global(P* p) {
  const unsigned long u = (unsigned long)vuni((long)p);
  return (P*)(__attribute__((address_space(1))) P*)u;
}

// the job's fields in scalar registers
__device__ __forceinline__ VoxJob vox_job(const VoxJob* J) {
  VoxJob r;
  r.n = vuni(J->n);
  r.dtil = vuni_global(J->dtil);
  r.phi = vuni_global(J->phi);
  r.sigma = vuni_global(J->sigma);
  r.lambda = vuni_global(J->lambda);
  r.R = vuni_global(J->R);
  r.crss = vuni_global(J->crss);
  r.acc = vuni_global(J->acc);
  r.C = vuni_global(J->C);
  r.g64 = vuni_global(J->g64);
  r.g32 = vuni_global(J->g32);
  r.partial = vuni_global(J->partial);
  r.ls = vuni_global(J->ls);
  r.rec = vuni_global(J->rec);
  r.prol = vuni_global(J->prol);
  r.health = vuni_global(J->health);
  return r;
}

// ---------------------------------------------------------------------------
// K4: per-voxel body of the fixed-point loop (MaterialGrid.cpp:946-1003),
// with the Newton solve in the crystal frame of the voxel.
// partial per block: [0..4] sum sigma, [5..9] sum edot, [10] sum|sigma-lambda|,
// [11] sum|edot-E-d|, [12] NR steps, [13] degenerate, [14] sum of 64 x
// per-wave max NR steps; then max NR error, max NR steps
// NE > 0: every slip system has rate-sensitivity exponent NE.
// ---------------------------------------------------------------------------
// MODE: NR_SINGLE takes the grid's pointers from the kernel arguments and
// d~ / phi from HBM; NR_BATCHED the pointers from the job table of grp
// (blocks of many grids in one launch); NR_ZFUSED (Nz = 256) is one
// fixed-point iteration's inverse z transform, this body and the forward z
// transform in one kernel: a block (one wave) takes one z row of all five
// components; its 3 pencils (component pairs) of the half spectrum
// zf.spec_in go through the inverse transforms into LDS, the wave runs the
// body on the row's 4 runs of 64 voxels, and the phi row goes through the
// forward transforms to zf.spec_out (another buffer: the caller alternates
// them, so d~ of the last iteration can be rebuilt after the loop).  Neither
// real field crosses HBM, and each wave moves between its transforms and
// its Newton solves on its own, so the memory-bound and the FP64-bound work
// of different waves overlap on a CU.  The operations are those of
// k_z_c2r256, the single form NR_ROWS256 (a block takes the 4 runs of a z
// row) and k_z_r2c256, so the bits are the same.  The lanes add up
// the 12 floating-point sums of their voxels over the row's runs in LDS (in
// the FP64 form, in the transforms' exchange buffer, idle during the runs),
// and the wave writes one partial row per z row: one reduction per wave
// instead of one per run.
// NR_ZFUSED512 (Nz = 512) is the same with rows of 512 voxels (8 runs):
// a pencil takes 32 lanes (the 16 x 16 x 2 form of k_z_c2r512 / k_z_r2c512;
// the single form with zf.linear = 1; one partial row per run of 64 voxels),
// so the wave transforms the component pairs (0, 1) and (2, 3) side by side
// and component 4 alone, and the transforms' LDS lies inside the row's
// 5 x 512 values (20 KB in FP64: eight waves fill a CU's 160 KB).
// NR_ROWS256 is the single form of the Nz = 256 fused kernel's Newton part:
// a block takes a z row of 256 voxels (d~ and phi in HBM), with the same
// LDS row sums and one partial row.
enum NrMode { NR_SINGLE = 0, NR_BATCHED = 1, NR_ZFUSED = 2, NR_ZFUSED512 = 3, NR_ROWS256 = 4 };

struct ZFusedArgs {
  const void* spec_in;
  void* spec_out;
  const void* tw;  // W256^j in the transform precision
  long nc;         // spectrum stride per component
  int nzp;         // spectrum row stride
  double scale;    // 1/N of the inverse transform
  int linear;      // NR_SINGLE: 0, a block takes a 4 x 16 tile; 1, 64 consecutive voxels (a run
                   // of the Nz = 512 fused form)
};

template <typename T, int NE, bool UNIFORM_TAU, int MODE>
__global__ void __launch_bounds__(kNrThreads) __attribute__((amdgpu_waves_per_eu(kNrWaves, kNrWaves)))
    k_nr(long n_, const T* __restrict__ dtil_, T* __restrict__ phi_, double* __restrict__ sigma_,
         double* __restrict__ lambda_, const double* __restrict__ R_, const double* __restrict__ crss_,
         const StepConst* __restrict__ Cp_, double* __restrict__ partial_, const LoopState* __restrict__ ls_,
         HealthRecord* __restrict__ health_, VoxGroup grp, ZFusedArgs zf) {
  constexpr bool FUSED = MODE == NR_ZFUSED || MODE == NR_ZFUSED512;
  constexpr int ZN = MODE == NR_ZFUSED512 ? 512 : 256;  // a fused block's row
  constexpr int RUNS = ZN / 64;
  using Cx = typename C2<T>::t;
  long n, blk;
  const T* dtil;
  T* phi;
  double *sigma, *lambda, *partial;
  const double *R, *crss;
  const StepConst* Cp;
  const LoopState* ls;
  HealthRecord* health;
  if constexpr (MODE == NR_BATCHED) {
    const VoxJob J = vox_job(grp.jobs + vox_member(grp, blockIdx.x, &blk));
    if (J.ls->done) return;
    ls = J.ls;
    health = J.health;
    n = J.n;
    dtil = (const T*)J.dtil;
    phi = (T*)J.phi;
    sigma = J.sigma;
    lambda = J.lambda;
    R = (const double*)J.R;
    crss = (const double*)J.crss;
    Cp = J.C;
    partial = J.partial;
  } else {
    if (ls_ && ls_->done) return;
    ls = ls_;
    health = health_;
    n = n_;
    dtil = dtil_;
    phi = phi_;
    sigma = sigma_;
    lambda = lambda_;
    R = R_;
    crss = crss_;
    Cp = Cp_;
    partial = partial_;
    blk = blockIdx.x;
  }
  // the fused form's LDS: the exchange buffer of the z transforms and the
  // row's d~ (then phi), [component][z]
  __shared__ T zbuf[MODE == NR_ZFUSED ? zbuf_reals<3>() : 1];
  __shared__ T fld[FUSED ? 5 * ZN : 1];
  const int zt16 = threadIdx.x & 15, zpair = threadIdx.x >> 4;
  const bool zactive = zpair < 3;
  const long zrow = blockIdx.x;
  if constexpr (MODE == NR_ZFUSED512) {  // d~ row: inverse z transforms into LDS
    // component 4 first, its transform's LDS in the rows of components 0 and
    // 1 (not written yet); then the pairs (0, 1) and (2, 3), whose LDS
    // (the same rows) the transform has left when the values are written
    // (each transform takes its lane coordinates afresh: the compiler would
    // otherwise keep the first one's twiddle and LDS addresses live, in
    // scratch, for the second)
    const Cx* si = (const Cx*)zf.spec_in;
    const int half = threadIdx.x >> 5;
    Cx v[16];
    {
      const int t = threadIdx.x & 31, k1 = t >> 1, t0 = t & 1;
      z_c2r512_pencil<T>(si + 4L * zf.nc + zrow * zf.nzp, nullptr, half == 0, (const Cx*)zf.tw, fld + half * 512, t, v);
      if (half == 0) {
#pragma unroll
        for (int m = 0; m < 16; m++) fld[4 * 512 + k1 + 16 * m + 256 * t0] = v[m].x * (T)zf.scale;
      }
    }
    {
      int t = threadIdx.x & 31;
      __asm__ volatile("" : "+v"(t));
      const int k1 = t >> 1, t0 = t & 1;
      const Cx* sa = si + (long)(2 * half) * zf.nc + zrow * zf.nzp;
      z_c2r512_pencil<T>(sa, sa + zf.nc, true, (const Cx*)zf.tw, fld + half * 512, t, v);
#pragma unroll
      for (int m = 0; m < 16; m++) {
        const int z = k1 + 16 * m + 256 * t0;
        fld[(2 * half) * 512 + z] = v[m].x * (T)zf.scale;
        fld[(2 * half + 1) * 512 + z] = v[m].y * (T)zf.scale;
      }
    }
    __syncthreads();
  }
  if constexpr (MODE == NR_ZFUSED) {  // d~ row: inverse z transforms into LDS
    const Cx* si = (const Cx*)zf.spec_in;
    const Cx* sa = si + (long)(2 * zpair) * zf.nc + zrow * zf.nzp;
    const Cx* sb = zpair < 2 ? si + (long)(2 * zpair + 1) * zf.nc + zrow * zf.nzp : nullptr;
    Cx v[16];
    z_c2r256_pencil<T, 3>(sa, sb, zactive, (const Cx*)zf.tw, zbuf, zpair, zt16, v);
    if (zactive) {
#pragma unroll
      for (int k2 = 0; k2 < 16; k2++) {
        fld[(2 * zpair) * 256 + zt16 + 16 * k2] = v[k2].x * (T)zf.scale;
        if (zpair < 2) fld[(2 * zpair + 1) * 256 + zt16 + 16 * k2] = v[k2].y * (T)zf.scale;
      }
    }
    __syncthreads();
  }
  const StepConst& C = *as_const_mem(Cp);
  // Rows of 256 voxels (the fused kernel, and the single form with
  // zf.linear = 4) write one partial row per wave: each lane keeps the 12
  // floating-point sums of its voxels in LDS across the runs (kRowSums x 64
  // doubles; the FP64 fused kernel uses its idle exchange buffer), the step
  // and degenerate counts, the maxima and the wave's steps in registers.
  constexpr int kRowSums = 12;
  constexpr bool ZBUF_ACC = MODE == NR_ZFUSED && sizeof(T) == 8;
  static_assert(!ZBUF_ACC || zbuf_reals<3>() >= kRowSums * 64, "row sums fit the exchange buffer");
  __shared__ double accbuf[(MODE == NR_ROWS256 || (MODE == NR_ZFUSED && !ZBUF_ACC)) ? kRowSums * 64 : 1];
  double* accl = ZBUF_ACC ? (double*)(void*)zbuf : accbuf;
  constexpr bool row_sums = MODE == NR_ZFUSED || MODE == NR_ROWS256;
  constexpr int nruns = FUSED ? RUNS : (MODE == NR_ROWS256 ? 4 : 1);
  // (the counts as integers: fewer registers live through the solves)
  int rsteps = 0, rdegen = 0, rmaxsteps = 0, wsteps = 0;
  double rmaxerr = 0.0;
  for (int q = 0; q < nruns; q++) {
    double acc[15];
#pragma unroll
    for (int w = 0; w < 15; w++) acc[w] = 0.0;
    double mx[2] = {0.0, 0.0};  // max NR error, max NR steps
    double run_steps = 0.0;
    // one voxel per lane: nothing but the voxel's own state is live
    long v, slot = 0;
    if constexpr (FUSED) {
      slot = q * 64 + (threadIdx.x & 63);
      v = zrow * ZN + slot;
      blk = zrow * RUNS + q;
    } else if (MODE == NR_ROWS256) {
      v = (blk * nruns + q) * kNrThreads + threadIdx.x;
    } else if (MODE == NR_SINGLE && zf.linear) {
      v = blk * kNrThreads + threadIdx.x;
    } else {
      v = tile_voxel<kNrThreads>(blk, threadIdx.x, C.ny, C.nz);
    }
    if (v < n) {
      double it[kSlips], r9[9];
      load_R(R, n, v, r9);
      if (UNIFORM_TAU) {
        // every slip system of the voxel has the same CRSS (one load) and
        // the same reference slip rate (the host checks both)
        const double t = 1.0 / crss[v];
#pragma unroll
        for (int k = 0; k < kSlips; k++) it[k] = t;
      } else {
        load_inv_tau(crss, n, v, it);
      }
      // yield-surface rescale, MaterialGrid.cpp:954-959, of the stress in
      // the crystal frame (the solve's start point)
      double mrss, sc[5];
      {
        double Q[25], s0[5];
#pragma unroll
        for (int i = 0; i < 5; i++) s0[i] = sigma[i * n + v];
        cf::dev_map(r9, Q);
        cf::rot(Q, s0, sc);
        mrss = cf::max_rss<UNIFORM_TAU>(sc, it);
      }
      if (!(mrss >= 1e-10)) acc[13] += 1.0;
      const double imrss = rcp_nr(mrss);
#pragma unroll
      for (int i = 0; i < 5; i++) sc[i] *= imrss;
      // q = L0^-1 lambda + edot + d~ ; lambda and d~ are re-read after the solve
      double qv[5];
#pragma unroll
      for (int i = 0; i < 5; i++) qv[i] = C.edot[i] + (FUSED ? (double)fld[i * ZN + slot] : ld(dtil + i * n + v));
#pragma unroll
      for (int j = 0; j < 5; j++) {
        const double lj = lambda[j * n + v];
#pragma unroll
        for (int i = 0; i < 5; i++) qv[i] += C.L0inv[j * 5 + i] * lj;
      }
      double err = 0.0;
      int steps = 0;
      double G[15];
      {
        double P[25], qc[5], A0[15];
        cf::inverse_map(r9, P);
        cf::rot_t(P, qv, qc);
        cf::congruence_chol(P, C.L0inv_w, A0);
        cf::gram(P, G);
        cf::constrained_nr<NE, UNIFORM_TAU && (NE > 0)>(sc, G, it, C, qc, A0, 5e-4, 1e9, 100, &err, &steps);
      }
      // the solution in the sample frame, s = P sc = Q^T G sc, and its strain
      // rate from the crystal frame, e = Q^T ec
      double s[5], e[5];
      {
        // Q(R) formed again from the orientation's 9 values, which stay live
        // through the solve (Q's 25 would not fit beside the solve's state)
        double Q[25], ec[5], gs[5];
        cf::dev_map(r9, Q);
        cf::apply_constitutive<NE, UNIFORM_TAU && (NE > 0)>(sc, it, C, ec);
        cf::rot_t(Q, ec, e);
        cf::symv(G, sc, gs);
        cf::rot_t(Q, gs, s);
      }
      // re-read d~ and lambda here rather than keep them live through the solve
      __asm__ volatile("" ::: "memory");
      double d[5], lam[5];
#pragma unroll
      for (int i = 0; i < 5; i++) {
        d[i] = FUSED ? (double)fld[i * ZN + slot] : ld(dtil + i * n + v);
        lam[i] = lambda[i * n + v];
      }
      double Ld[5], t[5], Lt[5];
      matvec5(C.L0, d, Ld);
#pragma unroll
      for (int i = 0; i < 5; i++) t[i] = C.edot[i] + d[i] - e[i];
      matvec5(C.L0, t, Lt);
      double se2 = 0.0, ee2 = 0.0;
#pragma unroll
      for (int i = 0; i < 5; i++) {
        lam[i] += Lt[i];
        if constexpr (FUSED) fld[i * ZN + slot] = (T)(s[i] - Ld[i]);
        else phi[i * n + v] = (T)(s[i] - Ld[i]);
        sigma[i * n + v] = s[i];
        lambda[i * n + v] = lam[i];
        const double se = s[i] - lam[i];
        const double ee = e[i] - C.edot[i] - d[i];
        se2 += se * se;
        ee2 += ee * ee;
        acc[i] += s[i];
        acc[5 + i] += e[i];
      }
      acc[10] += sqrt(se2);
      acc[11] += sqrt(ee2);
      acc[12] += (double)steps;
      mx[0] = fmax(mx[0], err);
      mx[1] = fmax(mx[1], (double)steps);
      run_steps = (double)steps;
      // the first non-finite value of the step: the FFT's output, then what
      // this kernel made of it
      const double fd = sum5(d), fs = sum5(s), fe = sum5(e), fl = sum5(lam);
      if (__builtin_expect(!__builtin_isfinite(fd + fs + fe + fl), 0)) {
        const int f = !__builtin_isfinite(fd) ? HF_DTILDE
                      : !__builtin_isfinite(fs) ? HF_STRESS
                      : !__builtin_isfinite(fe) ? HF_STRAIN_RATE : HF_LAMBDA;
        const double val = f == HF_DTILDE ? fd : f == HF_STRESS ? fs : f == HF_STRAIN_RATE ? fe : fl;
        report_health(health, HK_NR, f, ls ? ls->iter + 1 : 0, v, val);
      }
    }
    if constexpr (row_sums) {
      const int lane = threadIdx.x & 63;
#pragma unroll
      for (int w = 0; w < kRowSums; w++) accl[w * 64 + lane] = q == 0 ? acc[w] : accl[w * 64 + lane] + acc[w];
      rsteps += (int)acc[12];
      rdegen += (int)acc[13];
      rmaxerr = fmax(rmaxerr, mx[0]);
      rmaxsteps = max(rmaxsteps, (int)mx[1]);
      // lane efficiency: a wave runs as many steps as its slowest lane, per run
      wsteps += __builtin_amdgcn_readfirstlane((int)wave_max_all(run_steps));
    } else {
      // lane efficiency of the Newton loop: a wave runs as many steps as its
      // slowest lane, the wave's maximum of the steps (slot 14 is 0 in every
      // lane before the reduction and 64 x that maximum after it)
      wave_partial_row(acc, mx[0], mx[1], partial + blk * kPartialStride);
    }
  }
  if constexpr (row_sums) {
    const int lane = threadIdx.x & 63;
    double tot[15];
#pragma unroll
    for (int w = 0; w < kRowSums; w++) tot[w] = accl[w * 64 + lane];
    tot[12] = (double)rsteps;
    tot[13] = (double)rdegen;
    tot[14] = 0.0;
    wave_partial_row(tot, rmaxerr, (double)rmaxsteps, partial + (FUSED ? zrow : (long)blockIdx.x) * kPartialStride,
                     64.0 * (double)wsteps);
  }
  if constexpr (MODE == NR_ZFUSED512) {  // phi row: forward z transforms from LDS
    // the pairs (0, 1) and (2, 3) first, their transform's LDS in the rows of
    // components 0 and 1 once every lane has read its values; then
    // component 4 from its own row
    __syncthreads();
    const int half = threadIdx.x >> 5;
    Cx* so = (Cx*)zf.spec_out;
    Cx v[16];
    {
      int t = threadIdx.x & 31;
      __asm__ volatile("" : "+v"(t));
#pragma unroll
      for (int j = 0; j < 16; j++)
        v[j] = cmake<T>(fld[(2 * half) * 512 + t + 32 * j], fld[(2 * half + 1) * 512 + t + 32 * j]);
      __syncthreads();
      Cx* sa = so + (long)(2 * half) * zf.nc + zrow * zf.nzp;
      z_r2c512_pencil<T>(v, true, (const Cx*)zf.tw, fld + half * 512, t, sa, sa + zf.nc);
    }
    {
      int t = threadIdx.x & 31;
      __asm__ volatile("" : "+v"(t));
#pragma unroll
      for (int j = 0; j < 16; j++) v[j] = cmake<T>(half == 0 ? fld[4 * 512 + t + 32 * j] : (T)0, (T)0);
      z_r2c512_pencil<T>(v, half == 0, (const Cx*)zf.tw, fld + half * 512, t, so + 4L * zf.nc + zrow * zf.nzp,
                         nullptr);
    }
  }
  if constexpr (MODE == NR_ZFUSED) {  // phi row: forward z transforms from LDS
    __syncthreads();
    // the pencil coordinates afresh: the compiler would otherwise keep the
    // inverse transforms' twiddle addresses (the same tw[t k]) live through
    // the body, in scratch
    int t3 = threadIdx.x & 15;
    __asm__ volatile("" : "+v"(t3));
    Cx v[16];
#pragma unroll
    for (int j = 0; j < 16; j++) {
      const T a = zactive ? fld[(2 * zpair) * 256 + t3 + 16 * j] : (T)0;
      const T b = (zactive && zpair < 2) ? fld[(2 * zpair + 1) * 256 + t3 + 16 * j] : (T)0;
      v[j] = cmake<T>(a, b);
    }
    Cx* so = (Cx*)zf.spec_out;
    Cx* sa = so + (long)(2 * zpair) * zf.nc + zrow * zf.nzp;
    Cx* sb = zpair < 2 ? so + (long)(2 * zpair + 1) * zf.nc + zrow * zf.nzp : nullptr;
    z_r2c256_pencil<T, 3>(v, zactive, (const Cx*)zf.tw, zbuf, zpair, t3, sa, sb);
  }
}

// phi = sigma - L0 d~ after the loop of a fused step (the fused kernel keeps
// both rows in LDS): the values its body wrote there, by the same operations.
template <typename T>
__global__ void __launch_bounds__(256) k_phi_rebuild(long n, const T* __restrict__ dtil, T* __restrict__ phi,
                                                     const double* __restrict__ sigma, const StepConst* __restrict__ Cp) {
  const StepConst& C = *as_const_mem(Cp);
  for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (long)gridDim.x * blockDim.x) {
    double d[5], Ld[5];
#pragma unroll
    for (int i = 0; i < 5; i++) d[i] = ld(dtil + i * n + v);
    matvec5(C.L0, d, Ld);
#pragma unroll
    for (int i = 0; i < 5; i++) phi[i * n + v] = (T)(sigma[i * n + v] - Ld[i]);
  }
}

// Reduce rows [256 b, 256 b + 256) of the per-block partials into row b of
// out: 15 sums then 2 maxima, fixed tree order.
__global__ void __launch_bounds__(256) k_fold_partials(const double* __restrict__ in, int rows,
                                                       double* __restrict__ out, const int* __restrict__ done) {
  if (done && *done) return;
  // the block's 256 rows are one contiguous chunk: load it coalesced
  __shared__ double sh[256][kPartialStride];
  double* flat = &sh[0][0];
  const long r0 = (long)blockIdx.x * 256;
  const long valid = (rows - r0 < 256 ? rows - r0 : 256) * kPartialStride;
  for (int e = threadIdx.x; e < 256 * kPartialStride; e += 256) flat[e] = e < valid ? in[r0 * kPartialStride + e] : 0.0;
  __syncthreads();
  for (int stride = 128; stride > 0; stride >>= 1) {
    if (threadIdx.x < stride) {
      for (int w = 0; w < 15; w++) sh[threadIdx.x][w] += sh[threadIdx.x + stride][w];
      for (int w = 15; w < 17; w++) sh[threadIdx.x][w] = fmax(sh[threadIdx.x][w], sh[threadIdx.x + stride][w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int w = 0; w < 17; w++) out[(long)blockIdx.x * kPartialStride + w] = sh[0][w];
}

// LoopState: the iteration counter and the done flag, both on the device so
// a captured graph of several iterations needs no host round trip.
// The fixed-point convergence test from nblocks rows of NR partials (one block).
__device__ __forceinline__ void iter_finalize_body(long n, int nblocks, const double* __restrict__ partial,
                                                   const StepConst* __restrict__ C, double tol, int max_iter,
                                                   LoopState* ls, IterRecord* rec) {
  if (ls->done) return;
  const int it = ls->iter;
  __shared__ double sh[256][17];
  double acc[17];
  for (int w = 0; w < 17; w++) acc[w] = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
    for (int w = 0; w < 15; w++) acc[w] += partial[(long)b * kPartialStride + w];
    for (int w = 15; w < 17; w++) acc[w] = fmax(acc[w], partial[(long)b * kPartialStride + w]);
  }
  for (int w = 0; w < 17; w++) sh[threadIdx.x][w] = acc[w];
  __syncthreads();
  for (int stride = 128; stride > 0; stride >>= 1) {
    if (threadIdx.x < stride) {
      for (int w = 0; w < 15; w++) sh[threadIdx.x][w] += sh[threadIdx.x + stride][w];
      for (int w = 15; w < 17; w++) sh[threadIdx.x][w] = fmax(sh[threadIdx.x][w], sh[threadIdx.x + stride][w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    IterRecord& R = rec[it];
    // MaterialGrid.cpp:1036-1053
    const double nd = (double)n;
    for (int w = 0; w < 12; w++) R.sums[w] = sh[0][w];
    double es = sh[0][10] / nd, ee = sh[0][11] / nd;
    double avg[5], avge[5];
    for (int i = 0; i < 5; i++) {
      avg[i] = sh[0][i] / nd;
      avge[i] = sh[0][5 + i] / nd;
      R.avg_stress[i] = avg[i];
      R.avg_strain_rate[i] = avge[i];
    }
    es /= sqrt(dot5(avg, avg));
    ee /= sqrt(dot5(C->edot, C->edot));
    R.err_stress = es;
    R.err_strain_rate = ee;
    R.nr_steps = sh[0][12];
    R.degenerate = sh[0][13];
    R.nr_wave_steps = sh[0][14];
    R.max_nr_error = sh[0][15];
    R.max_nr_steps = sh[0][16];
    R.iteration = it + 1;
    // loop condition, MaterialGrid.cpp:1066-1068
    R.done = R.degenerate > 0 || !((it + 1) < max_iter && (es > tol || ee > tol));
    ls->iter = it + 1;
    ls->done = R.done;
  }
}

__global__ void __launch_bounds__(256) k_iter_finalize(long n, int nblocks, const double* __restrict__ partial,
                                                       const StepConst* __restrict__ C, double tol, int max_iter,
                                                       LoopState* ls, IterRecord* rec) {
  iter_finalize_body(n, nblocks, partial, C, tol, max_iter, ls, rec);
}

// ---------------------------------------------------------------------------
// Epilogue: R += dt * sum_k (R a0_k R^T) gdot_k sgn(r_k) with the old Schmid
// tensors, then Voce hardening with the new ones.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void epilogue_voxel(long v, long n, double* __restrict__ R, double* __restrict__ crss,
                                               double* __restrict__ acc_shear, const double* __restrict__ sigma,
                                               const StepConst& C, HealthRecord* __restrict__ health) {
  {
    double r[9], tau[kSlips], it[kSlips], m[kSlips][5], s[5];
    load_R(R, n, v, r);
#pragma unroll
    for (int k = 0; k < kSlips; k++) { tau[k] = crss[k * n + v]; it[k] = 1.0 / tau[k]; }
#pragma unroll
    for (int i = 0; i < 5; i++) s[i] = sigma[i * n + v];
    schmid_from_R(r, m);
    double w01 = 0, w02 = 0, w12 = 0;
#pragma unroll
    for (int k = 0; k < kSlips; k++) {
      const double rs = dot5(m[k], s) * it[k];
      const double g = gamma_n(C, k, fabs(rs)) * sgn(rs);
      double W[3];
      anti_from_R(r, k, W);
      w01 += W[0] * g;
      w02 += W[1] * g;
      w12 += W[2] * g;
    }
    const double dt = C.dt;
    const double Wm[9] = {0.0, w01, w02, -w01, 0.0, w12, -w02, -w12, 0.0};
    double rsum = 0.0;
#pragma unroll
    for (int e = 0; e < 9; e++) {
      r[e] += Wm[e] * dt;
      R[e * n + v] = r[e];
      rsum += r[e];
    }
    if (!__builtin_isfinite(rsum)) report_health(health, HK_EPILOGUE, HF_ORIENTATION, 0, v, rsum);
    // hardening with the updated Schmid tensors, Solvers.cpp:314-373
    schmid_from_R(r, m);
    double g[kSlips], local = 0.0;
#pragma unroll
    for (int k = 0; k < kSlips; k++) {
      const double rs = dot5(m[k], s) * it[k];
      g[k] = gamma_n(C, k, fabs(rs));
      local += fabs(g[k]) * dt;
    }
    const double A = acc_shear[v];
#pragma unroll
    for (int i = 0; i < kSlips; i++) {
      double dtau = 0.0;
#pragma unroll
      for (int j = 0; j < kSlips; j++) dtau += C.hmat[i * kSlips + j] * fabs(g[j]) * dt;
      const double tiny = 1e-4 * C.tau0[i];
      double voce = 0.0;
      if (fabs(C.theta0[i]) > tiny) {
        voce = C.theta1[i] * local;
        if (fabs(C.tau1[i]) > tiny) {
          const double fr = fabs(C.theta0[i] / C.tau1[i]);
          const double e_ini = exp(-A * fr);
          const double e_del = exp(-local * fr);
          voce = voce - (fr * C.tau1[i] * C.theta1[i]) / fr * e_ini * (e_del - 1.0) -
                 C.theta1[i] / fr * e_ini * (e_del * ((A + local) * fr + 1.0) - (A * fr + 1.0));
        }
      }
      const double t = tau[i] + dtau * voce / local;
      crss[i * n + v] = t;
      if (!__builtin_isfinite(t)) report_health(health, HK_EPILOGUE, HF_CRSS, 0, v, t);
    }
    acc_shear[v] = A + local;
  }
}

__global__ void __launch_bounds__(kPointThreads) k_epilogue(long n, double* __restrict__ R, double* __restrict__ crss,
                                                            double* __restrict__ acc_shear,
                                                            const double* __restrict__ sigma,
                                                            const StepConst* __restrict__ Cp,
                                                            HealthRecord* __restrict__ health) {
  // a step that met a non-finite value leaves the orientation and CRSS alone
  if (health->flag) return;
  for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (long)gridDim.x * blockDim.x)
    epilogue_voxel(v, n, R, crss, acc_shear, sigma, *Cp, health);
}

// ---------------------------------------------------------------------------
// Batched pointwise kernels: one launch covers many grids (jobs).  Blocks
// [off[m], off[m+1]) belong to jobs[ids[m]], one voxel per thread, so each
// job's partial rows and their reduction order depend on that job alone.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kPointThreads) __attribute__((amdgpu_waves_per_eu(2, 2)))
    k_prologue_b(VoxGroup grp) {
  __shared__ double mlds[kPointThreads / 64][kPrologueLdsDoubles];
  long local;
  const VoxJob J = vox_job(grp.jobs + vox_member(grp, blockIdx.x, &local));
  double acc[2] = {0.0, 0.0};
  double mx[1] = {0.0};
  const StepConst& C = *as_const_mem(J.C);
  const long v = cube_voxel(local, threadIdx.x, C.nx, C.ny, C.nz);
  if (v < J.n)
    prologue_voxel<T, false>(v, J.n, J.R, J.crss, C, (T*)J.phi, J.sigma, J.lambda, acc, J.health,
                             mlds[threadIdx.x >> 6]);
  block_sum<2>(acc, mx, J.partial + local * 18 + 15);
}

__global__ void __launch_bounds__(kPointThreads) k_secant_b(VoxGroup grp) {
  long local;
  const VoxJob J = vox_job(grp.jobs + vox_member(grp, blockIdx.x, &local));
  double acc[15];
#pragma unroll
  for (int w = 0; w < 15; w++) acc[w] = 0.0;
  double mx[1] = {0.0};
  const StepConst& C = *as_const_mem(J.C);
  const long v = cube_voxel(local, threadIdx.x, C.nx, C.ny, C.nz);
  if (v < J.n) secant_voxel(v, J.n, J.R, J.crss, J.sigma, C, acc, J.health);
  block_sum<15>(acc, mx, J.partial + ((long)pt_blocks(J.n) + local) * 18);
}

__global__ void __launch_bounds__(256) k_L0_finalize_b(VoxGroup grp) {
  const VoxJob J = vox_job(grp.jobs + vuni(grp.ids[blockIdx.x]));
  l0_finalize_body(J.n, pt_blocks(J.n), J.partial, J.C, J.g64, J.g32, J.prol, J.health);
}

__global__ void k_loop_reset_b(VoxGroup grp) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < grp.count) {
    LoopState* ls = grp.jobs[grp.ids[m]].ls;
    ls->iter = 0;
    ls->done = 0;
  }
}

__global__ void __launch_bounds__(256) k_iter_finalize_b(VoxGroup grp, double tol, int max_iter) {
  const VoxJob J = vox_job(grp.jobs + vuni(grp.ids[blockIdx.x]));
  iter_finalize_body(J.n, nr_blocks(J.n), J.partial, J.C, tol, max_iter, J.ls, J.rec);
}

__global__ void __launch_bounds__(kPointThreads) k_epilogue_b(VoxGroup grp) {
  long local;
  const VoxJob J = vox_job(grp.jobs + vox_member(grp, blockIdx.x, &local));
  if (J.health->flag) return;
  const long v = local * kPointThreads + threadIdx.x;
  if (v < J.n) epilogue_voxel(v, J.n, J.R, J.crss, J.acc, J.sigma, *as_const_mem(J.C), J.health);
}

// ---------------------------------------------------------------------------
// Layout conversions (host AoS <-> device SoA), chunked through staging.
// ---------------------------------------------------------------------------
__global__ void k_aos_to_soa(const double* __restrict__ aos, double* __restrict__ soa, long count, int width,
                             long stride, long offset) {
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < count * width; t += (long)gridDim.x * blockDim.x) {
    const long v = t / width;
    const int c = (int)(t - v * width);
    soa[c * stride + offset + v] = aos[t];
  }
}
__global__ void k_soa_to_aos(const double* __restrict__ soa, double* __restrict__ aos, long count, int width,
                             long stride, long offset) {
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < count * width; t += (long)gridDim.x * blockDim.x) {
    const long v = t / width;
    const int c = (int)(t - v * width);
    aos[t] = soa[c * stride + offset + v];
  }
}
template <typename D, typename S>
__global__ void k_convert(D* __restrict__ dst, const S* __restrict__ src, long count) {
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < count; t += (long)gridDim.x * blockDim.x)
    dst[t] = (D)src[t];
}

// ---------------------------------------------------------------------------
// Per-voxel solver batches, AoS inputs (voxel-contiguous), for kernel-level
// parity and the Solvers:: facade.
// ---------------------------------------------------------------------------
__global__ void k_constrained_batch(long n, const double* __restrict__ R, const double* __restrict__ crss,
                                    const StepConst* __restrict__ Cp, const double* __restrict__ s_in,
                                    const double* __restrict__ lam, const double* __restrict__ dtl,
                                    double* __restrict__ s_out, double* __restrict__ err, int* __restrict__ steps) {
  const StepConst& C = *Cp;
  for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (long)gridDim.x * blockDim.x) {
    double r[9], it[kSlips], P[25], s[5], l[5], d[5];
    for (int e = 0; e < 9; e++) r[e] = R[9 * v + e];
    for (int k = 0; k < kSlips; k++) it[k] = 1.0 / crss[12 * v + k];
    for (int i = 0; i < 5; i++) { s[i] = s_in[5 * v + i]; l[i] = lam[5 * v + i]; d[i] = dtl[5 * v + i]; }
    cf::inverse_map(r, P);
    double Ll[5], q[5], qc[5], A0[15];
    matvec5(C.L0inv, l, Ll);
    for (int i = 0; i < 5; i++) q[i] = Ll[i] + C.edot[i] + d[i];
    cf::rot_t(P, q, qc);
    cf::congruence_chol(P, C.L0inv_w, A0);
    double Q[25], sc[5], G[15], gs[5];
    cf::gram(P, G);
    cf::dev_map(r, Q);
    cf::rot(Q, s, sc);
    double e = err[v];
    int st = 0;
    cf::constrained_nr<0>(sc, G, it, C, qc, A0, 5e-4, 1e9, 100, &e, &st);
    cf::symv(G, sc, gs);
    cf::rot_t(Q, gs, s);
    for (int i = 0; i < 5; i++) s_out[5 * v + i] = s[i];
    err[v] = e;
    steps[v] = st;
  }
}

__global__ void k_taylor_batch(long n, const double* __restrict__ R, const double* __restrict__ crss,
                               const StepConst* __restrict__ Cp, const double* __restrict__ s_in,
                               double* __restrict__ s_out) {
  const StepConst& C = *Cp;
  for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (long)gridDim.x * blockDim.x) {
    double r[9], tau[kSlips], m[kSlips][5], s[5];
    for (int e = 0; e < 9; e++) r[e] = R[9 * v + e];
    for (int k = 0; k < kSlips; k++) tau[k] = crss[12 * v + k];
    for (int i = 0; i < 5; i++) s[i] = s_in[5 * v + i];
    faithful::schmid(r, m);
    faithful::taylor_nr(s, m, tau, C, 5e-4, 1000, s);
    for (int i = 0; i < 5; i++) s_out[5 * v + i] = s[i];
  }
}

// ---------------------------------------------------------------------------
// VPFFT::Solvers (Src/Solvers.h:51-153) over batches of voxels given the way
// the reference passes one: its 12 Schmid tensors in the sample frame (AoS,
// 60 per voxel) and its 12 CRSS; C holds the material, the strain rate and
// L0.  Sample-frame arithmetic in the oracle's operation order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void load_schmid_aos(const double* __restrict__ schmid, long v, double m[kSlips][5]) {
#pragma unroll
  for (int k = 0; k < kSlips; k++)
#pragma unroll
    for (int i = 0; i < 5; i++) m[k][i] = schmid[60 * v + 5 * k + i];
}

template <int OP>
__global__ __launch_bounds__(kPointThreads) void k_solvers(
    long n, const double* __restrict__ schmid, const double* __restrict__ anti, const double* __restrict__ crss,
    const StepConst* __restrict__ Cp, const double* __restrict__ a5, const double* __restrict__ b5,
    const double* __restrict__ c5, double eps, double max_rss, int max_it, double dt, double* __restrict__ acc,
    double* __restrict__ err, double* __restrict__ out) {
  const StepConst& C = *Cp;
  for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (long)gridDim.x * blockDim.x) {
    double m[kSlips][5], tau[kSlips], x[5];
    load_schmid_aos(schmid, v, m);
#pragma unroll
    for (int k = 0; k < kSlips; k++) tau[k] = crss[12 * v + k];
    switch (OP) {
      case SOLVER_TAYLOR: {  // SolveConstitutiveEquations: a5 = initial stress
        double s0[5];
#pragma unroll
        for (int i = 0; i < 5; i++) s0[i] = a5[5 * v + i];
        faithful::taylor_nr(s0, m, tau, C, eps, max_it, x);
#pragma unroll
        for (int i = 0; i < 5; i++) out[5 * v + i] = x[i];
        break;
      }
      case SOLVER_CONSTRAINED: {  // a5 = initial stress, b5 = lambda, c5 = d~
        double s0[5], l[5], d[5];
#pragma unroll
        for (int i = 0; i < 5; i++) { s0[i] = a5[5 * v + i]; l[i] = b5[5 * v + i]; d[i] = c5[5 * v + i]; }
        double e = err[v];
        faithful::constrained_literal(s0, m, tau, C, l, d, eps, max_rss, max_it, x, &e);
        err[v] = e;
#pragma unroll
        for (int i = 0; i < 5; i++) out[5 * v + i] = x[i];
        break;
      }
      case SOLVER_SACHS: {  // SolveSachEquation (Solvers.cpp:230-243)
        faithful::sachs(m, tau, C.edot, x);
#pragma unroll
        for (int i = 0; i < 5; i++) out[5 * v + i] = x[i];
        break;
      }
      case SOLVER_APPLY: {  // ApplyConstitutiveEquations (:255-274): a5 = stress
#pragma clang fp contract(off)
        double s[5], e[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int i = 0; i < 5; i++) s[i] = a5[5 * v + i];
#pragma unroll
        for (int k = 0; k < kSlips; k++) {
          const double rss = faithful::dot5(m[k], s) / tau[k];
          const double g = faithful::slip_rate(C, k, rss), sg = sgn(rss);
#pragma unroll
          for (int i = 0; i < 5; i++) e[i] += m[k][i] * g * sg;
        }
#pragma unroll
        for (int i = 0; i < 5; i++) out[5 * v + i] = e[i];
        break;
      }
      case SOLVER_SHEAR_RATE: {  // CalculateShearRate (:282-302): a5 = stress; out 9 per voxel
#pragma clang fp contract(off)
        double s[5], W[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int i = 0; i < 5; i++) s[i] = a5[5 * v + i];
#pragma unroll
        for (int k = 0; k < kSlips; k++) {
          const double rss = faithful::dot5(m[k], s) / tau[k];
          const double g = faithful::slip_rate(C, k, rss), sg = sgn(rss);
#pragma unroll
          for (int q = 0; q < 9; q++) W[q] += anti[108 * v + 9 * k + q] * g * sg;
        }
#pragma unroll
        for (int q = 0; q < 9; q++) out[9 * v + q] = W[q];
        break;
      }
      case SOLVER_HARDENING: {  // CalculateHardening (:314-373): a5 = stress, acc in/out; out 12 per voxel
#pragma clang fp contract(off)
        double s[5], g[kSlips], local = 0.0;
#pragma unroll
        for (int i = 0; i < 5; i++) s[i] = a5[5 * v + i];
#pragma unroll
        for (int k = 0; k < kSlips; k++) {
          g[k] = faithful::slip_rate(C, k, faithful::dot5(m[k], s) / tau[k]);
          local += fabs(g[k]) * dt;
        }
        const double A = acc[v];
#pragma unroll
        for (int i = 0; i < kSlips; i++) {
          double dtau = 0.0;
#pragma unroll
          for (int j = 0; j < kSlips; j++) dtau += C.hmat[i * kSlips + j] * fabs(g[j]) * dt;
          const double tiny = 1e-4 * C.tau0[i];
          double voce = 0.0;
          if (fabs(C.theta0[i]) > tiny) {
            voce = C.theta1[i] * local;
            if (fabs(C.tau1[i]) > tiny) {
              const double fr = fabs(C.theta0[i] / C.tau1[i]);
              const double e_ini = exp(-A * fr);
              const double e_del = exp(-local * fr);
              voce = voce - (fr * C.tau1[i] * C.theta1[i]) / fr * e_ini * (e_del - 1.0) -
                     C.theta1[i] / fr * e_ini * (e_del * ((A + local) * fr + 1.0) - (A * fr + 1.0));
            }
          }
          out[12 * v + i] = tau[i] + dtau * voce / local;
        }
        acc[v] = A + local;
        break;
      }
    }
  }
}

hipError_t launch_solvers(int op, long n, const double* schmid, const double* anti, const double* crss,
                          const StepConst* C, const double* a5, const double* b5, const double* c5, double eps,
                          double max_rss, int max_it, double dt, double* acc, double* err, double* out,
                          hipStream_t st) {
  auto go = [&](auto kernel) {
    hipLaunchKernelGGL(kernel, dim3(point_blocks(n)), dim3(kPointThreads), 0, st, n, schmid, anti, crss, C, a5, b5,
                       c5, eps, max_rss, max_it, dt, acc, err, out);
    return hipGetLastError();
  };
  switch (op) {
    case SOLVER_TAYLOR: return go(k_solvers<SOLVER_TAYLOR>);
    case SOLVER_CONSTRAINED: return go(k_solvers<SOLVER_CONSTRAINED>);
    case SOLVER_SACHS: return go(k_solvers<SOLVER_SACHS>);
    case SOLVER_APPLY: return go(k_solvers<SOLVER_APPLY>);
    case SOLVER_SHEAR_RATE: return go(k_solvers<SOLVER_SHEAR_RATE>);
    case SOLVER_HARDENING: return go(k_solvers<SOLVER_HARDENING>);
  }
  return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
hipError_t launch_prologue(int prec, bool uniform_crss, long n, const double* R, const double* crss,
                           const StepConst* C, void* field, double* sigma, double* lambda, double* partial,
                           HealthRecord* health, hipStream_t st) {
  const int nb = point_blocks(n);
#define PRO_LAUNCH(TT, UU)                                                                                        \
  hipLaunchKernelGGL((k_prologue<TT, UU>), dim3(nb), dim3(kPointThreads), 0, st, n, R, crss, C, (TT*)field, sigma, \
                     lambda, partial, health)
  if (prec == 0) {
    if (uniform_crss) PRO_LAUNCH(double, true); else PRO_LAUNCH(double, false);
  } else {
    if (uniform_crss) PRO_LAUNCH(float, true); else PRO_LAUNCH(float, false);
  }
#undef PRO_LAUNCH
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_secant, dim3(nb), dim3(kPointThreads), 0, st, n, R, crss, sigma, C, partial, health);
  return hipGetLastError();
}

hipError_t launch_L0_finalize(long n, int nblocks, const double* partial, StepConst* C, GreenConst<double>* g64,
                              GreenConst<float>* g32, PrologueRecord* rec, HealthRecord* health, hipStream_t st) {
  hipLaunchKernelGGL(k_L0_finalize, dim3(1), dim3(256), 0, st, n, nblocks, partial, C, g64, g32, rec, health);
  return hipGetLastError();
}

hipError_t launch_derive_from_L0(StepConst* C, GreenConst<double>* g64, GreenConst<float>* g32, hipStream_t st) {
  hipLaunchKernelGGL(k_derive, dim3(1), dim3(64), 0, st, C, g64, g32);
  return hipGetLastError();
}

hipError_t launch_nr(int prec, int nexp_uniform, bool uniform_tau, long n, const void* dtil, void* phi, double* sigma,
                     double* lambda, const double* R, const double* crss, const StepConst* C, double* partial,
                     const LoopState* ls, HealthRecord* health, int row_runs, hipStream_t st) {
  if (row_runs != 0 && row_runs != 1 && row_runs != 4) return hipErrorInvalidValue;
  if (row_runs == 4 && n % 256) return hipErrorInvalidValue;
  const int nb = nr_rows(n, row_runs);
  ZFusedArgs zf{};
  zf.linear = row_runs == 1 ? 1 : 0;
#define NR_LAUNCH(TT, NN, UU)                                                                                    \
  if (row_runs == 4)                                                                                             \
    hipLaunchKernelGGL((k_nr<TT, NN, UU, NR_ROWS256>), dim3(nb), dim3(kNrThreads), 0, st, n, (const TT*)dtil,   \
                       (TT*)phi, sigma, lambda, R, crss, C, partial, ls, health, VoxGroup{}, zf);                 \
  else                                                                                                           \
    hipLaunchKernelGGL((k_nr<TT, NN, UU, NR_SINGLE>), dim3(nb), dim3(kNrThreads), 0, st, n, (const TT*)dtil,    \
                       (TT*)phi, sigma, lambda, R, crss, C, partial, ls, health, VoxGroup{}, zf)
#define NR_N(TT, UU)                                 \
  switch (nexp_uniform) {                             \
    case 5: NR_LAUNCH(TT, 5, UU); break;              \
    case 10: NR_LAUNCH(TT, 10, UU); break;            \
    case 20: NR_LAUNCH(TT, 20, UU); break;            \
    default: NR_LAUNCH(TT, 0, UU); break;             \
  }
  if (prec == 0) {
    if (uniform_tau) { NR_N(double, true) } else { NR_N(double, false) }
  } else {
    if (uniform_tau) { NR_N(float, true) } else { NR_N(float, false) }
  }
#undef NR_N
#undef NR_LAUNCH
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // fold the per-block partials in a fixed order: nb -> ceil(nb / 256) rows
  const int nb2 = (nb + 255) / 256;
  hipLaunchKernelGGL(k_fold_partials, dim3(nb2), dim3(256), 0, st, partial, nb, partial + (long)nb * kPartialStride,
                     ls ? &ls->done : nullptr);
  return hipGetLastError();
}

bool zfused_eligible(const FftGeom& g) { return g.nz == 256 || g.nz == 512; }

hipError_t launch_zfused(int prec, int nexp_uniform, bool uniform_tau, long n, const void* spec_in, void* spec_out,
                         const FftGeom& g, const void* tw, double scale, double* sigma, double* lambda,
                         const double* R, const double* crss, const StepConst* C, double* partial,
                         const LoopState* ls, HealthRecord* health, hipStream_t st) {
  if (!zfused_eligible(g)) return hipErrorInvalidValue;
  const unsigned blocks = (unsigned)g.nrows;
  const ZFusedArgs zf{spec_in, spec_out, tw, g.nc, g.nzp, scale, 0};
#define ZF_LAUNCH(TT, NN, UU)                                                                                        \
  if (g.nz == 512)                                                                                                   \
    hipLaunchKernelGGL((k_nr<TT, NN, UU, NR_ZFUSED512>), dim3(blocks), dim3(kNrThreads), 0, st, n,                   \
                       (const TT*)nullptr, (TT*)nullptr, sigma, lambda, R, crss, C, partial, ls, health, VoxGroup{},  \
                       zf);                                                                                          \
  else                                                                                                               \
    hipLaunchKernelGGL((k_nr<TT, NN, UU, NR_ZFUSED>), dim3(blocks), dim3(kNrThreads), 0, st, n, (const TT*)nullptr,  \
                       (TT*)nullptr, sigma, lambda, R, crss, C, partial, ls, health, VoxGroup{}, zf)
#define ZF_N(TT, UU)                                 \
  switch (nexp_uniform) {                             \
    case 5: ZF_LAUNCH(TT, 5, UU); break;              \
    case 10: ZF_LAUNCH(TT, 10, UU); break;            \
    case 20: ZF_LAUNCH(TT, 20, UU); break;            \
    default: ZF_LAUNCH(TT, 0, UU); break;             \
  }
  if (prec == 0) {
    if (uniform_tau) { ZF_N(double, true) } else { ZF_N(double, false) }
  } else {
    if (uniform_tau) { ZF_N(float, true) } else { ZF_N(float, false) }
  }
#undef ZF_N
#undef ZF_LAUNCH
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // the same partial rows as launch_nr's with row_runs = 4 (Nz = 256: one per
  // z row) or 1 (Nz = 512: one per run of 64 voxels), folded the same way
  const int nb = nr_rows(n, g.nz == 256 ? 4 : 1);
  const int nb2 = (nb + 255) / 256;
  hipLaunchKernelGGL(k_fold_partials, dim3(nb2), dim3(256), 0, st, partial, nb, partial + (long)nb * kPartialStride,
                     ls ? &ls->done : nullptr);
  return hipGetLastError();
}

hipError_t launch_phi_rebuild(int prec, long n, const void* dtil, void* phi, const double* sigma, const StepConst* C,
                              hipStream_t st) {
  if (prec == 0)
    hipLaunchKernelGGL(k_phi_rebuild<double>, dim3(point_blocks(n)), dim3(256), 0, st, n, (const double*)dtil,
                       (double*)phi, sigma, C);
  else
    hipLaunchKernelGGL(k_phi_rebuild<float>, dim3(point_blocks(n)), dim3(256), 0, st, n, (const float*)dtil,
                       (float*)phi, sigma, C);
  return hipGetLastError();
}

hipError_t launch_iter_finalize(long n, int row_runs, const double* partial, const StepConst* C, double tol,
                                int max_iter, LoopState* ls, IterRecord* rec, hipStream_t st) {
  // launch_nr / launch_zfused left ceil(rows / 256) folded rows after the per-wave partials
  const int nb = nr_rows(n, row_runs);
  const int nb2 = (nb + 255) / 256;
  hipLaunchKernelGGL(k_iter_finalize, dim3(1), dim3(256), 0, st, n, nb2, partial + (long)nb * kPartialStride, C,
                     tol, max_iter, ls, rec);
  return hipGetLastError();
}

hipError_t launch_epilogue(long n, double* R, double* crss, double* acc, const double* sigma, const StepConst* C,
                           HealthRecord* health, hipStream_t st) {
  hipLaunchKernelGGL(k_epilogue, dim3(point_blocks(n)), dim3(kPointThreads), 0, st, n, R, crss, acc, sigma, C,
                     health);
  return hipGetLastError();
}

static int cvt_blocks(long work) {
  long b = (work + 255) / 256;
  if (b > 8192) b = 8192;
  return (int)(b < 1 ? 1 : b);
}

hipError_t launch_aos_to_soa(const double* aos, double* soa, long count, int width, long stride, long offset,
                             hipStream_t st) {
  hipLaunchKernelGGL(k_aos_to_soa, dim3(cvt_blocks(count * width)), dim3(256), 0, st, aos, soa, count, width, stride,
                     offset);
  return hipGetLastError();
}
hipError_t launch_soa_to_aos(const double* soa, double* aos, long count, int width, long stride, long offset,
                             hipStream_t st) {
  hipLaunchKernelGGL(k_soa_to_aos, dim3(cvt_blocks(count * width)), dim3(256), 0, st, soa, aos, count, width, stride,
                     offset);
  return hipGetLastError();
}
hipError_t launch_convert(int prec_dst, void* dst, int prec_src, const void* src, long count, hipStream_t st) {
  const dim3 grid(cvt_blocks(count));
  if (prec_dst == 0 && prec_src == 1)
    hipLaunchKernelGGL((k_convert<double, float>), grid, dim3(256), 0, st, (double*)dst, (const float*)src, count);
  else if (prec_dst == 1 && prec_src == 0)
    hipLaunchKernelGGL((k_convert<float, double>), grid, dim3(256), 0, st, (float*)dst, (const double*)src, count);
  else if (prec_dst == 0)
    hipLaunchKernelGGL((k_convert<double, double>), grid, dim3(256), 0, st, (double*)dst, (const double*)src, count);
  else
    hipLaunchKernelGGL((k_convert<float, float>), grid, dim3(256), 0, st, (float*)dst, (const float*)src, count);
  return hipGetLastError();
}

hipError_t launch_constrained_batch(long n, const double* R, const double* crss, const StepConst* C,
                                    const double* s_in, const double* lam, const double* dt, double* s_out,
                                    double* err, int* steps, hipStream_t st) {
  hipLaunchKernelGGL(k_constrained_batch, dim3(point_blocks(n)), dim3(kPointThreads), 0, st, n, R, crss, C, s_in,
                     lam, dt, s_out, err, steps);
  return hipGetLastError();
}
hipError_t launch_taylor_batch(long n, const double* R, const double* crss, const StepConst* C, const double* s_in,
                               double* s_out, hipStream_t st) {
  hipLaunchKernelGGL(k_taylor_batch, dim3(point_blocks(n)), dim3(kPointThreads), 0, st, n, R, crss, C, s_in, s_out);
  return hipGetLastError();
}

// ---- batched launches --------------------------------------------------------
hipError_t launch_prologue_group(int prec, const VoxGroup& grp, long blocks, hipStream_t st) {
  if (prec == 0)
    hipLaunchKernelGGL(k_prologue_b<double>, dim3((unsigned)blocks), dim3(kPointThreads), 0, st, grp);
  else
    hipLaunchKernelGGL(k_prologue_b<float>, dim3((unsigned)blocks), dim3(kPointThreads), 0, st, grp);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_secant_b, dim3((unsigned)blocks), dim3(kPointThreads), 0, st, grp);
  return hipGetLastError();
}
hipError_t launch_L0_finalize_group(const VoxGroup& grp, hipStream_t st) {
  hipLaunchKernelGGL(k_L0_finalize_b, dim3(grp.count), dim3(256), 0, st, grp);
  return hipGetLastError();
}
hipError_t launch_loop_reset_group(const VoxGroup& grp, hipStream_t st) {
  hipLaunchKernelGGL(k_loop_reset_b, dim3((grp.count + 63) / 64), dim3(64), 0, st, grp);
  return hipGetLastError();
}
hipError_t launch_nr_group(int prec, int nexp_uniform, bool uniform_tau, const VoxGroup& grp, long blocks,
                           hipStream_t st) {
#define NR_LAUNCH(TT, NN, UU)                                                                                   \
  hipLaunchKernelGGL((k_nr<TT, NN, UU, NR_BATCHED>), dim3((unsigned)blocks), dim3(kNrThreads), 0, st, 0L,       \
                     (const TT*)nullptr, (TT*)nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, \
                     nullptr, grp, ZFusedArgs{})
#define NR_N(TT, UU)                                 \
  switch (nexp_uniform) {                             \
    case 5: NR_LAUNCH(TT, 5, UU); break;              \
    case 10: NR_LAUNCH(TT, 10, UU); break;            \
    case 20: NR_LAUNCH(TT, 20, UU); break;            \
    default: NR_LAUNCH(TT, 0, UU); break;             \
  }
  if (prec == 0) {
    if (uniform_tau) { NR_N(double, true) } else { NR_N(double, false) }
  } else {
    if (uniform_tau) { NR_N(float, true) } else { NR_N(float, false) }
  }
#undef NR_N
#undef NR_LAUNCH
  return hipGetLastError();
}
hipError_t launch_iter_finalize_group(const VoxGroup& grp, double tol, int max_iter, hipStream_t st) {
  hipLaunchKernelGGL(k_iter_finalize_b, dim3(grp.count), dim3(256), 0, st, grp, tol, max_iter);
  return hipGetLastError();
}
hipError_t launch_epilogue_group(const VoxGroup& grp, long blocks, hipStream_t st) {
  hipLaunchKernelGGL(k_epilogue_b, dim3((unsigned)blocks), dim3(kPointThreads), 0, st, grp);
  return hipGetLastError();
}

}  // namespace vpfft
