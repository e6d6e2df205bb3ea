// fft_green.h -- plan/geometry structs and launchers for the FFT + Green passes.
#pragma once
#include <hip/hip_runtime.h>

namespace vpfft {

struct FftGeom {
  long n;      // voxels per component (Nx*Ny*Nz)
  long nc;     // spectrum stride per component (Nx*Ny*nzp)
  long nrows;  // Nx*Ny
  int nx, ny, nz, nzh;
  int nzp;     // spectrum row stride: Nz/2+1 rounded up to 16, so every
               // row starts on a 128-byte line in FP64 and FP32
};

inline int padded_half(int nz) { return ((nz / 2 + 1) + 15) / 16 * 16; }

struct AxisPlan {
  int n;
  int nstages;
  int radix[12];
  unsigned long long code;  // radix[s] in bits [4s, 4s+4): read with shifts, no indexed array
};

// Per-step Green-operator constants in the transform precision.
template <typename T>
struct GreenConst {
  T Kc[6][6];     // K(k)_r = sum_q Kc[r][q] mono_q(k); r over (00,11,22,12,02,01),
                  // q over (k0k0,k1k1,k2k2,k1k2,k0k2,k0k1)
  T kscale[3];    // 1/(N_a * voxel length_a)
  T L0inv[25];    // column-major
  int nx, ny, nz;
};

// Factor n into radices {8,4,2,3,5,7}; returns false for other prime factors.
bool make_axis_plan(int n, AxisPlan* ap);
// The same with the radices in the given order of preference (a permutation
// of {8,4,2,3,5,7}); the stage order changes the rounding, not the result.
bool make_axis_plan_ordered(int n, const int order[6], AxisPlan* ap);

// One transform problem of a batched launch (a whole grid).
struct FftJob {
  FftGeom g;
  AxisPlan ax, ay, az;
  const void* tw[3];  // twiddles for x, y, z in the transform precision
  const void* field;  // real input of the forward z pass (phi)
  void* dtil;         // real output of the inverse z pass
  void* spec;         // half spectrum
  const void* gc;     // GreenConst<T>
  const int* done;    // the job's loop flag (may be null)
};
// A batched launch: blocks [off[m], off[m+1]) belong to jobs[ids[m]].
struct FftGroup {
  const FftJob* jobs;
  const int* ids;
  const long* off;
  int count;
};

// Block shape of a pass for one grid: B pencils (rows or runs) per block,
// E values per thread (x pass), blocks per launch, dynamic LDS bytes.
// B = 0 picks the default; other B values are checked against the LDS budget.
struct PassShape {
  int B, E;
  long blocks;
  size_t lds;
  bool ok;
};
enum FftPass { kPassZ = 0, kPassY = 1, kPassX = 2 };
// The register form of a pass (one thread per pencil, the pencil in
// registers): pass_shape(prec, pass, g, kRegForm) gives B = -(axis length)
// when that length has one (batched y passes of 16, 24, 32 points, and 48
// and 64 in FP32), else ok = false.
constexpr int kRegForm = -1;
bool reg_form_supported(int prec, FftPass pass, int n);
// the register form's radix stages for length n, in the order they run; returns the count
int reg_form_radices(int n, int* out);
PassShape pass_shape(int prec, FftPass pass, const FftGeom& g, int B);
// candidate B values of a pass (for the tuner), returns the count
int pass_candidates(FftPass pass, int* out, int cap);

hipError_t launch_z_r2c_group(int prec, int B, const FftGroup& grp, long blocks, size_t lds, hipStream_t st);
hipError_t launch_z_c2r_group(int prec, int B, const FftGroup& grp, long blocks, size_t lds, hipStream_t st);
hipError_t launch_y_group(int prec, int B, const FftGroup& grp, long blocks, size_t lds, int sign, hipStream_t st);
hipError_t launch_x_green_group(int prec, int B, int E, const FftGroup& grp, long blocks, size_t lds,
                                hipStream_t st);

// prec: 0 = FP64 storage/compute, 1 = FP32.  B: block shape (pencils per
// block, see PassShape; 0 = the default).  done: optional device flag; a
// kernel whose flag is set returns at entry (graph-captured iterations after
// convergence do nothing).
hipError_t launch_z_r2c(int prec, const void* f, void* s, const FftGeom& g, const AxisPlan& ap, const void* tw, int B,
                        const int* done, hipStream_t st);
hipError_t launch_z_c2r(int prec, const void* s, void* f, const FftGeom& g, const AxisPlan& ap, const void* tw,
                        double scale, int B, const int* done, hipStream_t st);
hipError_t launch_y(int prec, void* s, const FftGeom& g, const AxisPlan& ap, const void* tw, int sign, int B,
                    const int* done, hipStream_t st);
hipError_t launch_x_green(int prec, void* s, const FftGeom& g, const AxisPlan& ap, const void* tw, const void* gc,
                          int B, const int* done, hipStream_t st);

}  // namespace vpfft
