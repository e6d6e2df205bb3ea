// fft_green.hip -- 3-D R2C/C2R transforms of the 5-component polarization
// field with the Green-operator convolution fused into the x pass.
//
// Replaces the reference's 5 forward + 5 backward in-place FFTW C2C plans and
// its per-frequency Green loop (Src/MaterialGrid.cpp:624-761).
//
// Layout in HBM (component-major, like the reference's one buffer per
// component, MaterialGrid.h:320-321):
//   real field   f[c][x][y][z]            c = 0..4, N = Nx*Ny*Nz per component
//   spectrum     s[c][x][y][kz]           kz = 0..Nz/2 in rows of nzp >= Nz/2+1
//                (padded to a multiple of 16 so runs of kz start on 128-byte
//                lines; the pad bins are never read or written)
// Passes per fixed-point iteration:
//   z_r2c : f -> s   (two real components packed in one complex FFT)
//   y     : s in place, forward
//   x     : s in place: forward, Green, inverse (all 5 components per block)
//   y     : s in place, inverse
//   z_c2r : s -> f   (scaled by 1/N)
// Each pass stages a batch of pencils in LDS and runs a mixed-radix
// (2,3,4,5,7,8) Stockham FFT there; global loads and stores move whole
// contiguous runs (rows along z, runs of consecutive kz for y and x).
//
// The reference's special frequency planes (index N/2 and N/2+1 on every
// even axis get L_inv instead of the Green operator, MaterialGrid.cpp:710-713)
// are not Hermitian-symmetric.  It keeps Re(IFFT(D)) (MaterialGrid.h:354-363),
// which equals the C2R transform of D'(k) = (D(k) + conj D(-k))/2, i.e. of the
// operator G_eff(k) = (G(k) + G(-k))/2.  That is what the x pass applies.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "device_common.h"
#include "fft_green.h"
#include "fft_regs.h"

// No multiply-add contraction in this file: with it the compiler's choice of
// which products to fuse can differ between the block-shape instantiations
// of a pass, and a transform's bits would depend on its block shape.  The
// passes are bound by HBM traffic, not by these instructions.
#pragma STDC FP_CONTRACT OFF

namespace vpfft {

// cos / sin (2 pi t / R) for the odd radices
__device__ __constant__ static const double kCos3[3] = {1.0, -0.5, -0.5};
__device__ __constant__ static const double kSin3[3] = {0.0, 0.86602540378443864676, -0.86602540378443864676};
__device__ __constant__ static const double kCos5[5] = {1.0, 0.30901699437494742410, -0.80901699437494742410,
                                                        -0.80901699437494742410, 0.30901699437494742410};
__device__ __constant__ static const double kSin5[5] = {0.0, 0.95105651629515357212, 0.58778525229247312917,
                                                        -0.58778525229247312917, -0.95105651629515357212};
__device__ __constant__ static const double kCos7[7] = {1.0, 0.62348980185873353053, -0.22252093395631440429,
                                                        -0.90096886790241912624, -0.90096886790241912624,
                                                        -0.22252093395631440429, 0.62348980185873353053};
__device__ __constant__ static const double kSin7[7] = {0.0, 0.78183148246802980871, 0.97492791218182360702,
                                                        0.43388373911755812048, -0.43388373911755812048,
                                                        -0.97492791218182360702, -0.78183148246802980871};

template <typename T, int R>
__device__ __forceinline__ void dft_odd(typename C2<T>::t* v, int sign) {
  using C = typename C2<T>::t;
  const double* cs = R == 3 ? kCos3 : (R == 5 ? kCos5 : kCos7);
  const double* sn = R == 3 ? kSin3 : (R == 5 ? kSin5 : kSin7);
  C out[R];
#pragma unroll
  for (int k = 0; k < R; k++) {
    T re = v[0].x, im = v[0].y;
#pragma unroll
    for (int r = 1; r < R; r++) {
      const int t = (r * k) % R;
      const T c = (T)cs[t], s = (T)(sign * sn[t]);
      re = fmad(-v[r].y, s, fmad(v[r].x, c, re));
      im = fmad(v[r].y, c, fmad(v[r].x, s, im));
    }
    out[k].x = re;
    out[k].y = im;
  }
#pragma unroll
  for (int k = 0; k < R; k++) v[k] = out[k];
}

template <typename T, int R>
__device__ __forceinline__ void dft(typename C2<T>::t* v, int sign) {
  using C = typename C2<T>::t;
  if constexpr (R == 2) {
    C a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 4) {
    dft4<T>(v, sign);
  } else if constexpr (R == 8) {
    C e[4] = {v[0], v[2], v[4], v[6]};
    C o[4] = {v[1], v[3], v[5], v[7]};
    dft<T, 4>(e, sign);
    dft<T, 4>(o, sign);
    const T h = (T)0.70710678118654752440;
    // W8^k = exp(sign i pi k / 4)
    C w1 = cmake<T>(h, sign * h);
    C w3 = cmake<T>(-h, sign * h);
    C t0 = o[0];
    C t1 = cmul(o[1], w1);
    C t2 = cmul_si(o[2], sign);
    C t3 = cmul(o[3], w3);
    v[0] = cadd(e[0], t0); v[4] = csub(e[0], t0);
    v[1] = cadd(e[1], t1); v[5] = csub(e[1], t1);
    v[2] = cadd(e[2], t2); v[6] = csub(e[2], t2);
    v[3] = cadd(e[3], t3); v[7] = csub(e[3], t3);
  } else {
    dft_odd<T, R>(v, sign);
  }
}

// One Stockham stage over P pencils held in LDS as buf[j * LS + p], LS >= P
// (an odd LS breaks the bank pattern of the row-transposing loads).
template <typename T, int P, int R, int LS = P>
__device__ __forceinline__ void stockham_stage(const typename C2<T>::t* __restrict__ in,
                                               typename C2<T>::t* __restrict__ out, int n, int ns,
                                               const typename C2<T>::t* __restrict__ tw, int sign) {
  using C = typename C2<T>::t;
  const int nb = n / R;
  const int total = nb * P;
  const int twstride = n / (ns * R);
  for (int u = threadIdx.x; u < total; u += blockDim.x) {
    const int p = u % P;
    const int j = u / P;
    C v[R];
#pragma unroll
    for (int r = 0; r < R; r++) v[r] = in[(j + r * nb) * LS + p];
    const int jm = j % ns;
    if (ns > 1) {
#pragma unroll
      for (int r = 1; r < R; r++) {
        C w = tw[r * jm * twstride];
        if (sign > 0) w.y = -w.y;
        v[r] = cmul(v[r], w);
      }
    }
    dft<T, R>(v, sign);
    const int idxd = (j - jm) * R + jm;
#pragma unroll
    for (int r = 0; r < R; r++) out[(idxd + r * ns) * LS + p] = v[r];
  }
}

// Values every lane of the block holds alike, moved to scalar registers so
// the FFT's loop bounds and radix switches stay uniform control flow.
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ long uni(long x) {
  const unsigned long u = (unsigned long)x;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
  return (long)(((unsigned long)hi << 32) | lo);
}
template <typename P> __device__ __forceinline__ P* uni(P* p) { return (P*)uni((long)p); }
// A uniform pointer read from a job table, made directly in address space AS
// (1 global, 4 constant) from its integer value.  A generic pointer cast to
// AS and back is folded back to generic, and its loads become flat loads (and
// constant structs land in vector registers instead of scalar ones).
template <int AS, typename P> __device__ __forceinline__ P* uni_as(P* p) {
  const unsigned long u = (unsigned long)uni((long)p);
  return (P*)(__attribute__((address_space(AS))) P*)u;
}
template <typename P> __device__ __forceinline__ P* guni(P* p) { return uni_as<1>(p); }

// The GPU hands consecutive blocks to its 8 XCDs in turn, and each XCD has
// its own L2.  Block b of n runs on XCD b % 8; renumbering so that each XCD
// takes a contiguous range of tiles keeps neighbouring tiles (the two halves
// of a 128-byte line, for 64-byte runs) in one L2.
__device__ __forceinline__ long xcd_remap(long b, long n) {
  const long x = b & 7, i = b >> 3;
  const long per = n >> 3, rem = n & 7;
  return x * per + (x < rem ? x : rem) + i;
}

// Full transform of P pencils of length n.  Data starts in buf0; returns the
// buffer holding the result (buf0 or buf1).  Ends with a barrier.
template <typename T, int P, int LS = P>
__device__ __forceinline__ typename C2<T>::t* lds_fft(typename C2<T>::t* buf0, typename C2<T>::t* buf1,
                                      const AxisPlan& ap, const typename C2<T>::t* tw, int sign) {
  using C = typename C2<T>::t;
  C* in = buf0;
  C* out = buf1;
  int ns = 1;
  const int n = uni(ap.n), nstages = uni(ap.nstages);
  const unsigned long long code = (unsigned long long)uni((long)ap.code);
  for (int s = 0; s < nstages; s++) {
    const int r = (int)((code >> (4 * s)) & 15);
    switch (r) {
      case 2: stockham_stage<T, P, 2, LS>(in, out, n, ns, tw, sign); break;
      case 3: stockham_stage<T, P, 3, LS>(in, out, n, ns, tw, sign); break;
      case 4: stockham_stage<T, P, 4, LS>(in, out, n, ns, tw, sign); break;
      case 5: stockham_stage<T, P, 5, LS>(in, out, n, ns, tw, sign); break;
      case 7: stockham_stage<T, P, 7, LS>(in, out, n, ns, tw, sign); break;
      default: stockham_stage<T, P, 8, LS>(in, out, n, ns, tw, sign); break;
    }
    __syncthreads();
    ns *= r;
    C* t = in;
    in = out;
    out = t;
  }
  return in;
}

// ---------------------------------------------------------------------------
// z pass, real -> half spectrum.  Block = B consecutive (x,y) rows.
// Pencil p = pair * B + b holds f[2 pair] + i f[2 pair + 1] of row b.
// ---------------------------------------------------------------------------
template <typename T, int B>
__device__ __forceinline__ void z_r2c_body(const T* __restrict__ f, typename C2<T>::t* __restrict__ s,
                                           const FftGeom& g, const AxisPlan& ap,
                                           const typename C2<T>::t* __restrict__ tw, long blk) {
  using C = typename C2<T>::t;
  constexpr int P = 3 * B;
  constexpr int LS = P;
  extern __shared__ unsigned char smem_raw[];
  C* buf0 = reinterpret_cast<C*>(smem_raw);
  C* buf1 = buf0 + LS * g.nz;
  const long row0 = blk * B;
  const int nrows = (int)min((long)B, g.nrows - row0);
  const int nz = g.nz, nzh = g.nzh, nzp = g.nzp;
  const int cnt = B * nz;
  for (int pair = 0; pair < 3; pair++) {
    const T* fa = f + (long)(2 * pair) * g.n + row0 * nz;
    const T* fb = pair < 2 ? f + (long)(2 * pair + 1) * g.n + row0 * nz : nullptr;
    for (int e = threadIdx.x; e < cnt; e += blockDim.x) {
      const int b = e / nz, j = e - b * nz;
      T a = 0, c = 0;
      if (b < nrows) {
        a = fa[e];
        if (fb) c = fb[e];
      }
      buf0[j * LS + pair * B + b] = cmake<T>(a, c);
    }
  }
  __syncthreads();
  C* res = lds_fft<T, P, LS>(buf0, buf1, ap, tw, -1);
  const int cnth = B * nzh;
  for (int pair = 0; pair < 3; pair++) {
    C* sa = s + (long)(2 * pair) * g.nc + row0 * nzp;
    C* sb = s + (long)(2 * pair + 1) * g.nc + row0 * nzp;
    for (int e = threadIdx.x; e < cnth; e += blockDim.x) {
      const int b = e / nzh, k = e - b * nzh;
      if (b >= nrows) continue;
      const C zk = res[k * LS + pair * B + b];
      const int kn = k == 0 ? 0 : nz - k;
      const C zn = res[kn * LS + pair * B + b];
      // X_a = (Z_k + conj Z_{-k}) / 2 ;  X_b = -i/2 (Z_k - conj Z_{-k})
      const T h = (T)0.5;
      sa[b * nzp + k] = cmake<T>(h * (zk.x + zn.x), h * (zk.y - zn.y));
      if (pair < 2) sb[b * nzp + k] = cmake<T>(h * (zk.y + zn.y), -h * (zk.x - zn.x));
    }
  }
}

template <typename T, int B>
__global__ void __launch_bounds__(256) k_z_r2c(const T* __restrict__ f, typename C2<T>::t* __restrict__ s,
                                               FftGeom g, AxisPlan ap, const typename C2<T>::t* __restrict__ tw,
                                               const int* __restrict__ done) {
  if (done && *done) return;
  z_r2c_body<T, B>(f, s, g, ap, tw, blockIdx.x);
}

// ---------------------------------------------------------------------------
// z pass, half spectrum -> real, scaled by 1/N.
// ---------------------------------------------------------------------------
template <typename T, int B>
__device__ __forceinline__ void z_c2r_body(const typename C2<T>::t* __restrict__ s, T* __restrict__ f,
                                           const FftGeom& g, const AxisPlan& ap,
                                           const typename C2<T>::t* __restrict__ tw, T scale, long blk) {
  using C = typename C2<T>::t;
  constexpr int P = 3 * B;
  constexpr int LS = P;
  extern __shared__ unsigned char smem_raw[];
  C* buf0 = reinterpret_cast<C*>(smem_raw);
  C* buf1 = buf0 + LS * g.nz;
  const long row0 = blk * B;
  const int nrows = (int)min((long)B, g.nrows - row0);
  const int nz = g.nz, nzh = g.nzh, nzp = g.nzp;
  const int cnth = B * nzh;
  const bool even = (nz & 1) == 0;
  for (int pair = 0; pair < 3; pair++) {
    const C* sa = s + (long)(2 * pair) * g.nc + row0 * nzp;
    const C* sb = s + (long)(2 * pair + 1) * g.nc + row0 * nzp;
    for (int e = threadIdx.x; e < cnth; e += blockDim.x) {
      const int b = e / nzh, k = e - b * nzh;
      C xa = cmake<T>(0, 0), xb = cmake<T>(0, 0);
      if (b < nrows) {
        xa = sa[b * nzp + k];
        if (pair < 2) xb = sb[b * nzp + k];
      }
      // C2R reads only the real part of the self-conjugate bins
      if (k == 0 || (even && 2 * k == nz)) { xa.y = 0; xb.y = 0; }
      // Z_k = X_a + i X_b ;  Z_{-k} = conj X_a + i conj X_b
      buf0[k * LS + pair * B + b] = cmake<T>(xa.x - xb.y, xa.y + xb.x);
      if (k > 0 && 2 * k != nz) buf0[(nz - k) * LS + pair * B + b] = cmake<T>(xa.x + xb.y, xb.x - xa.y);
    }
  }
  __syncthreads();
  C* res = lds_fft<T, P, LS>(buf0, buf1, ap, tw, +1);
  const int cnt = B * nz;
  for (int pair = 0; pair < 3; pair++) {
    T* fa = f + (long)(2 * pair) * g.n + row0 * nz;
    T* fb = f + (long)(2 * pair + 1) * g.n + row0 * nz;
    for (int e = threadIdx.x; e < cnt; e += blockDim.x) {
      const int b = e / nz, j = e - b * nz;
      if (b >= nrows) continue;
      const C z = res[j * LS + pair * B + b];
      fa[e] = z.x * scale;
      if (pair < 2) fb[e] = z.y * scale;
    }
  }
}

template <typename T, int B>
__global__ void __launch_bounds__(256) k_z_c2r(const typename C2<T>::t* __restrict__ s, T* __restrict__ f,
                                               FftGeom g, AxisPlan ap, const typename C2<T>::t* __restrict__ tw,
                                               T scale, const int* __restrict__ done) {
  if (done && *done) return;
  z_c2r_body<T, B>(s, f, g, ap, tw, scale, blockIdx.x);
}

// ---------------------------------------------------------------------------
// y pass, in place.  Block = (component c, x, run of B consecutive kz).
// ---------------------------------------------------------------------------
template <typename T, int B>
__device__ __forceinline__ void y_body(typename C2<T>::t* __restrict__ s, const FftGeom& g, const AxisPlan& ap,
                                       const typename C2<T>::t* __restrict__ tw, int sign, long blk) {
  using C = typename C2<T>::t;
  extern __shared__ unsigned char smem_raw[];
  C* buf0 = reinterpret_cast<C*>(smem_raw);
  C* buf1 = buf0 + B * g.ny;
  const int runs = (g.nzh + B - 1) / B;
  const int run = (int)(blk % runs);
  const long plane = blk / runs;  // c * nx + x
  const int kz0 = run * B;
  const int nb = min(B, g.nzh - kz0);
  C* base = s + plane * (long)g.ny * g.nzp + kz0;
  const int cnt = B * g.ny;
  for (int e = threadIdx.x; e < cnt; e += blockDim.x) {
    const int b = e % B, y = e / B;
    buf0[e] = b < nb ? base[(long)y * g.nzp + b] : cmake<T>(0, 0);
  }
  __syncthreads();
  C* res = lds_fft<T, B>(buf0, buf1, ap, tw, sign);
  for (int e = threadIdx.x; e < cnt; e += blockDim.x) {
    const int b = e % B, y = e / B;
    if (b < nb) base[(long)y * g.nzp + b] = res[e];
  }
}

template <typename T, int B>
__global__ void __launch_bounds__(256) k_y(typename C2<T>::t* __restrict__ s, FftGeom g, AxisPlan ap,
                                           const typename C2<T>::t* __restrict__ tw, int sign,
                                           const int* __restrict__ done) {
  if (done && *done) return;
  y_body<T, B>(s, g, ap, tw, sign, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Length-256 transforms in registers: 256 = 16 x 16.  Sixteen threads own a
// pencil; thread t holds x[t + 16 j] (j = 0..15), runs a 16-point DFT over j
// in registers, multiplies by W256^(t k1), and one exchange through LDS
// (real parts, then imaginary parts) gives thread k1 the values of all t for
// its k1; a second 16-point DFT yields X[k1 + 16 k2].  One LDS round trip
// replaces the three of the generic Stockham passes.
// ---------------------------------------------------------------------------
// Exchange through LDS: thread (p, t) holds V[k1] (k1 = 0..15) of pencil p;
// afterwards thread (p, t) holds W[tt] = V_tt[t] for tt = 0..15.  buf: 4096
// reals, laid out [k1][t][p] with p fastest so stores and loads are 128-byte
// runs.  Ends with a barrier.
template <typename T>
__device__ __forceinline__ void exchange16(typename C2<T>::t v[16], T* buf, int p, int t) {
  T re[16];
#pragma unroll
  for (int k = 0; k < 16; k++) buf[(k * 16 + t) * 16 + p] = v[k].x;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; k++) re[k] = buf[(t * 16 + k) * 16 + p];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; k++) buf[(k * 16 + t) * 16 + p] = v[k].y;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; k++) v[k] = cmake<T>(re[k], buf[(t * 16 + k) * 16 + p]);
  __syncthreads();
}

// y pass of length 256, in place: block = 16 consecutive kz columns of one
// (component, x) plane, 16 threads per column.
template <typename T>
__global__ void __launch_bounds__(256) k_y256(typename C2<T>::t* __restrict__ s, FftGeom g,
                                              const typename C2<T>::t* __restrict__ tw, int sign,
                                              const int* __restrict__ done) {
  using C = typename C2<T>::t;
  if (done && *done) return;
  __shared__ T buf[4096];
  const int runs = (g.nzh + 15) >> 4;
  const int run = (int)(blockIdx.x % runs);
  const long plane = blockIdx.x / runs;  // c * nx + x
  const int p = threadIdx.x & 15, t = threadIdx.x >> 4;
  const int kz = run * 16 + p;
  const bool active = kz < g.nzh;
  C* base = s + plane * (long)g.ny * g.nzp + kz;
  C v[16];
#pragma unroll
  for (int j = 0; j < 16; j++) v[j] = active ? base[(long)(t + 16 * j) * g.nzp] : cmake<T>(0, 0);
  dft16<T>(v, sign);
#pragma unroll
  for (int k = 1; k < 16; k++) {  // W256^(t k), sign as the transform
    C w = tw[t * k];
    if (sign > 0) w.y = -w.y;
    v[k] = cmul(v[k], w);
  }
  exchange16<T>(v, buf, p, t);
  dft16<T>(v, sign);
  if (active) {
#pragma unroll
    for (int k2 = 0; k2 < 16; k2++) base[(long)(t + 16 * k2) * g.nzp] = v[k2];
  }
}

// y pass of length 512, in place: 512 = 16 x 32, with the 32-point stage as
// 16 x 2.  Block = 16 consecutive kz columns of one (component, x) plane, 32
// threads per column (kz fastest: 256-byte runs in FP64).
//  1. thread t holds x[t + 32 j] (j < 16); a 16-point DFT over j and the
//     twiddle W512^(t k1) give Y_t[k1];
//  2. through LDS, thread u = 2 k1 + t0 gathers Y_{2 i + t0}[k1] (i < 16) and
//     takes a 16-point DFT over i: A_t0[m];
//  3. X[k1 + 16 m] = A_0[m] + W32^m A_1[m], X[k1 + 16 (m + 16)] = A_0[m] -
//     W32^m A_1[m], the partner thread (u ^ 1) one shuffle away: thread u
//     ends with X[k1 + 16 (m + 16 t0)] for m < 16.
template <typename T>
__global__ void __launch_bounds__(512) k_y512(typename C2<T>::t* __restrict__ s, FftGeom g,
                                              const typename C2<T>::t* __restrict__ tw, int sign,
                                              const int* __restrict__ done) {
  using C = typename C2<T>::t;
  if (done && *done) return;
  __shared__ T buf[16 * 32 * 16];  // [k1][t][p], one part (real or imaginary) at a time
  const int runs = (g.nzh + 15) >> 4;
  const int run = (int)(blockIdx.x % runs);
  const long plane = blockIdx.x / runs;  // c * nx + x
  const int p = threadIdx.x & 15, t = threadIdx.x >> 4;
  const int kz = run * 16 + p;
  const bool active = kz < g.nzh;
  C* base = s + plane * (long)g.ny * g.nzp + kz;
  C v[16];
#pragma unroll
  for (int j = 0; j < 16; j++) v[j] = active ? base[(long)(t + 32 * j) * g.nzp] : cmake<T>(0, 0);
  dft16<T>(v, sign);
#pragma unroll
  for (int k = 1; k < 16; k++) {  // W512^(t k1)
    C w = tw[t * k];
    if (sign > 0) w.y = -w.y;
    v[k] = cmul(v[k], w);
  }
  // stage 2 gather: thread u = t holds k1 = u >> 1, t0 = u & 1
  const int k1 = t >> 1, t0 = t & 1;
  T re[16];
#pragma unroll
  for (int k = 0; k < 16; k++) buf[(k * 32 + t) * 16 + p] = v[k].x;
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 16; i++) re[i] = buf[(k1 * 32 + 2 * i + t0) * 16 + p];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; k++) buf[(k * 32 + t) * 16 + p] = v[k].y;
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 16; i++) v[i] = cmake<T>(re[i], buf[(k1 * 32 + 2 * i + t0) * 16 + p]);
  dft16<T>(v, sign);  // A_t0[m]
  // stage 3: the partner (t ^ 1) is 16 lanes away in the same wave
#pragma unroll
  for (int m = 0; m < 16; m++) {
    C o;
    o.x = __shfl_xor(v[m].x, 16);
    o.y = __shfl_xor(v[m].y, 16);
    C w = tw[16 * m];  // W32^m = W512^(16 m)
    if (sign > 0) w.y = -w.y;
    // t0 = 0 holds A_0 and receives A_1; t0 = 1 the reverse
    const C a0 = t0 == 0 ? v[m] : o, a1 = cmul(t0 == 0 ? o : v[m], w);
    v[m] = t0 == 0 ? cadd(a0, a1) : csub(a0, a1);
  }
  if (active) {
#pragma unroll
    for (int m = 0; m < 16; m++) base[(long)(k1 + 16 * (m + 16 * t0)) * g.nzp] = v[m];
  }
}

// z passes of length 512 in the y512 form (16 x 16 x 2): 16 pencils per
// block, 32 threads per pencil, thread t the fastest index (256-byte runs
// along z in FP64).  The transforms are z_r2c512_pencil / z_c2r512_pencil.
template <typename T>
__global__ void __launch_bounds__(512) k_z_r2c512(const T* __restrict__ f, typename C2<T>::t* __restrict__ s,
                                                  FftGeom g, const typename C2<T>::t* __restrict__ tw,
                                                  const int* __restrict__ done) {
  using C = typename C2<T>::t;
  if (done && *done) return;
  __shared__ T buf[16 * 512];  // per pencil: the FFT gather, then the unpack
  const int t = threadIdx.x & 31, p = threadIdx.x >> 5;
  const long P = (long)blockIdx.x * 16 + p;
  const long row = P / 3;
  const int pair = (int)(P - row * 3);
  const bool active = row < g.nrows;
  const T* fa = f + (long)(2 * pair) * g.n + row * 512;
  const T* fb = f + (long)(2 * pair + 1) * g.n + row * 512;
  C v[16];
#pragma unroll
  for (int j = 0; j < 16; j++) {
    const T a = active ? fa[t + 32 * j] : (T)0;
    const T b = (active && pair < 2) ? fb[t + 32 * j] : (T)0;
    v[j] = cmake<T>(a, b);
  }
  C* sa = s + (long)(2 * pair) * g.nc + row * g.nzp;
  C* sb = pair < 2 ? s + (long)(2 * pair + 1) * g.nc + row * g.nzp : nullptr;
  z_r2c512_pencil<T>(v, active, tw, buf + p * 512, t, sa, sb);
}

template <typename T>
__global__ void __launch_bounds__(512) k_z_c2r512(const typename C2<T>::t* __restrict__ s, T* __restrict__ f,
                                                  FftGeom g, const typename C2<T>::t* __restrict__ tw, T scale,
                                                  const int* __restrict__ done) {
  using C = typename C2<T>::t;
  if (done && *done) return;
  __shared__ T buf[16 * 512];
  const int t = threadIdx.x & 31, p = threadIdx.x >> 5;
  const long P = (long)blockIdx.x * 16 + p;
  const long row = P / 3;
  const int pair = (int)(P - row * 3);
  const bool active = row < g.nrows;
  const C* sa = s + (long)(2 * pair) * g.nc + row * g.nzp;
  const C* sb = pair < 2 ? s + (long)(2 * pair + 1) * g.nc + row * g.nzp : nullptr;
  C v[16];
  z_c2r512_pencil<T>(sa, sb, active, tw, buf + p * 512, t, v);  // thread t holds z = k1 + 16 (m + 16 t0)
  if (active) {
    const int k1 = t >> 1, t0 = t & 1;
    T* oa = f + (long)(2 * pair) * g.n + row * 512;
    T* ob = f + (long)(2 * pair + 1) * g.n + row * 512;
#pragma unroll
    for (int m = 0; m < 16; m++) {
      const int z = k1 + 16 * m + 256 * t0;
      oa[z] = v[m].x * scale;
      if (pair < 2) ob[z] = v[m].y * scale;
    }
  }
}

// z passes of length 256 in the same 16 x 16 register form.  A pencil is
// one (row, pair) of real components packed as f[2 pair] + i f[2 pair + 1]
// (pair 2 holds component 4 alone); 16 pencils per block, thread t of a
// pencil is the fastest index so a wave reads 128-byte runs along z.
// LDS rows are padded to 257 reals so the transposing reads hit distinct banks.
template <typename T>
__global__ void __launch_bounds__(256) k_z_r2c256(const T* __restrict__ f, typename C2<T>::t* __restrict__ s,
                                                  FftGeom g, const typename C2<T>::t* __restrict__ tw,
                                                  const int* __restrict__ done) {
  using C = typename C2<T>::t;
  if (done && *done) return;
  __shared__ T buf[16 * 257];
  const int t = threadIdx.x & 15, p = threadIdx.x >> 4;
  const long P = (long)blockIdx.x * 16 + p;
  const long row = P / 3;
  const int pair = (int)(P - row * 3);
  const bool active = row < g.nrows;
  const T* fa = f + (long)(2 * pair) * g.n + row * 256;
  const T* fb = f + (long)(2 * pair + 1) * g.n + row * 256;
  C v[16];
#pragma unroll
  for (int j = 0; j < 16; j++) {
    const T a = active ? fa[t + 16 * j] : (T)0;
    const T b = (active && pair < 2) ? fb[t + 16 * j] : (T)0;
    v[j] = cmake<T>(a, b);
  }
  C* sa = s + (long)(2 * pair) * g.nc + row * g.nzp;
  C* sb = pair < 2 ? s + (long)(2 * pair + 1) * g.nc + row * g.nzp : nullptr;
  z_r2c256_pencil<T>(v, active, tw, buf, p, t, sa, sb);
}

template <typename T>
__global__ void __launch_bounds__(256) k_z_c2r256(const typename C2<T>::t* __restrict__ s, T* __restrict__ f,
                                                  FftGeom g, const typename C2<T>::t* __restrict__ tw, T scale,
                                                  const int* __restrict__ done) {
  using C = typename C2<T>::t;
  if (done && *done) return;
  __shared__ T buf[16 * 257];
  const int t = threadIdx.x & 15, p = threadIdx.x >> 4;
  const long P = (long)blockIdx.x * 16 + p;
  const long row = P / 3;
  const int pair = (int)(P - row * 3);
  const bool active = row < g.nrows;
  const C* sa = s + (long)(2 * pair) * g.nc + row * g.nzp;
  const C* sb = pair < 2 ? s + (long)(2 * pair + 1) * g.nc + row * g.nzp : nullptr;
  C v[16];
  z_c2r256_pencil<T>(sa, sb, active, tw, buf, p, t, v);
  if (active) {
    T* fa = f + (long)(2 * pair) * g.n + row * 256;
    T* fb = f + (long)(2 * pair + 1) * g.n + row * 256;
#pragma unroll
    for (int k2 = 0; k2 < 16; k2++) {
      fa[t + 16 * k2] = v[k2].x * scale;
      if (pair < 2) fb[t + 16 * k2] = v[k2].y * scale;
    }
  }
}

// Green operator applied to one frequency: phi <- G_eff(k) phi  (complex 5-vector).
template <typename T>
__device__ __forceinline__ void green_apply(typename C2<T>::t ph[5], int ix, int iy, int iz,
                                            const GreenConst<T>& G) {
  using C = typename C2<T>::t;
  if (ix == 0 && iy == 0 && iz == 0) {  // DC -> 0, MaterialGrid.cpp:733-740
#pragma unroll
    for (int a = 0; a < 5; a++) ph[a] = cmake<T>(0, 0);
    return;
  }
  const int nx = G.nx, ny = G.ny, nz = G.nz;
  const int cx = nx / 2, cy = ny / 2, cz = nz / 2;
  const bool ex = (nx & 1) == 0, ey = (ny & 1) == 0, ez = (nz & 1) == 0;
  const int jx = ix == 0 ? 0 : nx - ix, jy = iy == 0 ? 0 : ny - iy, jz = iz == 0 ? 0 : nz - iz;
  const bool sp = (ex && (ix == cx || ix == cx + 1)) || (ey && (iy == cy || iy == cy + 1)) ||
                  (ez && (iz == cz || iz == cz + 1));
  const bool spn = (ex && (jx == cx || jx == cx + 1)) || (ey && (jy == cy || jy == cy + 1)) ||
                   (ez && (jz == cz || jz == cz + 1));
  C out[5];
#pragma unroll
  for (int a = 0; a < 5; a++) out[a] = cmake<T>(0, 0);
  if (!(sp && spn)) {
    // Gamma(k):  d = V N V^T phi,  V_a = B_a k,  N = K^-1 - K^-1 k k^T K^-1 / (k^T K^-1 k)
    const T k0 = (T)(ix > cx ? ix - nx : ix) * G.kscale[0];
    const T k1 = (T)(iy > cy ? iy - ny : iy) * G.kscale[1];
    const T k2 = (T)(iz > cz ? iz - nz : iz) * G.kscale[2];
    const T mono[6] = {k0 * k0, k1 * k1, k2 * k2, k1 * k2, k0 * k2, k0 * k1};
    T K[6];
#pragma unroll
    for (int r = 0; r < 6; r++) {
      T acc = 0;
#pragma unroll
      for (int q = 0; q < 6; q++) acc = fmad(G.Kc[r][q], mono[q], acc);
      K[r] = acc;
    }
    // K = [[a f e],[f b d],[e d c]]
    const T a = K[0], b = K[1], c = K[2], d = K[3], e = K[4], f = K[5];
    const T A00 = fmad(b, c, -(d * d)), A11 = fmad(a, c, -(e * e)), A22 = fmad(a, b, -(f * f));
    const T A12 = fmad(e, f, -(a * d)), A02 = fmad(f, d, -(b * e)), A01 = fmad(d, e, -(c * f));
    const T det = fmad(e, A02, fmad(f, A01, a * A00));
    const T p0 = fmad(A02, k2, fmad(A01, k1, A00 * k0));
    const T p1 = fmad(A12, k2, fmad(A11, k1, A01 * k0));
    const T p2 = fmad(A22, k2, fmad(A12, k1, A02 * k0));
    const T sk = fmad(k2, p2, fmad(k1, p1, k0 * p0));
    const T inv = (T)1 / (sk * det);
    // N = (adj * sk - p p^T) / (sk det)
    const T N00 = fmad(A00, sk, -(p0 * p0)) * inv, N11 = fmad(A11, sk, -(p1 * p1)) * inv,
            N22 = fmad(A22, sk, -(p2 * p2)) * inv;
    const T N12 = fmad(A12, sk, -(p1 * p2)) * inv, N02 = fmad(A02, sk, -(p0 * p2)) * inv,
            N01 = fmad(A01, sk, -(p0 * p1)) * inv;
    const T s2 = (T)0.70710678118654752440, s6 = (T)0.40824829046386301637;
    // u = V^T phi
    C u0, u1, u2;
    u0.x = -s2 * k0 * ph[0].x - s6 * k0 * ph[1].x + s2 * k2 * ph[3].x + s2 * k1 * ph[4].x;
    u0.y = -s2 * k0 * ph[0].y - s6 * k0 * ph[1].y + s2 * k2 * ph[3].y + s2 * k1 * ph[4].y;
    u1.x = s2 * k1 * ph[0].x - s6 * k1 * ph[1].x + s2 * k2 * ph[2].x + s2 * k0 * ph[4].x;
    u1.y = s2 * k1 * ph[0].y - s6 * k1 * ph[1].y + s2 * k2 * ph[2].y + s2 * k0 * ph[4].y;
    u2.x = (T)2 * s6 * k2 * ph[1].x + s2 * k1 * ph[2].x + s2 * k0 * ph[3].x;
    u2.y = (T)2 * s6 * k2 * ph[1].y + s2 * k1 * ph[2].y + s2 * k0 * ph[3].y;
    C w0, w1, w2;
    w0.x = fmad(N02, u2.x, fmad(N01, u1.x, N00 * u0.x));
    w0.y = fmad(N02, u2.y, fmad(N01, u1.y, N00 * u0.y));
    w1.x = fmad(N12, u2.x, fmad(N11, u1.x, N01 * u0.x));
    w1.y = fmad(N12, u2.y, fmad(N11, u1.y, N01 * u0.y));
    w2.x = fmad(N22, u2.x, fmad(N12, u1.x, N02 * u0.x));
    w2.y = fmad(N22, u2.y, fmad(N12, u1.y, N02 * u0.y));
    const T wt = (sp || spn) ? (T)0.5 : (T)1;
    out[0].x = wt * s2 * (-k0 * w0.x + k1 * w1.x);
    out[0].y = wt * s2 * (-k0 * w0.y + k1 * w1.y);
    out[1].x = wt * s6 * (-k0 * w0.x - k1 * w1.x + (T)2 * k2 * w2.x);
    out[1].y = wt * s6 * (-k0 * w0.y - k1 * w1.y + (T)2 * k2 * w2.y);
    out[2].x = wt * s2 * (k2 * w1.x + k1 * w2.x);
    out[2].y = wt * s2 * (k2 * w1.y + k1 * w2.y);
    out[3].x = wt * s2 * (k2 * w0.x + k0 * w2.x);
    out[3].y = wt * s2 * (k2 * w0.y + k0 * w2.y);
    out[4].x = wt * s2 * (k1 * w0.x + k0 * w1.x);
    out[4].y = wt * s2 * (k1 * w0.y + k0 * w1.y);
  }
  if (sp || spn) {
    // special planes: d = -L_inv phi = L0^-1 phi  (MaterialGrid.cpp:628-632, 713)
    const T wt = (sp && spn) ? (T)1 : (T)0.5;
#pragma unroll
    for (int i = 0; i < 5; i++) {
      T re = 0, im = 0;
#pragma unroll
      for (int j = 0; j < 5; j++) {
        re = fmad(G.L0inv[j * 5 + i], ph[j].x, re);
        im = fmad(G.L0inv[j * 5 + i], ph[j].y, im);
      }
      out[i].x = fmad(wt, re, out[i].x);
      out[i].y = fmad(wt, im, out[i].y);
    }
  }
#pragma unroll
  for (int a = 0; a < 5; a++) ph[a] = out[a];
}

// ---------------------------------------------------------------------------
// x pass with the Green operator: block = run of B consecutive columns
// q = y * nzh + kz, all 5 components; forward FFT along x per component,
// frequencies kept in registers (E per thread), Green applied, inverse FFT.
// ---------------------------------------------------------------------------
template <typename T, int B, int E>
__device__ __forceinline__ void x_green_body(typename C2<T>::t* __restrict__ s, const FftGeom& g, const AxisPlan& ap,
                                             const typename C2<T>::t* __restrict__ tw,
                                             const GreenConst<T>* __restrict__ gc, long blk) {
  using C = typename C2<T>::t;
  extern __shared__ unsigned char smem_raw[];
  C* buf0 = reinterpret_cast<C*>(smem_raw);
  C* buf1 = buf0 + B * g.nx;
  const long ncol = (long)g.ny * g.nzp;  // x stride
  const int runs = (g.nzh + B - 1) / B;
  const int iy = (int)(blk / runs);
  const int kz0 = (int)(blk - (long)iy * runs) * B;
  const long q0 = (long)iy * g.nzp + kz0;
  const int nb = min(B, g.nzh - kz0);
  const int cnt = B * g.nx;
  C reg[5][E];
  // Every load of all 5 components is issued before the first transform, so
  // a block keeps 5 B nx values in flight instead of waiting on one component
  // at a time.  Component loops are unrolled so reg[][] stays in registers.
#pragma unroll
  for (int c = 0; c < 5; c++) {
    const C* base = s + (long)c * g.nc + q0;
#pragma unroll
    for (int i = 0; i < E; i++) {
      const int e = threadIdx.x + i * blockDim.x;
      const int b = e % B, x = e / B;
      reg[c][i] = (e < cnt && b < nb) ? base[(long)x * ncol + b] : cmake<T>(0, 0);
    }
  }
#pragma unroll
  for (int c = 0; c < 5; c++) {
#pragma unroll
    for (int i = 0; i < E; i++) {
      const int e = threadIdx.x + i * blockDim.x;
      if (e < cnt) buf0[e] = reg[c][i];
    }
    __syncthreads();
    C* res = lds_fft<T, B>(buf0, buf1, ap, tw, -1);
#pragma unroll
    for (int i = 0; i < E; i++) {
      const int e = threadIdx.x + i * blockDim.x;
      reg[c][i] = e < cnt ? res[e] : cmake<T>(0, 0);
    }
    __syncthreads();
  }
  const GreenConst<T> G = *as_const_mem(gc);
#pragma unroll
  for (int i = 0; i < E; i++) {
    const int e = threadIdx.x + i * blockDim.x;
    const int b = e % B, x = e / B;
    if (e < cnt && b < nb) {
      const int iz = kz0 + b;
      C ph[5];
#pragma unroll
      for (int c = 0; c < 5; c++) ph[c] = reg[c][i];
      green_apply<T>(ph, x, iy, iz, G);
#pragma unroll
      for (int c = 0; c < 5; c++) reg[c][i] = ph[c];
    }
  }
#pragma unroll
  for (int c = 0; c < 5; c++) {
#pragma unroll
    for (int i = 0; i < E; i++) {
      const int e = threadIdx.x + i * blockDim.x;
      if (e < cnt) buf0[e] = reg[c][i];
    }
    __syncthreads();
    C* res = lds_fft<T, B>(buf0, buf1, ap, tw, +1);
    C* base = s + (long)c * g.nc + q0;
    for (int e = threadIdx.x; e < cnt; e += blockDim.x) {
      const int b = e % B, x = e / B;
      if (b < nb) base[(long)x * ncol + b] = res[e];
    }
    __syncthreads();
  }
}

template <typename T, int B, int E, int NT>
__global__ void __launch_bounds__(NT) k_x_green(typename C2<T>::t* __restrict__ s, FftGeom g, AxisPlan ap,
                                                 const typename C2<T>::t* __restrict__ tw,
                                                 const GreenConst<T>* __restrict__ gc, const int* __restrict__ done) {
  if (done && *done) return;
  x_green_body<T, B, E>(s, g, ap, tw, gc, xcd_remap(blockIdx.x, gridDim.x));
}

// ---------------------------------------------------------------------------
// x pass of length 256 in the 16 x 16 register form, Green operator fused.
// Block = P consecutive kz columns of one y, all 5 components: component c
// is threads [16 P c, 16 P (c + 1)), thread (c, t, p) owns x = t + 16 j of
// pencil p (p fastest, so a load covers 16 rows of P consecutive kz).  The
// five components transform side by side; the Green operator then needs all
// five at one frequency, so the spectra go through LDS, in two halves of the
// frequencies, to the threads that apply it, and back.
// LDS: per component a 16 x 17 x P array for the FFT exchange (rows padded
// to 17 so the transposing reads spread over the banks), reused for the
// 2 x 5 x 128 x P reals of a half of the Green exchange.
// ---------------------------------------------------------------------------
template <int P> __device__ __forceinline__ int xex(int row, int col, int p) { return (row * 17 + col) * P + p; }

// thread (t, p) holds V[k] (k = 0..15) of pencil p; afterwards it holds
// W[k] = V_k[t] (as exchange16).  buf: this component's 16 x 17 x P reals.
template <typename T, int P>
__device__ __forceinline__ void exchange16_x(typename C2<T>::t v[16], T* buf, int t, int p) {
  T re[16];
#pragma unroll
  for (int k = 0; k < 16; k++) buf[xex<P>(k, t, p)] = v[k].x;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; k++) re[k] = buf[xex<P>(t, k, p)];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; k++) buf[xex<P>(k, t, p)] = v[k].y;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; k++) v[k] = cmake<T>(re[k], buf[xex<P>(t, k, p)]);
  __syncthreads();
}

// 256-point transform of the pencil whose x = t + 16 j thread t holds in
// v[j]; on return v[k2] holds frequency t + 16 k2 (natural order).
template <typename T, int P>
__device__ __forceinline__ void fft256_x(typename C2<T>::t v[16], T* buf, const typename C2<T>::t* __restrict__ tw,
                                         int t, int p, int sign) {
  using C = typename C2<T>::t;
  dft16<T>(v, sign);
#pragma unroll
  for (int k = 1; k < 16; k++) {  // W256^(t k), sign as the transform
    C w = tw[t * k];
    if (sign > 0) w.y = -w.y;
    v[k] = cmul(v[k], w);
  }
  exchange16_x<T, P>(v, buf, t, p);
  dft16<T>(v, sign);
}

// Four waves per SIMD: three blocks per CU.  In FP64 that spills 44 B per
// lane (FFT values of the second half through the first half's Green
// phase) and still runs faster than three waves (two blocks per CU) without
// spills (256^3: 0.53 against 0.61 ms).
template <typename T, int P>
__global__ void __launch_bounds__(80 * P) __attribute__((amdgpu_waves_per_eu(4))) k_x256_green(typename C2<T>::t* __restrict__ s, FftGeom g,
                                                        const typename C2<T>::t* __restrict__ tw,
                                                        const GreenConst<T>* __restrict__ gc,
                                                        const int* __restrict__ done) {
  using C = typename C2<T>::t;
  constexpr int NT = 80 * P;  // threads: 5 components x 16 P
  if (done && *done) return;
  __shared__ T buf[5 * 16 * 17 * P];
  const int runs = (g.nzh + P - 1) / P;
  const long blk = xcd_remap(blockIdx.x, gridDim.x);
  const int iy = (int)(blk / runs);
  const int kz0 = (int)(blk - (long)iy * runs) * P;
  const int tid = threadIdx.x;
  const int c = __builtin_amdgcn_readfirstlane(tid / (16 * P));  // whole waves per component
  const int r = tid - c * 16 * P;
  const int p = r % P, t = r / P;
  const bool active = kz0 + p < g.nzh;
  // a uniform base and 32-bit element offsets (x <= 255 rows of ny nzp), so
  // the addresses need one register each instead of a 64-bit pair
  const unsigned ncol = (unsigned)(g.ny * g.nzp);  // x stride
  C* base = s + (long)c * g.nc + (long)iy * g.nzp + kz0;
  T* cbuf = buf + c * (16 * 17 * P);
  C v[16];
#pragma unroll
  for (int j = 0; j < 16; j++) v[j] = active ? base[(unsigned)(t + 16 * j) * ncol + (unsigned)p] : cmake<T>(0, 0);
  fft256_x<T, P>(v, cbuf, tw, t, p, -1);
  // v[k2] = X_c(f = t + 16 k2) of pencil p.  The Green operator runs in two
  // halves of the frequencies (f < 128 and f >= 128): each half of all five
  // spectra, real and imaginary parts, fits in the same LDS, and the threads
  // apply the operator in place one point at a time.  In half h the point
  // q = (f - 128 h) P + pp holds component cc at buf[cc 128 P + q] (real)
  // and buf[640 P + cc 128 P + q] (imaginary).
  constexpr int NPT = (128 * P + NT - 1) / NT;  // Green points per thread per half
  const GreenConst<T>* G = as_const_mem(gc);
#pragma unroll
  for (int h = 0; h < 2; h++) {
#pragma unroll
    for (int kk = 0; kk < 8; kk++) {
      const int i = (c * 128 + t + 16 * kk) * P + p;
      buf[i] = v[8 * h + kk].x;
      buf[640 * P + i] = v[8 * h + kk].y;
    }
    __syncthreads();
#pragma unroll 1
    for (int m = 0; m < NPT; m++) {
      const int q = tid + NT * m;
      const int iz = kz0 + q % P;
      if (q < 128 * P && iz < g.nzh) {
        C ph[5];
#pragma unroll
        for (int cc = 0; cc < 5; cc++) ph[cc] = cmake<T>(buf[cc * 128 * P + q], buf[640 * P + cc * 128 * P + q]);
        green_apply<T>(ph, 128 * h + q / P, iy, iz, *G);
#pragma unroll
        for (int cc = 0; cc < 5; cc++) {
          buf[cc * 128 * P + q] = ph[cc].x;
          buf[640 * P + cc * 128 * P + q] = ph[cc].y;
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 8; kk++) {
      const int i = (c * 128 + t + 16 * kk) * P + p;
      v[8 * h + kk] = cmake<T>(buf[i], buf[640 * P + i]);
    }
    __syncthreads();
  }
  // inverse: v[j] holds frequency t + 16 j, the layout the transform takes.
  // The twiddles are loaded again (an opaque copy of the pointer) rather
  // than held in registers from the forward transform.
  const C* tw2 = tw;
  __asm__ volatile("" : "+s"(tw2));
  fft256_x<T, P>(v, cbuf, tw2, t, p, +1);
  // the thread's pencil coordinates afresh for the stores, rather than kept
  // live (in scratch) through the transforms and the Green phase
  int r2 = threadIdx.x - c * 16 * P;
  __asm__ volatile("" : "+v"(r2));
  const int p2 = r2 % P, t2 = r2 / P;
  if (kz0 + p2 < g.nzh) {
#pragma unroll
    for (int k2 = 0; k2 < 16; k2++) base[(unsigned)(t2 + 16 * k2) * ncol + (unsigned)p2] = v[k2];
  }
}

// ---------------------------------------------------------------------------
// Batched passes: one launch covers the same pass of many transform jobs
// (independent grids of any size).  Blocks [off[m], off[m+1]) of the launch
// belong to job ids[m]; a job whose loop flag is set returns at entry.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int group_member(const FftGroup& grp, long b, long* local) {
  int lo = 0, hi = grp.count;  // off[lo] <= b < off[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (uni(grp.off[mid]) <= b) lo = mid;
    else hi = mid;
  }
  *local = b - uni(grp.off[lo]);
  return uni(grp.ids[lo]);
}

// the fields lds_fft reads (the radix array itself is not copied: an indexed
// local array would live in scratch)
__device__ __forceinline__ AxisPlan uni_plan(const AxisPlan& a) {
  AxisPlan r;
  r.n = uni(a.n);
  r.nstages = uni(a.nstages);
  r.code = (unsigned long long)uni((long)a.code);
  return r;
}

__device__ __forceinline__ FftGeom uni(const FftGeom& g) {
  FftGeom r;
  r.n = uni(g.n);
  r.nc = uni(g.nc);
  r.nrows = uni(g.nrows);
  r.nx = uni(g.nx);
  r.ny = uni(g.ny);
  r.nz = uni(g.nz);
  r.nzh = uni(g.nzh);
  r.nzp = uni(g.nzp);
  return r;
}

template <typename T, int B>
__global__ void __launch_bounds__(256) k_z_r2c_b(FftGroup grp) {
  using C = typename C2<T>::t;
  long local;
  const FftJob* J = grp.jobs + group_member(grp, blockIdx.x, &local);
  const int* done = guni(J->done);
  if (done && *done) return;
  const FftGeom g = uni(J->g);
  const AxisPlan ap = uni_plan(J->az);
  z_r2c_body<T, B>(guni((const T*)J->field), guni((C*)J->spec), g, ap, guni((const C*)J->tw[2]), local);
}
template <typename T, int B>
__global__ void __launch_bounds__(256) k_z_c2r_b(FftGroup grp) {
  using C = typename C2<T>::t;
  long local;
  const FftJob* J = grp.jobs + group_member(grp, blockIdx.x, &local);
  const int* done = guni(J->done);
  if (done && *done) return;
  const FftGeom g = uni(J->g);
  const AxisPlan ap = uni_plan(J->az);
  z_c2r_body<T, B>(guni((const C*)J->spec), guni((T*)J->dtil), g, ap, guni((const C*)J->tw[2]),
                   (T)(1.0 / (double)g.n), local);
}
template <typename T, int B>
__global__ void __launch_bounds__(256) k_y_b(FftGroup grp, int sign) {
  using C = typename C2<T>::t;
  long local;
  const FftJob* J = grp.jobs + group_member(grp, blockIdx.x, &local);
  const int* done = guni(J->done);
  if (done && *done) return;
  const FftGeom g = uni(J->g);
  const AxisPlan ap = uni_plan(J->ay);
  y_body<T, B>(guni((C*)J->spec), g, ap, guni((const C*)J->tw[1]), sign, local);
}
template <typename T, int B, int E, int NT>
__global__ void __launch_bounds__(NT) k_x_green_b(FftGroup grp) {
  using C = typename C2<T>::t;
  long local;
  const FftJob* J = grp.jobs + group_member(grp, xcd_remap(blockIdx.x, gridDim.x), &local);
  const int* done = guni(J->done);
  if (done && *done) return;
  const FftGeom g = uni(J->g);
  const AxisPlan ap = uni_plan(J->ax);
  x_green_body<T, B, E>(guni((C*)J->spec), g, ap, guni((const C*)J->tw[0]), uni_as<4>((const GreenConst<T>*)J->gc), local);
}

// ---------------------------------------------------------------------------
// Register form of a pass for short axes: one thread per pencil, the whole
// pencil in registers, a mixed-radix decimation-in-time transform unrolled
// at compile time (N = R M: M-point transforms of the R interleaved
// subsequences, twiddles W_N^(r k1), R-point transforms across them).  No
// LDS and no barrier; the lanes of a wave take consecutive kz, so each of the
// pencil's loads and stores is one contiguous run of the wave's 64 values.
// Twiddles come from the axis's table (W_n^j) at compile-time offsets.
// ---------------------------------------------------------------------------
template <int N>
constexpr int reg_radix() {
  return N == 16 ? 4
         : (N % 8 == 0 && N > 8) ? 8
         : (N % 4 == 0 && N > 4) ? 4
         : (N % 2 == 0 && N > 2) ? 2
         : (N % 3 == 0 && N > 3) ? 3
         : (N % 5 == 0 && N > 5) ? 5
         : (N % 7 == 0 && N > 7) ? 7 : N;
}

// W_N^j = exp(-2 pi i j / N) evaluated at compile time: the angle reduced by
// exact integer arithmetic to [0, pi/4] (quadrant, then the complement),
// then Taylor series to x^23 (error below 1e-20 there).
struct CxCS {
  double c, s;
};
constexpr double cx_sin(double x) {
  double t = x, r = x;
  for (int i = 1; i <= 11; i++) {
    t *= -x * x / (double)((2 * i) * (2 * i + 1));
    r += t;
  }
  return r;
}
constexpr double cx_cos(double x) {
  double t = 1.0, r = 1.0;
  for (int i = 1; i <= 11; i++) {
    t *= -x * x / (double)((2 * i - 1) * (2 * i));
    r += t;
  }
  return r;
}
constexpr CxCS cx_twiddle(int j, int n) {  // cos and sin of 2 pi j / n
  j %= n;
  const int q = (4 * j) / n;  // quadrant
  const int rr = 4 * j - q * n;  // angle within it: (rr / n) pi / 2
  const double hp = 1.57079632679489661923;
  double c0, s0;
  if (2 * rr <= n) {
    const double a = (double)rr / (double)n * hp;
    c0 = cx_cos(a);
    s0 = cx_sin(a);
  } else {
    const double a = (double)(n - rr) / (double)n * hp;
    c0 = cx_sin(a);
    s0 = cx_cos(a);
  }
  switch (q) {
    case 0: return CxCS{c0, s0};
    case 1: return CxCS{-s0, c0};
    case 2: return CxCS{-c0, -s0};
    default: return CxCS{s0, -c0};
  }
}
template <int X, int N>
struct TwC {  // W_N^X (forward sign: exp(-2 pi i X / N))
  static constexpr CxCS v = cx_twiddle(X, N);
};

// compile-time loop: f(std::integral_constant<int, I>) for I in [B, E)
template <int B, int E, typename F>
__device__ __forceinline__ void sfor(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    sfor<B + 1, E>(f);
  }
}

// transform of v (natural order in and out), sign -1 forward, +1 inverse
template <typename T, int N>
__device__ __forceinline__ void reg_fft(typename C2<T>::t (&v)[N], int sign) {
  using C = typename C2<T>::t;
  constexpr int R = reg_radix<N>();
  if constexpr (R == N) {
    dft<T, N>(v, sign);
  } else {
    constexpr int M = N / R;
    C y[R][M];
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
      for (int m = 0; m < M; m++) y[r][m] = v[r + R * m];
#pragma unroll
    for (int r = 0; r < R; r++) reg_fft<T, M>(y[r], sign);
    sfor<0, M>([&](auto k1c) {
      constexpr int k1 = decltype(k1c)::value;
      C t[R];
      sfor<0, R>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        if constexpr (r * k1 == 0) {
          t[r] = y[r][k1];
        } else {
          constexpr CxCS w = TwC<(r * k1) % N, N>::v;
          t[r] = cmul(y[r][k1], cmake<T>((T)w.c, (T)(sign > 0 ? w.s : -w.s)));
        }
      });
      dft<T, R>(t, sign);
#pragma unroll
      for (int k2 = 0; k2 < R; k2++) v[k1 + M * k2] = t[k2];
    });
  }
}

// Batched y pass in register form, in place: pencil p of a job (kz fastest,
// then x, then the component) is thread p of the job's blocks.
template <typename T, int N>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_y_reg_b(FftGroup grp, int sign) {
  using C = typename C2<T>::t;
  long local;
  const FftJob* J = grp.jobs + group_member(grp, blockIdx.x, &local);
  const int* done = guni(J->done);
  if (done && *done) return;
  const FftGeom g = uni(J->g);
  const long p = local * 256 + threadIdx.x;
  if (p >= 5L * g.nx * g.nzh) return;  // no barrier below
  const int kz = (int)(p % g.nzh);
  const long plane = p / g.nzh;  // c * nx + x
  C* base = guni((C*)J->spec) + plane * (long)g.ny * g.nzp + kz;
  C v[N];
#pragma unroll
  for (int y = 0; y < N; y++) v[y] = base[(long)y * g.nzp];
  reg_fft<T, N>(v, sign);
#pragma unroll
  for (int y = 0; y < N; y++) base[(long)y * g.nzp] = v[y];
}

// ---------------------------------------------------------------------------
// Host-side launchers
// ---------------------------------------------------------------------------
bool make_axis_plan(int n, AxisPlan* ap) {
  const int order[6] = {8, 4, 2, 3, 5, 7};
  return make_axis_plan_ordered(n, order, ap);
}

bool make_axis_plan_ordered(int n, const int order[6], AxisPlan* ap) {
  ap->n = n;
  ap->nstages = 0;
  int m = n;
  for (int oi = 0; oi < 6; oi++) {
    const int r = order[oi];
    while (m % r == 0 && m > 1) {
      if (ap->nstages >= 12) return false;
      ap->radix[ap->nstages++] = r;
      m /= r;
    }
  }
  ap->code = 0;
  for (int s = 0; s < ap->nstages; s++) ap->code |= (unsigned long long)ap->radix[s] << (4 * s);
  return m == 1 && n > 1;
}

static int pick_rows(int nz, int budget) {
  int b = 1;
  while (b < 16 && 3 * (2 * b) * nz <= budget) b *= 2;
  return b;
}
static int pick_run(int n, int budget) {
  int b = 1;
  while (b < 64 && (2 * b) * n <= budget) b *= 2;
  return b;
}

template <typename T>
static hipError_t launch_z_r2c_t(const T* f, void* s, const FftGeom& g, const AxisPlan& ap, const void* tw, int Bq,
                                 const int* done, hipStream_t st) {
  using C = typename C2<T>::t;
  if (g.nz == 256 && Bq == 0) {  // the register 16 x 16 form
    const unsigned grid = (unsigned)((3 * g.nrows + 15) / 16);
    hipLaunchKernelGGL((k_z_r2c256<T>), dim3(grid), dim3(256), 0, st, f, (C*)s, g, (const C*)tw, done);
    return hipGetLastError();
  }
  if (g.nz == 512 && Bq == 0) {  // the register 16 x 16 x 2 form
    const unsigned grid = (unsigned)((3 * g.nrows + 15) / 16);
    hipLaunchKernelGGL((k_z_r2c512<T>), dim3(grid), dim3(512), 0, st, f, (C*)s, g, (const C*)tw, done);
    return hipGetLastError();
  }
  const PassShape ps = pass_shape(sizeof(T) == 8 ? 0 : 1, kPassZ, g, Bq);
  if (!ps.ok) return hipErrorInvalidValue;
  const unsigned grid = (unsigned)ps.blocks;
  switch (ps.B) {
#define CASE(BB) case BB: hipLaunchKernelGGL((k_z_r2c<T, BB>), dim3(grid), dim3(256), ps.lds, st, f, (C*)s, g, ap, (const C*)tw, done); break;
    CASE(1) CASE(2) CASE(4) CASE(8) CASE(16)
#undef CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T>
static hipError_t launch_z_c2r_t(const void* s, T* f, const FftGeom& g, const AxisPlan& ap, const void* tw,
                                 double scale, int Bq, const int* done, hipStream_t st) {
  using C = typename C2<T>::t;
  if (g.nz == 256 && Bq == 0) {  // the register 16 x 16 form
    const unsigned grid = (unsigned)((3 * g.nrows + 15) / 16);
    hipLaunchKernelGGL((k_z_c2r256<T>), dim3(grid), dim3(256), 0, st, (const C*)s, f, g, (const C*)tw, (T)scale,
                       done);
    return hipGetLastError();
  }
  if (g.nz == 512 && Bq == 0) {  // the register 16 x 16 x 2 form
    const unsigned grid = (unsigned)((3 * g.nrows + 15) / 16);
    hipLaunchKernelGGL((k_z_c2r512<T>), dim3(grid), dim3(512), 0, st, (const C*)s, f, g, (const C*)tw, (T)scale,
                       done);
    return hipGetLastError();
  }
  const PassShape ps = pass_shape(sizeof(T) == 8 ? 0 : 1, kPassZ, g, Bq);
  if (!ps.ok) return hipErrorInvalidValue;
  const unsigned grid = (unsigned)ps.blocks;
  switch (ps.B) {
#define CASE(BB) case BB: hipLaunchKernelGGL((k_z_c2r<T, BB>), dim3(grid), dim3(256), ps.lds, st, (const C*)s, f, g, ap, (const C*)tw, (T)scale, done); break;
    CASE(1) CASE(2) CASE(4) CASE(8) CASE(16)
#undef CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T>
static hipError_t launch_y_t(void* s, const FftGeom& g, const AxisPlan& ap, const void* tw, int sign, int Bq,
                             const int* done, hipStream_t st) {
  using C = typename C2<T>::t;
  if (g.ny == 256 && Bq == 0) {  // the register 16 x 16 form
    const unsigned grid = (unsigned)(5L * g.nx * ((g.nzh + 15) / 16));
    hipLaunchKernelGGL((k_y256<T>), dim3(grid), dim3(256), 0, st, (C*)s, g, (const C*)tw, sign, done);
    return hipGetLastError();
  }
  if (g.ny == 512 && Bq == 0) {  // the register 16 x 16 x 2 form
    const unsigned grid = (unsigned)(5L * g.nx * ((g.nzh + 15) / 16));
    hipLaunchKernelGGL((k_y512<T>), dim3(grid), dim3(512), 0, st, (C*)s, g, (const C*)tw, sign, done);
    return hipGetLastError();
  }
  const PassShape ps = pass_shape(sizeof(T) == 8 ? 0 : 1, kPassY, g, Bq);
  if (!ps.ok) return hipErrorInvalidValue;
  const unsigned grid = (unsigned)ps.blocks;
  switch (ps.B) {
#define CASE(BB) case BB: hipLaunchKernelGGL((k_y<T, BB>), dim3(grid), dim3(256), ps.lds, st, (C*)s, g, ap, (const C*)tw, sign, done); break;
    CASE(1) CASE(2) CASE(4) CASE(8) CASE(16) CASE(32) CASE(64)
#undef CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T>
static hipError_t launch_x_green_t(void* s, const FftGeom& g, const AxisPlan& ap, const void* tw,
                                   const void* gc, int Bq, const int* done, hipStream_t st) {
  using C = typename C2<T>::t;
  if (g.nx == 256 && Bq == 0) {  // the register 16 x 16 form: 64-byte runs of kz
    constexpr int P = sizeof(T) == 8 ? 4 : 8;
    const unsigned grid = (unsigned)((long)g.ny * ((g.nzh + P - 1) / P));
    hipLaunchKernelGGL((k_x256_green<T, P>), dim3(grid), dim3(80 * P), 0, st, (C*)s, g, (const C*)tw,
                       (const GreenConst<T>*)gc, done);
    return hipGetLastError();
  }
  const PassShape ps = pass_shape(sizeof(T) == 8 ? 0 : 1, kPassX, g, Bq);
  if (!ps.ok) return hipErrorInvalidValue;
  const unsigned grid = (unsigned)ps.blocks;
  const int E = ps.E;
  const GreenConst<T>* G = (const GreenConst<T>*)gc;
#define LAUNCH(BB, EE, TT) hipLaunchKernelGGL((k_x_green<T, BB, EE, TT>), dim3(grid), dim3(TT), ps.lds, st, (C*)s, g, ap, (const C*)tw, G, done)
#define EC(BB, TT)                              \
  case BB:                                      \
    if (E == 1) LAUNCH(BB, 1, TT);              \
    else if (E == 2) LAUNCH(BB, 2, TT);         \
    else if (E == 4) LAUNCH(BB, 4, TT);         \
    else LAUNCH(BB, 8, TT);                     \
    break;
  switch (ps.B) {
    EC(4, 256) EC(8, 512) EC(16, 512) EC(32, 512) EC(64, 512)
    default: return hipErrorInvalidValue;
  }
#undef EC
#undef LAUNCH
  return hipGetLastError();
}

hipError_t launch_z_r2c(int prec, const void* f, void* s, const FftGeom& g, const AxisPlan& ap, const void* tw, int B,
                        const int* done, hipStream_t st) {
  return prec == 0 ? launch_z_r2c_t<double>((const double*)f, s, g, ap, tw, B, done, st)
                   : launch_z_r2c_t<float>((const float*)f, s, g, ap, tw, B, done, st);
}
hipError_t launch_z_c2r(int prec, const void* s, void* f, const FftGeom& g, const AxisPlan& ap, const void* tw,
                        double scale, int B, const int* done, hipStream_t st) {
  return prec == 0 ? launch_z_c2r_t<double>(s, (double*)f, g, ap, tw, scale, B, done, st)
                   : launch_z_c2r_t<float>(s, (float*)f, g, ap, tw, scale, B, done, st);
}
hipError_t launch_y(int prec, void* s, const FftGeom& g, const AxisPlan& ap, const void* tw, int sign, int B,
                    const int* done, hipStream_t st) {
  return prec == 0 ? launch_y_t<double>(s, g, ap, tw, sign, B, done, st)
                   : launch_y_t<float>(s, g, ap, tw, sign, B, done, st);
}
hipError_t launch_x_green(int prec, void* s, const FftGeom& g, const AxisPlan& ap, const void* tw, const void* gc,
                          int B, const int* done, hipStream_t st) {
  return prec == 0 ? launch_x_green_t<double>(s, g, ap, tw, gc, B, done, st)
                   : launch_x_green_t<float>(s, g, ap, tw, gc, B, done, st);
}

// ---- batched launches ------------------------------------------------------
static constexpr size_t kMaxLds = 65536;

// threads of an x-pass block: 256 for runs of up to 4 columns, else 512
static int x_threads(int b) { return b <= 4 ? 256 : 512; }

int pass_candidates(FftPass pass, int* out, int cap) {
  static const int z[] = {1, 2, 4, 8, 16};
  static const int y[] = {1, 2, 4, 8, 16, 32, 64};
  static const int x[] = {4, 8, 16, 32, 64};
  const int* src = pass == kPassZ ? z : (pass == kPassY ? y : x);
  const int cnt = pass == kPassZ ? 5 : (pass == kPassY ? 7 : 5);
  int k = 0;
  for (; k < cnt && k < cap; k++) out[k] = src[k];
  return k;
}

// the radices reg_fft applies for length n, first stage first (host side of
// reg_radix: the outermost split runs last)
int reg_form_radices(int n, int* out) {
  int chain[12], k = 0;
  while (n > 1 && k < 12) {
    const int r = n == 16 ? 4 : (n % 8 == 0 && n > 8) ? 8 : (n % 4 == 0 && n > 4) ? 4 : (n % 2 == 0 && n > 2) ? 2
                  : (n % 3 == 0 && n > 3) ? 3 : (n % 5 == 0 && n > 5) ? 5 : (n % 7 == 0 && n > 7) ? 7 : n;
    chain[k++] = r;
    n /= r;
  }
  for (int i = 0; i < k; i++) out[i] = chain[k - 1 - i];
  return k;
}

bool reg_form_supported(int prec, FftPass pass, int n) {
  if (pass != kPassY) return false;
  switch (n) {  // at two waves per SIMD without spilling (FP32 64 points: 116 B)
    case 16: case 24: case 32: return true;
    case 48: case 64: return prec != 0;
    default: return false;
  }
}

PassShape pass_shape(int prec, FftPass pass, const FftGeom& g, int B) {
  const size_t csz = prec == 0 ? 16 : 8;
  PassShape ps{0, 1, 0, 0, false};
  if (B == kRegForm || B < 0) {  // one thread per pencil, B = -(axis length)
    const int n = pass == kPassZ ? g.nz : (pass == kPassY ? g.ny : g.nx);
    if (!reg_form_supported(prec, pass, n) || (B < 0 && B != kRegForm && B != -n)) return ps;
    ps.B = -n;
    ps.blocks = (5L * g.nx * g.nzh + 255) / 256;
    ps.lds = 0;
    ps.ok = true;
    return ps;
  }
  if (pass == kPassZ) {
    ps.B = B > 0 ? B : pick_rows(g.nz, prec == 0 ? 768 : 1536);
    ps.blocks = (g.nrows + ps.B - 1) / ps.B;
    ps.lds = 2 * csz * 3 * ps.B * g.nz;
  } else if (pass == kPassY) {
    ps.B = B > 0 ? B : pick_run(g.ny, prec == 0 ? 1024 : 2048);
    ps.blocks = 5L * g.nx * ((g.nzh + ps.B - 1) / ps.B);
    ps.lds = 2 * csz * ps.B * g.ny;
  } else {
    int b = B;
    if (b <= 0) {
      // runs of at least 64 bytes along kz; measured at 256^3: 4 columns of
      // complex doubles in 256-thread blocks, 8 of complex floats in 512
      const int minrun = prec == 0 ? 4 : 8;
      b = pick_run(g.nx, prec == 0 ? 1024 : 2048);
      if (b < minrun) b = minrun;
    }
    ps.B = b;
    ps.blocks = (long)g.ny * ((g.nzh + b - 1) / b);
    ps.lds = 2 * csz * b * g.nx;
    const int nt = x_threads(b);
    ps.E = (b * g.nx + nt - 1) / nt;
    if (ps.E > 8) return ps;  // register budget of the x pass
    ps.E = ps.E <= 1 ? 1 : (ps.E <= 2 ? 2 : (ps.E <= 4 ? 4 : 8));
  }
  ps.ok = ps.lds <= kMaxLds;
  return ps;
}

template <typename T>
static hipError_t z_r2c_group_t(int B, const FftGroup& grp, long blocks, size_t lds, hipStream_t st) {
  switch (B) {
#define CASE(BB) case BB: hipLaunchKernelGGL((k_z_r2c_b<T, BB>), dim3((unsigned)blocks), dim3(256), lds, st, grp); break;
    CASE(1) CASE(2) CASE(4) CASE(8) CASE(16)
#undef CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
template <typename T>
static hipError_t z_c2r_group_t(int B, const FftGroup& grp, long blocks, size_t lds, hipStream_t st) {
  switch (B) {
#define CASE(BB) case BB: hipLaunchKernelGGL((k_z_c2r_b<T, BB>), dim3((unsigned)blocks), dim3(256), lds, st, grp); break;
    CASE(1) CASE(2) CASE(4) CASE(8) CASE(16)
#undef CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
template <typename T>
static hipError_t y_group_t(int B, const FftGroup& grp, long blocks, size_t lds, int sign, hipStream_t st) {
  switch (B) {
#define RCASE(NN) case -NN: hipLaunchKernelGGL((k_y_reg_b<T, NN>), dim3((unsigned)blocks), dim3(256), 0, st, grp, sign); break;
    RCASE(16) RCASE(24) RCASE(32)
#undef RCASE
    case -48: case -64:
      if constexpr (sizeof(T) == 4) {
        if (B == -48) hipLaunchKernelGGL((k_y_reg_b<T, 48>), dim3((unsigned)blocks), dim3(256), 0, st, grp, sign);
        else hipLaunchKernelGGL((k_y_reg_b<T, 64>), dim3((unsigned)blocks), dim3(256), 0, st, grp, sign);
        break;
      } else {
        return hipErrorInvalidValue;
      }
#define CASE(BB) case BB: hipLaunchKernelGGL((k_y_b<T, BB>), dim3((unsigned)blocks), dim3(256), lds, st, grp, sign); break;
    CASE(1) CASE(2) CASE(4) CASE(8) CASE(16) CASE(32) CASE(64)
#undef CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
template <typename T>
static hipError_t x_group_t(int B, int E, const FftGroup& grp, long blocks, size_t lds, hipStream_t st) {
#define LAUNCH(BB, EE, TT) hipLaunchKernelGGL((k_x_green_b<T, BB, EE, TT>), dim3((unsigned)blocks), dim3(TT), lds, st, grp)
#define EC(BB, TT)                              \
  case BB:                                      \
    if (E == 1) LAUNCH(BB, 1, TT);              \
    else if (E == 2) LAUNCH(BB, 2, TT);         \
    else if (E == 4) LAUNCH(BB, 4, TT);         \
    else if (E == 8) LAUNCH(BB, 8, TT);         \
    else return hipErrorInvalidValue;           \
    break;
  switch (B) {
    EC(4, 256) EC(8, 512) EC(16, 512) EC(32, 512) EC(64, 512)
    default: return hipErrorInvalidValue;
  }
#undef EC
#undef LAUNCH
  return hipGetLastError();
}

hipError_t launch_z_r2c_group(int prec, int B, const FftGroup& grp, long blocks, size_t lds, hipStream_t st) {
  return prec == 0 ? z_r2c_group_t<double>(B, grp, blocks, lds, st) : z_r2c_group_t<float>(B, grp, blocks, lds, st);
}
hipError_t launch_z_c2r_group(int prec, int B, const FftGroup& grp, long blocks, size_t lds, hipStream_t st) {
  return prec == 0 ? z_c2r_group_t<double>(B, grp, blocks, lds, st) : z_c2r_group_t<float>(B, grp, blocks, lds, st);
}
hipError_t launch_y_group(int prec, int B, const FftGroup& grp, long blocks, size_t lds, int sign, hipStream_t st) {
  return prec == 0 ? y_group_t<double>(B, grp, blocks, lds, sign, st)
                   : y_group_t<float>(B, grp, blocks, lds, sign, st);
}
hipError_t launch_x_green_group(int prec, int B, int E, const FftGroup& grp, long blocks, size_t lds,
                                hipStream_t st) {
  return prec == 0 ? x_group_t<double>(B, E, grp, blocks, lds, st) : x_group_t<float>(B, E, grp, blocks, lds, st);
}

}  // namespace vpfft
