// fft_regs.h -- the complex helpers and the 256- and 512-point register
// transforms of the z passes, shared by fft_green.hip (the z passes) and voxel_kernels.hip
// (the fused inverse-z / Newton / forward-z kernel), so both compile the same
// operations.  Every function body turns multiply-add contraction off: the
// fused multiply-adds are written out (fmad), so a transform's bits do not
// depend on the translation unit's contraction setting.
#pragma once
#include <hip/hip_runtime.h>

namespace vpfft {

template <typename T> struct C2;
template <> struct C2<double> { using t = double2; };
template <> struct C2<float> { using t = float2; };

template <typename T> __device__ __forceinline__ typename C2<T>::t cmake(T a, T b) {
  typename C2<T>::t r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename C> __device__ __forceinline__ C cadd(C a, C b) {
#pragma clang fp contract(off)
  C r;
  r.x = a.x + b.x;
  r.y = a.y + b.y;
  return r;
}
template <typename C> __device__ __forceinline__ C csub(C a, C b) {
#pragma clang fp contract(off)
  C r;
  r.x = a.x - b.x;
  r.y = a.y - b.y;
  return r;
}
__device__ __forceinline__ double fmad(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmad(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
#pragma clang fp contract(off)
  C r;
  r.x = fmad(a.x, b.x, -(a.y * b.y));
  r.y = fmad(a.x, b.y, a.y * b.x);
  return r;
}
// multiply by (s * i), s = +-1
template <typename C> __device__ __forceinline__ C cmul_si(C a, int s) {
  C r;
  r.x = s > 0 ? -a.y : a.y;
  r.y = s > 0 ? a.x : -a.x;
  return r;
}

// 4-point DFT in place, W = exp(sign 2 pi i / 4)
template <typename T>
__device__ __forceinline__ void dft4(typename C2<T>::t* v, int sign) {
  using C = typename C2<T>::t;
  C a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
  C c = cadd(v[1], v[3]), d = cmul_si(csub(v[1], v[3]), sign);
  v[0] = cadd(a, c);
  v[2] = csub(a, c);
  v[1] = cadd(b, d);
  v[3] = csub(b, d);
}

// ---------------------------------------------------------------------------
// 256 = 16 x 16 in registers: sixteen threads own a pencil; thread t holds
// x[t + 16 j] (j = 0..15), takes a 16-point DFT over j, multiplies by
// W256^(t k1), and one exchange through LDS (real parts, then imaginary parts)
// gives thread k1 the values of all t for its k1; a second 16-point DFT yields
// X[k1 + 16 k2].  One LDS round trip replaces the three of a Stockham FFT.
// ---------------------------------------------------------------------------
// 16-point DFT of v in place (natural order), W = exp(sign 2 pi i / 16):
// 4 DFT-4 over m of v[a + 4m], twiddles W^(a b), 4 DFT-4 over a.
template <typename T>
__device__ __forceinline__ void dft16(typename C2<T>::t v[16], int sign) {
  using C = typename C2<T>::t;
  const T c1 = (T)0.92387953251128675613, s1 = (T)0.38268343236508977173, h = (T)0.70710678118654752440;
  C u[4][4];
#pragma unroll
  for (int a = 0; a < 4; a++) {
    C w[4] = {v[a], v[a + 4], v[a + 8], v[a + 12]};
    dft4<T>(w, sign);
#pragma unroll
    for (int b = 0; b < 4; b++) u[a][b] = w[b];
  }
  // W16^(ab) for ab = 1, 2, 3, 4, 6, 9 (a, b in 1..3)
  const T sg = (T)sign;
  const C w1 = cmake<T>(c1, sg * s1), w2 = cmake<T>(h, sg * h), w3 = cmake<T>(s1, sg * c1);
  const C w6 = cmake<T>(-h, sg * h), w9 = cmake<T>(-c1, -sg * s1);
  u[1][1] = cmul(u[1][1], w1);
  u[1][2] = cmul(u[1][2], w2);
  u[1][3] = cmul(u[1][3], w3);
  u[2][1] = cmul(u[2][1], w2);
  u[2][2] = cmul_si(u[2][2], sign);  // W^4 = sign i
  u[2][3] = cmul(u[2][3], w6);
  u[3][1] = cmul(u[3][1], w3);
  u[3][2] = cmul(u[3][2], w6);
  u[3][3] = cmul(u[3][3], w9);
#pragma unroll
  for (int b = 0; b < 4; b++) {
    C w[4] = {u[0][b], u[1][b], u[2][b], u[3][b]};
    dft4<T>(w, sign);
#pragma unroll
    for (int c = 0; c < 4; c++) v[b + 4 * c] = w[c];
  }
}

// The z passes' exchange: pencil p (0..P-1) of a block, thread t of the
// pencil the fastest index.  buf: 16 rows of 16 P + 1 reals (the odd row
// length puts the transposing reads on distinct banks).  Ends with a
// barrier; every thread of the block takes part.
template <typename T, int P = 16>
__device__ __forceinline__ void exchange16_zt(typename C2<T>::t v[16], T* buf, int p, int t) {
  constexpr int S = 16 * P + 1;
  T re[16];
#pragma unroll
  for (int k = 0; k < 16; k++) buf[k * S + p * 16 + t] = v[k].x;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; k++) re[k] = buf[t * S + p * 16 + k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; k++) buf[k * S + p * 16 + t] = v[k].y;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; k++) v[k] = cmake<T>(re[k], buf[t * S + p * 16 + k]);
  __syncthreads();
}

// reals of the z passes' LDS buffer for P pencils: the exchange (16 rows of
// 16 P + 1) and the forward pass's unpack (P rows of 257)
template <int P> constexpr int zbuf_reals() { return 16 * (16 * P + 1) > 257 * P ? 16 * (16 * P + 1) : 257 * P; }

// Inverse z transform of one pencil of length 256, P pencils per block:
// the half spectra sa, sb (bins 0..128) of real components a and b packed as
// a + i b (sb null: b is 0).  On return thread t holds a + i b at z = t + 16 k2 in v[k2], unscaled.
// tw: W256^j, j = 0..255 (forward sign).
template <typename T, int P = 16>
__device__ __forceinline__ void z_c2r256_pencil(const typename C2<T>::t* __restrict__ sa,
                                                const typename C2<T>::t* __restrict__ sb, bool active,
                                                const typename C2<T>::t* __restrict__ tw, T* buf, int p, int t,
                                                typename C2<T>::t v[16]) {
#pragma clang fp contract(off)
  using C = typename C2<T>::t;
#pragma unroll
  for (int j = 0; j < 16; j++) {
    const int k = t + 16 * j;
    const int kk = k <= 128 ? k : 256 - k;  // Z_-k from the conjugates
    C xa = cmake<T>(0, 0), xb = cmake<T>(0, 0);
    if (active) {
      xa = sa[kk];
      if (sb) xb = sb[kk];
    }
    if (kk == 0 || kk == 128) {  // C2R reads only the real part of the self-conjugate bins
      xa.y = 0;
      xb.y = 0;
    }
    // Z_k = X_a + i X_b ;  Z_-k = conj X_a + i conj X_b
    v[j] = k <= 128 ? cmake<T>(xa.x - xb.y, xa.y + xb.x) : cmake<T>(xa.x + xb.y, xb.x - xa.y);
  }
  dft16<T>(v, +1);
#pragma unroll
  for (int k = 1; k < 16; k++) {
    C w = tw[t * k];
    w.y = -w.y;
    v[k] = cmul(v[k], w);
  }
  exchange16_zt<T, P>(v, buf, p, t);
  dft16<T>(v, +1);  // thread t holds z = t + 16 k2
}

// Forward z transform of one pencil of length 256 from v (thread t holds
// a + i b at z = t + 16 j in v[j]) to the half spectra sa, sb (bins 0..128;
// sb null: b is not stored).  Every thread of the block takes part.
template <typename T, int P = 16>
__device__ __forceinline__ void z_r2c256_pencil(typename C2<T>::t v[16], bool active,
                                                const typename C2<T>::t* __restrict__ tw, T* buf, int p, int t,
                                                typename C2<T>::t* __restrict__ sa, typename C2<T>::t* __restrict__ sb) {
#pragma clang fp contract(off)
  dft16<T>(v, -1);
#pragma unroll
  for (int k = 1; k < 16; k++) v[k] = cmul(v[k], tw[t * k]);
  exchange16_zt<T, P>(v, buf, p, t);
  dft16<T>(v, -1);  // thread t now holds Z[t + 16 k2]
  // Hermitian unpack: X_a = (Z_k + conj Z_-k) / 2, X_b = -i/2 (Z_k - conj Z_-k)
  // for k = t + 16 j <= 128; Z in natural order through LDS, real parts first
  T zr[9], zi[9], nr[9], ni[9];
#pragma unroll
  for (int ph = 0; ph < 2; ph++) {
#pragma unroll
    for (int k2 = 0; k2 < 16; k2++) buf[p * 257 + t + 16 * k2] = ph == 0 ? v[k2].x : v[k2].y;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 9; j++) {
      const int k = t + 16 * j;
      const int kn = (256 - k) & 255;
      const T a = k <= 128 ? buf[p * 257 + k] : (T)0;
      const T b = k <= 128 ? buf[p * 257 + kn] : (T)0;
      if (ph == 0) { zr[j] = a; nr[j] = b; } else { zi[j] = a; ni[j] = b; }
    }
    __syncthreads();
  }
  if (active) {
    const T h = (T)0.5;
#pragma unroll
    for (int j = 0; j < 9; j++) {
      const int k = t + 16 * j;
      if (k > 128) continue;
      sa[k] = cmake<T>(h * (zr[j] + nr[j]), h * (zi[j] - ni[j]));
      if (sb) sb[k] = cmake<T>(h * (zi[j] + ni[j]), -h * (zr[j] - nr[j]));
    }
  }
}

// ---------------------------------------------------------------------------
// 512 = 16 x 16 x 2: thirty-two threads own a pencil.
//  1. thread t holds x[t + 32 j] (j < 16); a 16-point DFT over j and the
//     twiddle W512^(t k1) give Y_t[k1];
//  2. through LDS, thread u = 2 k1 + t0 gathers Y_{2 i + t0}[k1] (i < 16) and
//     takes a 16-point DFT over i: A_t0[m];
//  3. X[k1 + 16 m] = A_0[m] + W32^m A_1[m], X[k1 + 16 (m + 16)] = A_0[m] -
//     W32^m A_1[m], the partner thread (u ^ 1) the neighbouring lane: thread
//     u ends with X[k1 + 16 (m + 16 t0)] for m < 16.
// pb: this pencil's 512 reals of LDS, the 16 x 32 gather array with row k
// rotated by 2 k so the gathering reads of one i spread over the banks.
// Every lane of the block takes part (the barriers); ends with a barrier.
// tw: W512^j, j = 0..511 (forward sign).
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void fft512_z(typename C2<T>::t v[16], T* pb, const typename C2<T>::t* __restrict__ tw,
                                         int t, int sign) {
#pragma clang fp contract(off)
  using C = typename C2<T>::t;
  dft16<T>(v, sign);
#pragma unroll
  for (int k = 1; k < 16; k++) {  // W512^(t k1)
    C w = tw[t * k];
    if (sign > 0) w.y = -w.y;
    v[k] = cmul(v[k], w);
  }
  const int k1 = t >> 1, t0 = t & 1;
  T re[16];
#pragma unroll
  for (int k = 0; k < 16; k++) pb[k * 32 + ((t + 2 * k) & 31)] = v[k].x;
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 16; i++) re[i] = pb[k1 * 32 + ((2 * i + t0 + 2 * k1) & 31)];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; k++) pb[k * 32 + ((t + 2 * k) & 31)] = v[k].y;
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 16; i++) v[i] = cmake<T>(re[i], pb[k1 * 32 + ((2 * i + t0 + 2 * k1) & 31)]);
  __syncthreads();
  dft16<T>(v, sign);  // A_t0[m]
#pragma unroll
  for (int m = 0; m < 16; m++) {  // the partner (t ^ 1) is the neighbouring lane
    C o;
    o.x = __shfl_xor(v[m].x, 1);
    o.y = __shfl_xor(v[m].y, 1);
    C w = tw[16 * m];  // W32^m
    if (sign > 0) w.y = -w.y;
    const C a0 = t0 == 0 ? v[m] : o, a1 = cmul(t0 == 0 ? o : v[m], w);
    v[m] = t0 == 0 ? cadd(a0, a1) : csub(a0, a1);
  }
}

// Inverse z transform of one pencil of length 512 (32 threads, thread t of
// the pencil): the half spectra sa, sb (bins 0..256) of real components a
// and b packed as a + i b (sb null: b is 0).  On return thread t holds a + i b
// at z = k1 + 16 (m + 16 t0) in v[m] (k1 = t >> 1, t0 = t & 1), unscaled.
template <typename T>
__device__ __forceinline__ void z_c2r512_pencil(const typename C2<T>::t* __restrict__ sa,
                                                const typename C2<T>::t* __restrict__ sb, bool active,
                                                const typename C2<T>::t* __restrict__ tw, T* pb, int t,
                                                typename C2<T>::t v[16]) {
#pragma clang fp contract(off)
  using C = typename C2<T>::t;
#pragma unroll
  for (int j = 0; j < 16; j++) {
    const int k = t + 32 * j;
    const int kk = k <= 256 ? k : 512 - k;  // Z_-k from the conjugates
    C xa = cmake<T>(0, 0), xb = cmake<T>(0, 0);
    if (active) {
      xa = sa[kk];
      if (sb) xb = sb[kk];
    }
    if (kk == 0 || kk == 256) {  // C2R reads only the real part of the self-conjugate bins
      xa.y = 0;
      xb.y = 0;
    }
    // Z_k = X_a + i X_b ;  Z_-k = conj X_a + i conj X_b
    v[j] = k <= 256 ? cmake<T>(xa.x - xb.y, xa.y + xb.x) : cmake<T>(xa.x + xb.y, xb.x - xa.y);
  }
  fft512_z<T>(v, pb, tw, t, +1);
}

// Forward z transform of one pencil of length 512 from v (thread t holds
// a + i b at z = t + 32 j in v[j]) to the half spectra sa, sb (bins 0..256;
// sb null: b is not stored).  Every lane of the block takes part.
template <typename T>
__device__ __forceinline__ void z_r2c512_pencil(typename C2<T>::t v[16], bool active,
                                                const typename C2<T>::t* __restrict__ tw, T* pb, int t,
                                                typename C2<T>::t* __restrict__ sa, typename C2<T>::t* __restrict__ sb) {
#pragma clang fp contract(off)
  fft512_z<T>(v, pb, tw, t, -1);
  // Hermitian unpack: X_a = (Z_k + conj Z_-k) / 2, X_b = -i/2 (Z_k - conj Z_-k)
  // for k = t + 32 j <= 256; Z in natural order through LDS, real parts first
  const int k1 = t >> 1, t0 = t & 1;
  T zr[9], zi[9], nr[9], ni[9];
#pragma unroll
  for (int ph = 0; ph < 2; ph++) {
#pragma unroll
    for (int m = 0; m < 16; m++) pb[k1 + 16 * m + 256 * t0] = ph == 0 ? v[m].x : v[m].y;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 9; j++) {
      const int k = t + 32 * j;
      const int kn = (512 - k) & 511;
      const T a = k <= 256 ? pb[k] : (T)0;
      const T b = k <= 256 ? pb[kn] : (T)0;
      if (ph == 0) { zr[j] = a; nr[j] = b; } else { zi[j] = a; ni[j] = b; }
    }
    __syncthreads();
  }
  if (active) {
    const T h = (T)0.5;
#pragma unroll
    for (int j = 0; j < 9; j++) {
      const int k = t + 32 * j;
      if (k > 256) continue;
      sa[k] = cmake<T>(h * (zr[j] + nr[j]), h * (zi[j] - ni[j]));
      if (sb) sb[k] = cmake<T>(h * (zi[j] + ni[j]), -h * (zr[j] - nr[j]));
    }
  }
}

}  // namespace vpfft
