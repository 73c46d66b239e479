// fft_line.h -- in-workgroup FFT of one "line" (a row or a column of the
// four-step N = N1 x N2 decomposition) for gfx950.
//
// A line of length L (8..256, power of two) is owned by T = L/8 consecutive
// lanes; lane t holds the 8 points x[t + i*T], i = 0..7, in registers. The
// transform is a Stockham autosort sequence of radix-8 stages followed by at
// most one radix-4 or radix-2 stage; stages exchange data through the line's
// LDS slice (L complex values). On return lane t holds X[t + i*T] in natural
// order, so the caller's prologue and epilogue address memory the same way.
//
// Twiddles come from a table tw[j] = exp(-2*pi*i*j/L), j < L, computed on the
// host in long double and rounded once; the inverse sign conjugates them.
#pragma once

#include <hip/hip_runtime.h>

namespace mmp {

struct cd {
  double re, im;
};

__device__ __forceinline__ cd cmul(cd a, cd b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ cd cmulc(cd a, cd b) {  // a * conj(b)
  return {a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im};
}
__device__ __forceinline__ cd cadd(cd a, cd b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cd csub(cd a, cd b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cd cscale(cd a, double s) { return {a.re * s, a.im * s}; }
__device__ __forceinline__ cd cconj(cd a) { return {a.re, -a.im}; }

// multiply by (SIGN * i)
template <int SIGN>
__device__ __forceinline__ cd mul_si(cd a) {
  return SIGN < 0 ? cd{a.im, -a.re} : cd{-a.im, a.re};
}

// DFT of length 2, 4, 8 with kernel exp(SIGN * 2*pi*i*m*q/R), in place, natural order.
template <int SIGN>
__device__ __forceinline__ void dft2(cd& a0, cd& a1) {
  cd t = a0;
  a0 = cadd(t, a1);
  a1 = csub(t, a1);
}

template <int SIGN>
__device__ __forceinline__ void dft4(cd& a0, cd& a1, cd& a2, cd& a3) {
  cd c0 = cadd(a0, a2), c1 = csub(a0, a2);
  cd d0 = cadd(a1, a3), d1 = mul_si<SIGN>(csub(a1, a3));
  a0 = cadd(c0, d0);
  a2 = csub(c0, d0);
  a1 = cadd(c1, d1);
  a3 = csub(c1, d1);
}

template <int SIGN>
__device__ __forceinline__ void dft8(cd* v) {
  constexpr double r = 0.70710678118654752440;  // cos(pi/4)
  cd a[4], b[4];
#pragma unroll
  for (int m = 0; m < 4; m++) {
    a[m] = cadd(v[m], v[m + 4]);
    b[m] = csub(v[m], v[m + 4]);
  }
  // b[m] *= w8^m, w8 = exp(SIGN*i*pi/4)
  b[1] = cscale(SIGN < 0 ? cd{b[1].re + b[1].im, b[1].im - b[1].re} : cd{b[1].re - b[1].im, b[1].im + b[1].re}, r);
  b[2] = mul_si<SIGN>(b[2]);
  b[3] = cscale(SIGN < 0 ? cd{-b[3].re + b[3].im, -b[3].im - b[3].re} : cd{-b[3].re - b[3].im, -b[3].im + b[3].re}, r);
  dft4<SIGN>(a[0], a[1], a[2], a[3]);
  dft4<SIGN>(b[0], b[1], b[2], b[3]);
#pragma unroll
  for (int q = 0; q < 4; q++) {
    v[2 * q] = a[q];
    v[2 * q + 1] = b[q];
  }
}

template <int SIGN>
__device__ __forceinline__ cd twiddle(const cd* __restrict__ tw, int j) {
  cd w = tw[j];
  if (SIGN > 0) w.im = -w.im;
  return w;
}

// Number of radix-8 stages and the trailing radix (1, 2 or 4) for length L.
template <int L>
struct LinePlan {
  static constexpr int n8 = (L >= 512) ? 3 : (L >= 64 ? 2 : 1);
  static constexpr int tail = L / (n8 == 3 ? 512 : (n8 == 2 ? 64 : 8));
  static_assert(L >= 8 && L <= 256 && (L & (L - 1)) == 0, "line length must be a power of two in [8, 256]");
  static constexpr int T = L / 8;
};

// LDS placement of element pos of a line. Row passes give each line its own
// slice padded by one slot every 8 (lanes writing 8t+q then land in distinct
// banks); column passes interleave the CB lines of neighbouring columns
// (stride CB), so the CB consecutive lanes of a column group touch
// consecutive 16-byte slots.
template <int STRIDE, int PAD>
__device__ __forceinline__ int lds_pos(int pos) {
  return PAD ? (pos + (pos >> 3)) * STRIDE : pos * STRIDE;
}

// LDS slots one line occupies (before the stride).
template <int L, int PAD>
struct LineSlots {
  static constexpr int value = PAD ? L + L / 8 : L;
};

// One radix-8 Stockham stage with span NS (NS > 1) that reads the line from LDS
// and writes it back to LDS (not the last stage).
template <int L, int NS, int SIGN, int STRIDE, int PAD>
__device__ __forceinline__ void stage8_lds(cd* v, cd* sm, const cd* __restrict__ tw, int t) {
  constexpr int T = L / 8;
#pragma unroll
  for (int i = 0; i < 8; i++) v[i] = sm[lds_pos<STRIDE, PAD>(t + i * T)];
  const int k = t % NS;
  constexpr int step = L / (NS * 8);
#pragma unroll
  for (int m = 1; m < 8; m++) v[m] = cmul(v[m], twiddle<SIGN>(tw, k * m * step));
  dft8<SIGN>(v);
  __syncthreads();
  const int base = (t / NS) * NS * 8 + k;
#pragma unroll
  for (int q = 0; q < 8; q++) sm[lds_pos<STRIDE, PAD>(base + q * NS)] = v[q];
  __syncthreads();
}

// In-place transform of the line held as v[i] = x[t + i*T]; sm is the line's
// LDS slice (L values). All lanes of the workgroup must call this together.
// sm points at the line's first slot; element pos is at sm[lds_pos<STRIDE, PAD>(pos)].
template <int L, int SIGN, int STRIDE, int PAD>
__device__ __forceinline__ void line_fft(cd* v, cd* sm, const cd* __restrict__ tw, int t) {
  using P = LinePlan<L>;
  constexpr int T = P::T;
  // stage 1: radix-8, span 1, inputs x[t + m*T]
  dft8<SIGN>(v);
  if constexpr (L == 8) return;
  __syncthreads();  // the slice may still be read by a previous use
  // write x'[8t + q]
#pragma unroll
  for (int q = 0; q < 8; q++) sm[lds_pos<STRIDE, PAD>(8 * t + q)] = v[q];
  __syncthreads();
  if constexpr (P::n8 == 2 && P::tail == 1) {
    // last stage: radix-8, span 8 = T, outputs stay in registers (positions t + q*T)
#pragma unroll
    for (int i = 0; i < 8; i++) v[i] = sm[lds_pos<STRIDE, PAD>(t + i * T)];
    const int k = t % 8;
#pragma unroll
    for (int m = 1; m < 8; m++) v[m] = cmul(v[m], twiddle<SIGN>(tw, k * m * (L / 64)));
    dft8<SIGN>(v);
    __syncthreads();
    return;
  } else {
    constexpr int NS = (P::n8 == 2) ? 64 : 8;
    if constexpr (P::n8 == 2) stage8_lds<L, 8, SIGN, STRIDE, PAD>(v, sm, tw, t);
    // tail stage with span NS = L / tail: registers slot i = b + tail' * m
#pragma unroll
    for (int i = 0; i < 8; i++) v[i] = sm[lds_pos<STRIDE, PAD>(t + i * T)];
    __syncthreads();
    if constexpr (P::tail == 4) {
      // two butterflies b = 0, 1: j = t + b*T, inputs slots b + 2m, twiddle W_L^(j*m)
#pragma unroll
      for (int b = 0; b < 2; b++) {
        const int j = t + b * T;
        cd a0 = v[b], a1 = cmul(v[b + 2], twiddle<SIGN>(tw, j)), a2 = cmul(v[b + 4], twiddle<SIGN>(tw, 2 * j)),
           a3 = cmul(v[b + 6], twiddle<SIGN>(tw, 3 * j));
        dft4<SIGN>(a0, a1, a2, a3);
        v[b] = a0;
        v[b + 2] = a1;
        v[b + 4] = a2;
        v[b + 6] = a3;
      }
    } else if constexpr (P::tail == 2) {
#pragma unroll
      for (int b = 0; b < 4; b++) {
        const int j = t + b * T;
        cd a0 = v[b], a1 = cmul(v[b + 4], twiddle<SIGN>(tw, j));
        dft2<SIGN>(a0, a1);
        v[b] = a0;
        v[b + 4] = a1;
      }
    } else {
      // tail == 1 with n8 == 1 cannot happen for L > 8
    }
    (void)NS;
  }
}

}  // namespace mmp
