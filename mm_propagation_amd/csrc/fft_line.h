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

// Complex value of the work precision T (double, or float for the mixed-precision path),
// aligned to its size so that one value moves in one 16-byte (8-byte) memory access.
template <class T>
struct alignas(2 * sizeof(T)) cpx {
  T re, im;
};
using cd = cpx<double>;
using cf = cpx<float>;

template <class T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <class T>
__device__ __forceinline__ cpx<T> cmulc(cpx<T> a, cpx<T> b) {  // a * conj(b)
  return {a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im};
}
template <class T>
__device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) {
  return {a.re + b.re, a.im + b.im};
}
template <class T>
__device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) {
  return {a.re - b.re, a.im - b.im};
}
template <class T, class S>
__device__ __forceinline__ cpx<T> cscale(cpx<T> a, S s) {
  const T f = (T)s;
  return {a.re * f, a.im * f};
}
template <class T>
__device__ __forceinline__ cpx<T> cconj(cpx<T> a) {
  return {a.re, -a.im};
}
// precision conversions (float -> double is exact; double -> float rounds to nearest)
__device__ __forceinline__ cd widen(cf a) { return {(double)a.re, (double)a.im}; }
__device__ __forceinline__ cd widen(cd a) { return a; }
template <class T>
__device__ __forceinline__ cpx<T> narrow(cd a) {
  return {(T)a.re, (T)a.im};
}

// multiply by (SIGN * i)
template <int SIGN, class T>
__device__ __forceinline__ cpx<T> mul_si(cpx<T> a) {
  return SIGN < 0 ? cpx<T>{a.im, -a.re} : cpx<T>{-a.im, a.re};
}

// DFT of length 2, 4, 8 with kernel exp(SIGN * 2*pi*i*m*q/R), in place, natural order.
template <int SIGN, class T>
__device__ __forceinline__ void dft2(cpx<T>& a0, cpx<T>& a1) {
  cpx<T> t = a0;
  a0 = cadd(t, a1);
  a1 = csub(t, a1);
}

template <int SIGN, class T>
__device__ __forceinline__ void dft4(cpx<T>& a0, cpx<T>& a1, cpx<T>& a2, cpx<T>& a3) {
  cpx<T> c0 = cadd(a0, a2), c1 = csub(a0, a2);
  cpx<T> d0 = cadd(a1, a3), d1 = mul_si<SIGN>(csub(a1, a3));
  a0 = cadd(c0, d0);
  a2 = csub(c0, d0);
  a1 = cadd(c1, d1);
  a3 = csub(c1, d1);
}

template <int SIGN, class T>
__device__ __forceinline__ void dft8(cpx<T>* v) {
  constexpr T r = (T)0.70710678118654752440L;  // cos(pi/4), rounded once to T
  cpx<T> a[4], b[4];
#pragma unroll
  for (int m = 0; m < 4; m++) {
    a[m] = cadd(v[m], v[m + 4]);
    b[m] = csub(v[m], v[m + 4]);
  }
  // b[m] *= w8^m, w8 = exp(SIGN*i*pi/4)
  b[1] = cscale(SIGN < 0 ? cpx<T>{b[1].re + b[1].im, b[1].im - b[1].re} : cpx<T>{b[1].re - b[1].im, b[1].im + b[1].re},
                r);
  b[2] = mul_si<SIGN>(b[2]);
  b[3] = cscale(SIGN < 0 ? cpx<T>{-b[3].re + b[3].im, -b[3].im - b[3].re}
                         : cpx<T>{-b[3].re - b[3].im, -b[3].im + b[3].re},
                r);
  dft4<SIGN>(a[0], a[1], a[2], a[3]);
  dft4<SIGN>(b[0], b[1], b[2], b[3]);
#pragma unroll
  for (int q = 0; q < 4; q++) {
    v[2 * q] = a[q];
    v[2 * q + 1] = b[q];
  }
}

template <int SIGN, class T>
__device__ __forceinline__ cpx<T> twiddle(const cpx<T>* __restrict__ tw, int j) {
  cpx<T> w = tw[j];
  if (SIGN > 0) w.im = -w.im;
  return w;
}

// Number of radix-8 stages and the trailing radix (1, 2 or 4) for length L.
template <int L>
struct LinePlan {
  static constexpr int n8 = (L >= 512) ? 3 : (L >= 64 ? 2 : 1);
  static constexpr int tail = L / (n8 == 3 ? 512 : (n8 == 2 ? 64 : 8));
  static_assert(L >= 8 && L <= 1024 && (L & (L - 1)) == 0, "line length must be a power of two in [8, 1024]");
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

// Stage twiddles a lane needs for one line transform, loaded before the data
// arrives so their latency overlaps the data loads. Slots:
//   [0..6]   radix-8 stage of span 8 (L >= 64):   W_L^((t%8) * m * L/64), m = 1..7
//   [7..13]  radix-8 stage of span 64 (L >= 512): W_L^((t%64) * m * L/512), m = 1..7
//   then the trailing stage, butterflies j = t + b*T:
//     radix-4: W_L^(j*m), m = 1..3 (3 per butterfly, b = 0, 1)
//     radix-2: W_L^j (b = 0..3)
template <int L, class T = double>
struct LineTw {
  static constexpr int n8 = LinePlan<L>::n8;
  static constexpr int tail = LinePlan<L>::tail;
  static constexpr int base = (n8 == 3) ? 14 : ((n8 == 2) ? 7 : 0);
  static constexpr int count = (L == 8) ? 1 : base + (tail == 4 ? 6 : (tail == 2 ? 4 : 0));
  cpx<T> w[count];
};

// Loads the forward-sign twiddles (the inverse transform conjugates them on use).
template <int L, class T>
__device__ __forceinline__ void load_line_tw(LineTw<L, T>& lt, const cpx<T>* __restrict__ tw, int t) {
  using P = LinePlan<L>;
  constexpr int LT = P::T;
  if constexpr (L == 8) {
    (void)tw;
    (void)t;
    return;
  } else {
    if constexpr (P::n8 >= 2) {
      const int k = t % 8;
#pragma unroll
      for (int m = 1; m < 8; m++) lt.w[m - 1] = tw[k * m * (L / 64)];
    }
    if constexpr (P::n8 == 3) {
      const int k = t % 64;
#pragma unroll
      for (int m = 1; m < 8; m++) lt.w[6 + m] = tw[k * m * (L / 512)];
    }
    constexpr int b0 = LineTw<L, T>::base;
    if constexpr (P::tail == 4) {
#pragma unroll
      for (int b = 0; b < 2; b++)
#pragma unroll
        for (int m = 1; m < 4; m++) lt.w[b0 + 3 * b + m - 1] = tw[(t + b * LT) * m];
    } else if constexpr (P::tail == 2) {
#pragma unroll
      for (int b = 0; b < 4; b++) lt.w[b0 + b] = tw[t + b * LT];
    }
  }
}

// Orders the LDS accesses of one wavefront around a data exchange between its lanes.
// LDS operations of a single wavefront are performed in order, so no wait is needed;
// the fences only keep the compiler from moving accesses across this point.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Exchange barrier of a line transform: the whole workgroup, or only the wavefront when
// every line lies inside one wavefront (L/8 lanes per line, lines aligned to it).
template <bool WAVE>
__device__ __forceinline__ void line_sync() {
  if constexpr (WAVE) wave_lds_sync();
  else __syncthreads();
}

template <int SIGN, class T>
__device__ __forceinline__ cpx<T> tw_dir(cpx<T> w) {
  return SIGN < 0 ? w : cpx<T>{w.re, -w.im};
}

// In-place transform of the line held as v[i] = x[t + i*T]; sm points at the
// line's first LDS slot and element pos lives at sm[lds_pos<STRIDE, PAD>(pos)].
// All lanes of the workgroup (WAVE = false) or of the wavefront (WAVE = true) must
// call this together.
template <int L, int SIGN, int STRIDE, int PAD, class T, bool WAVE = false>
__device__ __forceinline__ void line_fft(cpx<T>* v, cpx<T>* sm, const LineTw<L, T>& lt, int t) {
  using P = LinePlan<L>;
  constexpr int LT = P::T;
  // stage 1: radix-8, span 1, inputs x[t + m*T]
  dft8<SIGN>(v);
  if constexpr (L == 8) {
    (void)sm;
    (void)lt;
    (void)t;
    return;
  } else {
    line_sync<WAVE>();  // the slice may still be read by a previous use
#pragma unroll
    for (int q = 0; q < 8; q++) sm[lds_pos<STRIDE, PAD>(8 * t + q)] = v[q];
    line_sync<WAVE>();
#pragma unroll
    for (int i = 0; i < 8; i++) v[i] = sm[lds_pos<STRIDE, PAD>(t + i * LT)];
    if constexpr (P::n8 >= 2) {
      // radix-8 stage of span 8
#pragma unroll
      for (int m = 1; m < 8; m++) v[m] = cmul(v[m], tw_dir<SIGN>(lt.w[m - 1]));
      dft8<SIGN>(v);
      if constexpr (P::n8 == 2 && P::tail == 1) {
        line_sync<WAVE>();
        return;  // outputs at positions t + q*LT
      }
      line_sync<WAVE>();
      const int base = (t / 8) * 64 + t % 8;
#pragma unroll
      for (int q = 0; q < 8; q++) sm[lds_pos<STRIDE, PAD>(base + q * 8)] = v[q];
      line_sync<WAVE>();
#pragma unroll
      for (int i = 0; i < 8; i++) v[i] = sm[lds_pos<STRIDE, PAD>(t + i * LT)];
    }
    if constexpr (P::n8 == 3) {
      // radix-8 stage of span 64
#pragma unroll
      for (int m = 1; m < 8; m++) v[m] = cmul(v[m], tw_dir<SIGN>(lt.w[6 + m]));
      dft8<SIGN>(v);
      if constexpr (P::tail == 1) {
        line_sync<WAVE>();
        return;  // outputs at positions t + q*LT
      }
      line_sync<WAVE>();
      const int base = (t / 64) * 512 + t % 64;
#pragma unroll
      for (int q = 0; q < 8; q++) sm[lds_pos<STRIDE, PAD>(base + q * 64)] = v[q];
      line_sync<WAVE>();
#pragma unroll
      for (int i = 0; i < 8; i++) v[i] = sm[lds_pos<STRIDE, PAD>(t + i * LT)];
    }
    line_sync<WAVE>();
    constexpr int b0 = LineTw<L, T>::base;
    if constexpr (P::tail == 4) {
#pragma unroll
      for (int b = 0; b < 2; b++) {
        cpx<T> a0 = v[b], a1 = cmul(v[b + 2], tw_dir<SIGN>(lt.w[b0 + 3 * b])),
           a2 = cmul(v[b + 4], tw_dir<SIGN>(lt.w[b0 + 3 * b + 1])), a3 = cmul(v[b + 6], tw_dir<SIGN>(lt.w[b0 + 3 * b + 2]));
        dft4<SIGN>(a0, a1, a2, a3);
        v[b] = a0;
        v[b + 2] = a1;
        v[b + 4] = a2;
        v[b + 6] = a3;
      }
    } else if constexpr (P::tail == 2) {
#pragma unroll
      for (int b = 0; b < 4; b++) {
        cpx<T> a0 = v[b], a1 = cmul(v[b + 4], tw_dir<SIGN>(lt.w[b0 + b]));
        dft2<SIGN>(a0, a1);
        v[b] = a0;
        v[b + 4] = a1;
      }
    }
  }
}

}  // namespace mmp
