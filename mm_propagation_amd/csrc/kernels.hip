// kernels.hip -- gfx950 kernels of the Lawson-Runge-Kutta propagation step.
//
// One RK stage i of a step is, in the four-step layout of kernels.h:
//   row_head(i)        rows: finish N_{i-1} (row DFT of sigma, conj(E) phase,
//                      RK accumulate) and build E_i * (psi + h sum a_ij N_j),
//                      then the inverse row DFT
//   col_inverse        columns: inverse column DFT -> U (time domain)
//   col_pairs_forward  columns: pair products of U, forward column DFT   (Raman on)
//   row_raman          rows: DFT, multiply by the Raman spectrum, inverse DFT
//   col_inverse        columns: -> Raman-convolved pair products (time domain)
//   gather             samples: Kerr + Raman sum sigma_p
//   col_forward        columns: forward column DFT of sigma
// and row_head(s) closes the step. The state psi is kept as its spectrum, so the
// reference's F^-1 F round trips between stages vanish (exact algebra; results
// differ from the reference by rounding only).
#include "kernels.h"

namespace mmp {

// ---------------------------------------------------------------------------
// configuration of the row and column passes
// ---------------------------------------------------------------------------
template <int L>
struct RowCfg {
  static constexpr int T = L / 8;                  // lanes per row
  static constexpr int G = T >= 64 ? 1 : 64 / T;   // rows per workgroup
  static constexpr int THREADS = G * T;            // 64
  static constexpr int SLOTS = LineSlots<L, 1>::value;
};

constexpr int kCB = 8;  // columns per column-pass workgroup (8 x 16 B = one 128 B line)

template <int L>
struct ColCfg {
  static constexpr int T = L / 8;                                 // lanes per column line
  static constexpr int G = (kCB * T >= 64) ? 1 : 64 / (kCB * T);  // lines per workgroup
  static constexpr int THREADS = G * kCB * T;
};

__device__ __forceinline__ cd phase(double imL, double tau) {
  double s, c;
  sincos(imL * tau, &s, &c);
  return {c, s};
}

// ---------------------------------------------------------------------------
// row stage head
// ---------------------------------------------------------------------------
template <int L>
__global__ void __launch_bounds__(64) k_row_head(Geometry g, const cd* __restrict__ rowtw, HeadArgs a) {
  using C = RowCfg<L>;
  __shared__ cd smem[C::G * C::SLOTS];
  const int tid = threadIdx.x;
  const int gl = tid / C::T, t = tid % C::T;
  const int row = blockIdx.x * C::G + gl;
  const bool valid = row < g.M * g.N1;
  const int p = valid ? row / g.N1 : 0;
  const int k1 = valid ? row % g.N1 : 0;
  const size_t base = (size_t)p * g.N + (size_t)k1 * g.N2;
  cd* sm = smem + gl * C::SLOTS;
  const double z = a.st->z;
  const size_t MN = (size_t)g.M * g.N;

  cd v[8];
  cd nh[8];
#pragma unroll
  for (int i = 0; i < 8; i++) nh[i] = {0., 0.};
  if (a.has_prev) {
#pragma unroll
    for (int i = 0; i < 8; i++) v[i] = valid ? a.ysig[base + t + i * C::T] : cd{0., 0.};
    line_fft<L, -1, 1, 1>(v, sm, rowtw, t);
    if (valid) {
      const double tau = z + a.ch_prev;
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const size_t idx = base + t + i * C::T;
        const cd E = phase(a.imL[idx], tau);
        const cd x = cmulc(v[i], E);                 // sigma_hat * conj(E)
        const cd n = {-x.im * a.sgn, x.re * a.sgn};  // * (i sgn)
        nh[i] = n;
        cd acc = a.prev_is_first ? a.psi_hat[idx] : a.acc[idx];
        if (a.bh_prev != 0.) acc = cadd(acc, cscale(n, a.bh_prev));
        if (a.final_stage) a.psi_hat[idx] = acc;
        else a.acc[idx] = acc;
        if (a.stash_prev) a.stash[(size_t)a.prev_stage * MN + idx] = n;
      }
    }
  }
  if (a.final_stage) return;
  const double tau = z + a.ch_cur;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const size_t idx = base + t + i * C::T;
    if (!valid) {
      v[i] = {0., 0.};
      continue;
    }
    cd psi = a.psi_hat[idx];
    for (int e = 0; e < a.n_a; e++) {
      const int j = a.a_idx[e];
      const cd nj = (j == a.prev_stage && a.has_prev) ? nh[i] : a.stash[(size_t)j * MN + idx];
      psi = cadd(psi, cscale(nj, a.a_h[e]));
    }
    v[i] = cmul(phase(a.imL[idx], tau), psi);
  }
  line_fft<L, 1, 1, 1>(v, sm, rowtw, t);
  if (valid) {
#pragma unroll
    for (int i = 0; i < 8; i++) a.yout[base + t + i * C::T] = v[i];
  }
}

// ---------------------------------------------------------------------------
// column passes
// ---------------------------------------------------------------------------
struct ColIdx {
  int c, t, gl;
};

template <int L>
__device__ __forceinline__ ColIdx col_index() {
  using C = ColCfg<L>;
  const int tid = threadIdx.x;
  return {tid % kCB, (tid / kCB) % C::T, tid / (kCB * C::T)};
}

// inverse column DFT: in = spectral-side rows (k1), out = time domain, times scale, divided by divisor
template <int L>
__global__ void __launch_bounds__(256) k_col_inverse(Geometry g, const cd* __restrict__ coltw,
                                                     const cd* __restrict__ fourtw, const cd* in, cd* out, int lines,
                                                     double scale, double divisor) {
  using C = ColCfg<L>;
  __shared__ cd smem[C::G * L * kCB];
  const ColIdx ix = col_index<L>();
  const int q = blockIdx.y * C::G + ix.gl;
  const int n2 = blockIdx.x * kCB + ix.c;
  const bool valid = q < lines;
  const size_t base = (size_t)(valid ? q : 0) * g.N + n2;
  cd* sm = smem + ix.gl * L * kCB + ix.c;
  cd v[8];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const int k1 = ix.t + i * C::T;
    v[i] = valid ? cmulc(in[base + (size_t)k1 * g.N2], fourtw[n2 * k1]) : cd{0., 0.};
  }
  line_fft<L, 1, kCB, 0>(v, sm, coltw, ix.t);
  if (valid) {
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int n1 = ix.t + i * C::T;
      cd x = cscale(v[i], scale);
      if (divisor != 1.0) x = {x.re / divisor, x.im / divisor};
      out[base + (size_t)n1 * g.N2] = x;
    }
  }
}

// forward column DFT + four-step twiddle: in = time domain, out = spectral-side rows
template <int L>
__global__ void __launch_bounds__(256) k_col_forward(Geometry g, const cd* __restrict__ coltw,
                                                     const cd* __restrict__ fourtw, const cd* in, cd* out, int lines) {
  using C = ColCfg<L>;
  __shared__ cd smem[C::G * L * kCB];
  const ColIdx ix = col_index<L>();
  const int q = blockIdx.y * C::G + ix.gl;
  const int n2 = blockIdx.x * kCB + ix.c;
  const bool valid = q < lines;
  const size_t base = (size_t)(valid ? q : 0) * g.N + n2;
  cd* sm = smem + ix.gl * L * kCB + ix.c;
  cd v[8];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const int n1 = ix.t + i * C::T;
    v[i] = valid ? in[base + (size_t)n1 * g.N2] : cd{0., 0.};
  }
  line_fft<L, -1, kCB, 0>(v, sm, coltw, ix.t);
  if (valid) {
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int k1 = ix.t + i * C::T;
      out[base + (size_t)k1 * g.N2] = cmul(v[i], fourtw[n2 * k1]);
    }
  }
}

// forward column DFT of one real line (the Raman response)
template <int L>
__global__ void __launch_bounds__(256) k_col_forward_real(Geometry g, const cd* __restrict__ coltw,
                                                          const cd* __restrict__ fourtw, const double* __restrict__ in,
                                                          cd* __restrict__ out) {
  using C = ColCfg<L>;
  __shared__ cd smem[C::G * L * kCB];
  const ColIdx ix = col_index<L>();
  const bool valid = blockIdx.y * C::G + ix.gl < 1;
  const int n2 = blockIdx.x * kCB + ix.c;
  cd* sm = smem + ix.gl * L * kCB + ix.c;
  cd v[8];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const int n1 = ix.t + i * C::T;
    v[i] = valid ? cd{in[(size_t)n1 * g.N2 + n2], 0.} : cd{0., 0.};
  }
  line_fft<L, -1, kCB, 0>(v, sm, coltw, ix.t);
  if (valid) {
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int k1 = ix.t + i * C::T;
      out[(size_t)k1 * g.N2 + n2] = cmul(v[i], fourtw[n2 * k1]);
    }
  }
}

// pair products of the time-domain fields U, then forward column DFT
template <int L>
__global__ void __launch_bounds__(256) k_col_pairs_forward(Geometry g, const cd* __restrict__ coltw,
                                                           const cd* __restrict__ fourtw, const cd* __restrict__ U,
                                                           cd* __restrict__ out, int lines, PairLineArgs pl) {
  using C = ColCfg<L>;
  __shared__ cd smem[C::G * L * kCB];
  const ColIdx ix = col_index<L>();
  const int q = blockIdx.y * C::G + ix.gl;
  const int n2 = blockIdx.x * kCB + ix.c;
  const bool valid = q < lines;
  const int qq = valid ? q : 0;
  const uint32_t la = pl.line_a[qq], lb = pl.line_b[qq], self = pl.line_self[qq];
  cd* sm = smem + ix.gl * L * kCB + ix.c;
  cd v[8];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const int n1 = ix.t + i * C::T;
    const size_t tt = (size_t)n1 * g.N2 + n2;
    cd x = {0., 0.};
    if (valid) {
      const cd ua = U[(size_t)la * g.N + tt];
      if (self) {
        x.re = ua.re * ua.re + ua.im * ua.im;
        if (lb != 0xffffffffu) {
          const cd ub = U[(size_t)lb * g.N + tt];
          x.im = ub.re * ub.re + ub.im * ub.im;
        }
      } else {
        x = cmulc(ua, U[(size_t)lb * g.N + tt]);
      }
    }
    v[i] = x;
  }
  line_fft<L, -1, kCB, 0>(v, sm, coltw, ix.t);
  if (valid) {
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int k1 = ix.t + i * C::T;
      out[(size_t)q * g.N + (size_t)k1 * g.N2 + n2] = cmul(v[i], fourtw[n2 * k1]);
    }
  }
}

// ---------------------------------------------------------------------------
// row passes
// ---------------------------------------------------------------------------
template <int L>
__global__ void __launch_bounds__(64) k_row_forward(Geometry g, const cd* __restrict__ rowtw, cd* __restrict__ data,
                                                    int lines) {
  using C = RowCfg<L>;
  __shared__ cd smem[C::G * C::SLOTS];
  const int tid = threadIdx.x;
  const int gl = tid / C::T, t = tid % C::T;
  const int row = blockIdx.x * C::G + gl;
  const bool valid = row < lines * g.N1;
  const size_t base = valid ? (size_t)row * g.N2 : 0;  // rows of consecutive lines are contiguous
  cd* sm = smem + gl * C::SLOTS;
  cd v[8];
#pragma unroll
  for (int i = 0; i < 8; i++) v[i] = valid ? data[base + t + i * C::T] : cd{0., 0.};
  line_fft<L, -1, 1, 1>(v, sm, rowtw, t);
  if (valid) {
#pragma unroll
    for (int i = 0; i < 8; i++) data[base + t + i * C::T] = v[i];
  }
}

// optional phase exp(i ImL z) (the reference's exp(L z)), then inverse row DFT
template <int L>
__global__ void __launch_bounds__(64) k_row_inverse_phase(Geometry g, const cd* __restrict__ rowtw,
                                                          const cd* __restrict__ in, cd* __restrict__ out,
                                                          const double* __restrict__ imL,
                                                          const StepState* __restrict__ st, int use_phase) {
  using C = RowCfg<L>;
  __shared__ cd smem[C::G * C::SLOTS];
  const int tid = threadIdx.x;
  const int gl = tid / C::T, t = tid % C::T;
  const int row = blockIdx.x * C::G + gl;
  const bool valid = row < g.M * g.N1;
  const size_t base = valid ? (size_t)row * g.N2 : 0;
  cd* sm = smem + gl * C::SLOTS;
  const double z = st->z;
  cd v[8];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const size_t idx = base + t + i * C::T;
    cd x = valid ? in[idx] : cd{0., 0.};
    if (use_phase && valid) x = cmul(x, phase(imL[idx], z));
    v[i] = x;
  }
  line_fft<L, 1, 1, 1>(v, sm, rowtw, t);
  if (valid) {
#pragma unroll
    for (int i = 0; i < 8; i++) out[base + t + i * C::T] = v[i];
  }
}

// Raman convolution in the spectral rows: DFT, * H * hscale, inverse DFT (in place)
template <int L>
__global__ void __launch_bounds__(64) k_row_raman(Geometry g, const cd* __restrict__ rowtw, cd* __restrict__ data,
                                                  int lines, const cd* __restrict__ hspec, double hscale) {
  using C = RowCfg<L>;
  __shared__ cd smem[C::G * C::SLOTS];
  const int tid = threadIdx.x;
  const int gl = tid / C::T, t = tid % C::T;
  const int row = blockIdx.x * C::G + gl;
  const bool valid = row < lines * g.N1;
  const size_t base = valid ? (size_t)row * g.N2 : 0;
  const size_t hbase = valid ? (size_t)(row % g.N1) * g.N2 : 0;
  cd* sm = smem + gl * C::SLOTS;
  cd v[8];
#pragma unroll
  for (int i = 0; i < 8; i++) v[i] = valid ? data[base + t + i * C::T] : cd{0., 0.};
  line_fft<L, -1, 1, 1>(v, sm, rowtw, t);
#pragma unroll
  for (int i = 0; i < 8; i++) v[i] = cscale(cmul(v[i], hspec[hbase + t + i * C::T]), hscale);
  line_fft<L, 1, 1, 1>(v, sm, rowtw, t);
  if (valid) {
#pragma unroll
    for (int i = 0; i < 8; i++) data[base + t + i * C::T] = v[i];
  }
}

// ---------------------------------------------------------------------------
// Kerr + Raman gather over a tile of TS consecutive samples:
//   X_slot = kv * U_a conj(U_b) + kr * VR_slot        (slot mix, into LDS)
//   sigma_p = sum over the nonzeros (k, slot, q) of Q^(p):  q * X_slot * U_k
// U and X live in LDS; each (sample, mode) lane accumulates its own sum.
// ---------------------------------------------------------------------------
template <int CAP>
__global__ void __launch_bounds__(256) k_gather(Geometry g, const cd* __restrict__ U, const cd* __restrict__ VR,
                                                cd* __restrict__ sigma, GatherArgs a, int TS) {
  __shared__ cd sh[CAP];
  const int M = g.M, P = a.n_slots;
  cd* Ush = sh;            // [M][TS]
  cd* Xsh = sh + M * TS;   // [P][TS]
  const int t0 = blockIdx.x * TS;
  const int tid = threadIdx.x;
  for (int e = tid; e < M * TS; e += 256) {
    const int k = e / TS, j = e % TS;
    Ush[e] = U[(size_t)k * g.N + t0 + j];
  }
  __syncthreads();
  for (int e = tid; e < P * TS; e += 256) {
    const int s = e / TS, j = e % TS;
    const cd V = cmulc(Ush[a.slot_a[s] * TS + j], Ush[a.slot_b[s] * TS + j]);
    cd X = cscale(V, a.kv);
    if (a.raman) {
      const cd r = VR[(size_t)a.slot_line[s] * g.N + t0 + j];
      cd vr;
      switch (a.slot_kind[s]) {
        case kDirect: vr = r; break;
        case kConj: vr = cconj(r); break;
        case kRe: vr = {r.re, 0.}; break;
        default: vr = {r.im, 0.}; break;
      }
      X = cadd(X, cscale(vr, a.kr));
    }
    Xsh[e] = X;
  }
  __syncthreads();
  const int j = tid % TS;
  for (int p = tid / TS; p < M; p += 256 / TS) {
    cd acc = {0., 0.};
    const uint32_t e1 = a.off[p + 1];
    for (uint32_t e = a.off[p]; e < e1; e++) {
      const cd W = cmul(Xsh[a.slot[e] * TS + j], Ush[a.k[e] * TS + j]);
      acc = cadd(acc, cscale(W, a.q[e]));
    }
    sigma[(size_t)p * g.N + t0 + j] = acc;
  }
}

__global__ void k_advance_z(StepState* st, double h) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    st->z += h;
    st->steps += 1;
  }
}

// per-mode energy sum_j |A_p,j|^2 * dt; one workgroup per mode, fixed reduction order
__global__ void __launch_bounds__(256) k_energy(Geometry g, const cd* __restrict__ f, double* __restrict__ out,
                                                double dt) {
  __shared__ double red[256];
  const int p = blockIdx.x;
  double s = 0.;
  for (int j = threadIdx.x; j < g.N; j += 256) {
    const cd x = f[(size_t)p * g.N + j];
    s += x.re * x.re + x.im * x.im;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[p] = red[0] * dt;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
#define MMP_DISPATCH_L(Lv, FN, ...)                        \
  switch (Lv) {                                            \
    case 8: return FN<8>(__VA_ARGS__);                     \
    case 16: return FN<16>(__VA_ARGS__);                   \
    case 32: return FN<32>(__VA_ARGS__);                   \
    case 64: return FN<64>(__VA_ARGS__);                   \
    case 128: return FN<128>(__VA_ARGS__);                 \
    case 256: return FN<256>(__VA_ARGS__);                 \
    default: return hipErrorInvalidValue;                  \
  }

template <int L>
static hipError_t row_head_impl(const Geometry& g, const Twiddles& tw, const HeadArgs& a, hipStream_t s) {
  using C = RowCfg<L>;
  const int rows = g.M * g.N1;
  dim3 grid((rows + C::G - 1) / C::G);
  hipLaunchKernelGGL(k_row_head<L>, grid, dim3(C::THREADS), 0, s, g, tw.rowtw, a);
  return hipGetLastError();
}
hipError_t launch_row_head(const Geometry& g, const Twiddles& tw, const HeadArgs& a, hipStream_t s) {
  MMP_DISPATCH_L(g.N2, row_head_impl, g, tw, a, s);
}

template <int L>
static hipError_t col_inverse_impl(const Geometry& g, const Twiddles& tw, const cd* in, cd* out, int lines,
                                   double scale, double divisor, hipStream_t s) {
  using C = ColCfg<L>;
  dim3 grid(g.N2 / kCB, (lines + C::G - 1) / C::G);
  hipLaunchKernelGGL(k_col_inverse<L>, grid, dim3(C::THREADS), 0, s, g, tw.coltw, tw.fourtw, in, out, lines, scale,
                     divisor);
  return hipGetLastError();
}
hipError_t launch_col_inverse(const Geometry& g, const Twiddles& tw, const cd* in, cd* out, int lines, double scale,
                              double divisor, hipStream_t s) {
  MMP_DISPATCH_L(g.N1, col_inverse_impl, g, tw, in, out, lines, scale, divisor, s);
}

template <int L>
static hipError_t col_forward_impl(const Geometry& g, const Twiddles& tw, const cd* in, cd* out, int lines,
                                   hipStream_t s) {
  using C = ColCfg<L>;
  dim3 grid(g.N2 / kCB, (lines + C::G - 1) / C::G);
  hipLaunchKernelGGL(k_col_forward<L>, grid, dim3(C::THREADS), 0, s, g, tw.coltw, tw.fourtw, in, out, lines);
  return hipGetLastError();
}
hipError_t launch_col_forward(const Geometry& g, const Twiddles& tw, const cd* in, cd* out, int lines,
                              hipStream_t s) {
  MMP_DISPATCH_L(g.N1, col_forward_impl, g, tw, in, out, lines, s);
}

template <int L>
static hipError_t col_forward_real_impl(const Geometry& g, const Twiddles& tw, const double* in, cd* out,
                                        hipStream_t s) {
  using C = ColCfg<L>;
  dim3 grid(g.N2 / kCB, 1);
  hipLaunchKernelGGL(k_col_forward_real<L>, grid, dim3(C::THREADS), 0, s, g, tw.coltw, tw.fourtw, in, out);
  return hipGetLastError();
}
hipError_t launch_col_forward_real(const Geometry& g, const Twiddles& tw, const double* in, cd* out,
                                   hipStream_t s) {
  MMP_DISPATCH_L(g.N1, col_forward_real_impl, g, tw, in, out, s);
}

template <int L>
static hipError_t col_pairs_forward_impl(const Geometry& g, const Twiddles& tw, const cd* U, cd* out, int lines,
                                         const PairLineArgs& pl, hipStream_t s) {
  using C = ColCfg<L>;
  dim3 grid(g.N2 / kCB, (lines + C::G - 1) / C::G);
  hipLaunchKernelGGL(k_col_pairs_forward<L>, grid, dim3(C::THREADS), 0, s, g, tw.coltw, tw.fourtw, U, out, lines,
                     pl);
  return hipGetLastError();
}
hipError_t launch_col_pairs_forward(const Geometry& g, const Twiddles& tw, const cd* U, cd* out, int lines,
                                    const PairLineArgs& pl, hipStream_t s) {
  MMP_DISPATCH_L(g.N1, col_pairs_forward_impl, g, tw, U, out, lines, pl, s);
}

template <int L>
static hipError_t row_forward_impl(const Geometry& g, const Twiddles& tw, cd* data, int lines, hipStream_t s) {
  using C = RowCfg<L>;
  const int rows = lines * g.N1;
  hipLaunchKernelGGL(k_row_forward<L>, dim3((rows + C::G - 1) / C::G), dim3(C::THREADS), 0, s, g, tw.rowtw, data,
                     lines);
  return hipGetLastError();
}
hipError_t launch_row_forward(const Geometry& g, const Twiddles& tw, cd* data, int lines, hipStream_t s) {
  MMP_DISPATCH_L(g.N2, row_forward_impl, g, tw, data, lines, s);
}

template <int L>
static hipError_t row_inverse_phase_impl(const Geometry& g, const Twiddles& tw, const cd* in, cd* out,
                                         const double* imL, const StepState* st, int use_phase, hipStream_t s) {
  using C = RowCfg<L>;
  const int rows = g.M * g.N1;
  hipLaunchKernelGGL(k_row_inverse_phase<L>, dim3((rows + C::G - 1) / C::G), dim3(C::THREADS), 0, s, g, tw.rowtw,
                     in, out, imL, st, use_phase);
  return hipGetLastError();
}
hipError_t launch_row_inverse_phase(const Geometry& g, const Twiddles& tw, const cd* in, cd* out,
                                    const double* imL, const StepState* st, int use_phase, hipStream_t s) {
  MMP_DISPATCH_L(g.N2, row_inverse_phase_impl, g, tw, in, out, imL, st, use_phase, s);
}

template <int L>
static hipError_t row_raman_impl(const Geometry& g, const Twiddles& tw, cd* data, int lines, const cd* hspec,
                                 double hscale, hipStream_t s) {
  using C = RowCfg<L>;
  const int rows = lines * g.N1;
  hipLaunchKernelGGL(k_row_raman<L>, dim3((rows + C::G - 1) / C::G), dim3(C::THREADS), 0, s, g, tw.rowtw, data,
                     lines, hspec, hscale);
  return hipGetLastError();
}
hipError_t launch_row_raman(const Geometry& g, const Twiddles& tw, cd* data, int lines, const cd* hspec,
                            double hscale, hipStream_t s) {
  MMP_DISPATCH_L(g.N2, row_raman_impl, g, tw, data, lines, hspec, hscale, s);
}

// Largest sample tile (a power of two dividing N, at most 64, keeping N / TS >= 256
// workgroups when possible) whose U and slot tiles fit the LDS capacity.
int gather_tile(const Geometry& g, int n_slots, int cap) {
  int ts = 64;
  while (ts > 8 && (g.N / ts < 256 || (g.M + n_slots) * ts > cap)) ts >>= 1;
  return ((g.M + n_slots) * ts <= cap) ? ts : 0;
}

hipError_t launch_gather(const Geometry& g, const cd* U, const cd* VR, cd* sigma, const GatherArgs& a,
                         hipStream_t s) {
  const int caps[3] = {2048, 4096, 8192};
  for (int ci = 0; ci < 3; ci++) {
    const int ts = gather_tile(g, a.n_slots, caps[ci]);
    if (!ts || g.N % ts) continue;
    dim3 grid(g.N / ts);
    switch (ci) {
      case 0: hipLaunchKernelGGL(k_gather<2048>, grid, dim3(256), 0, s, g, U, VR, sigma, a, ts); break;
      case 1: hipLaunchKernelGGL(k_gather<4096>, grid, dim3(256), 0, s, g, U, VR, sigma, a, ts); break;
      default: hipLaunchKernelGGL(k_gather<8192>, grid, dim3(256), 0, s, g, U, VR, sigma, a, ts); break;
    }
    return hipGetLastError();
  }
  return hipErrorInvalidValue;
}

hipError_t launch_advance_z(StepState* st, double h, hipStream_t s) {
  hipLaunchKernelGGL(k_advance_z, dim3(1), dim3(64), 0, s, st, h);
  return hipGetLastError();
}

hipError_t launch_energy(const Geometry& g, const cd* fields, double* out, double dt, hipStream_t s) {
  hipLaunchKernelGGL(k_energy, dim3(g.M), dim3(256), 0, s, g, fields, out, dt);
  return hipGetLastError();
}

}  // namespace mmp
