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

#include <hip/hip_ext.h>

#include <mutex>
#include <type_traits>

namespace mmp {

thread_local LaunchEvents g_launch_events;

// Every launch goes through hipExtLaunchKernelGGL so that, when g_launch_events is set, the
// dispatch itself records its start and stop (no separate event packets around the kernel).
#define MMP_LAUNCH(KERNEL, GRID, BLOCK, SHMEM, STREAM, ...)                                               \
  hipExtLaunchKernelGGL(KERNEL, GRID, BLOCK, SHMEM, STREAM, g_launch_events.start, g_launch_events.stop, 0, \
                        __VA_ARGS__)

// ---------------------------------------------------------------------------
// configuration of the row and column passes
// ---------------------------------------------------------------------------
template <int L>
struct RowCfg {
  static constexpr int T = L / 8;                  // lanes per row
  static constexpr int G = T >= 64 ? 1 : 64 / T;   // rows per workgroup
  static constexpr int THREADS = G * T;            // 64 (128 for rows of 1024)
  static constexpr int SLOTS = LineSlots<L, 1>::value;
};

// Column passes: a workgroup transforms CB neighbouring columns of G lines.
// CB = 8 reads whole 128 B lines; CB = 4 or 2 gives more workgroups for small
// problems (the LDS layout is then padded to keep lanes on distinct banks).
template <int L, int CB>
struct ColCfg {
  static constexpr int T = L / 8;                               // lanes per column line
  static constexpr int G = (CB * T >= 64) ? 1 : 64 / (CB * T);  // lines per workgroup
  static constexpr int THREADS = G * CB * T;
  static constexpr int PAD = CB < 8 ? 1 : 0;
  static constexpr int SLOTS = LineSlots<L, PAD>::value;        // per column
};

// Block ids are dealt round-robin to the 8 XCDs; remap them so that consecutive
// logical blocks (neighbouring columns) run on one XCD and share its L2.
__device__ __forceinline__ int xcd_block(int bid, int nb) {
  if (nb % 8) return bid;
  return (bid % 8) * (nb / 8) + bid / 8;
}

// Streaming store for arrays the next kernel reads (it runs on other XCDs, whose L2s do not
// share lines with this one).
template <class T>
__device__ __forceinline__ void store_stream(cpx<T>* p, cpx<T> v) {
  if constexpr (sizeof(T) == 4) {
    // one 8-byte store per FP32 value (p is 8-byte aligned: cpx<float> is)
    const unsigned long long bits =
        (unsigned long long)__float_as_uint(v.re) | ((unsigned long long)__float_as_uint(v.im) << 32);
    __builtin_nontemporal_store(bits, reinterpret_cast<unsigned long long*>(p));
  } else {
    __builtin_nontemporal_store(v.re, &p->re);
    __builtin_nontemporal_store(v.im, &p->im);
  }
}
// The same in FP64 only: FP32 stores of the column kernels measured slower streamed.
template <class T>
__device__ __forceinline__ void store_stream64(cpx<T>* p, cpx<T> v) {
  if constexpr (std::is_same<T, double>::value) store_stream(p, v);
  else *p = v;
}

__device__ __forceinline__ cd phase(double imL, double tau) {
  double s, c;
  sincos(imL * tau, &s, &c);
  return {c, s};
}

// ---------------------------------------------------------------------------
// row stage head
// ---------------------------------------------------------------------------
// E <- E * E1 with the reference's rounding (no FMA contraction), so stored
// tables stay bit-identical to src/MultimodePropagation_CPU.cpp:41.
__device__ __forceinline__ cd cmul_exact(cd a, cd b) {
  return {__dsub_rn(__dmul_rn(a.re, b.re), __dmul_rn(a.im, b.im)),
          __dadd_rn(__dmul_rn(a.re, b.im), __dmul_rn(a.im, b.re))};
}

// RE: Lawson factors computed from Im L and the fiber's z (HeadArgs::recompute_e) instead of
// read from the tables; they are formed where they are used, so no table registers stay live.
template <int L, class W, bool RE>
__global__ void __launch_bounds__(RowCfg<L>::THREADS) k_row_head(Geometry g, const cpx<W>* __restrict__ rowtw, HeadArgs a, Batch b) {
  using C = RowCfg<L>;
  using cw = cpx<W>;
  __shared__ cw smem[C::G * C::SLOTS];
  const int tid = threadIdx.x;
  const int gl = tid / C::T, t = tid % C::T;
  const int row = blockIdx.x * C::G + gl;
  const bool valid = row < g.M * g.N1;
  const size_t base = valid ? (size_t)row * g.N2 + t : t;
  cw* sm = smem + gl * C::SLOTS;
  const size_t MN = (size_t)g.M * g.N;
  // fiber blockIdx.y of the batch: its arrays and scalars
  const int fib = blockIdx.y;
  const FiberParams fp = b.fp[fib];
  if (!fp.live) return;  // uniform over the workgroup
  const double* __restrict__ hc = b.coef + ((size_t)fib * b.nheads + a.head) * kHeadCoef;  // uniform
  cd* __restrict__ psi_hat = a.psi_hat + fib * b.sf;
  cd* __restrict__ accp = a.acc + fib * b.sf;
  cd* __restrict__ stash = a.stash ? a.stash + fib * b.ss : nullptr;
  cd* __restrict__ e_prev = a.e_prev ? a.e_prev + fib * b.se : nullptr;
  const cd* __restrict__ e_cur = a.e_cur ? a.e_cur + fib * b.se : nullptr;
  const cd* __restrict__ e1p = a.e1 ? a.e1 + fib * b.sf : nullptr;
  const double* __restrict__ imLp = a.imL ? a.imL + fib * b.sf : nullptr;
  const cw* __restrict__ ysig = static_cast<const cw*>(a.ysig) + fib * b.sf;
  cw* __restrict__ yout = static_cast<cw*>(a.yout) + fib * b.sf;
  LineTw<L, W> lt;
  load_line_tw<L>(lt, rowtw, t);

  // issue every load this lane needs before the first barrier
  cw v[8];
  cd x[8], accb[8];
  cd ep[RE ? 1 : 8], ec[RE ? 1 : 8], e1[RE ? 1 : 8];  // Lawson tables (not RE)
  double lam[RE ? 8 : 1];                              // Im L (RE)
  const cd zero = {0., 0.};
  const bool want_cur = !a.final_stage || a.chain_next;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const size_t idx = base + i * C::T;
    if constexpr (RE) lam[i] = valid ? imLp[idx] : 0.;
    v[i] = (valid && a.has_prev) ? ysig[idx] : cw{0, 0};
    if constexpr (!RE) ep[i] = (valid && a.has_prev) ? e_prev[idx] : zero;
    const cd psi = (valid && (a.prev_is_first || !a.final_stage)) ? psi_hat[idx] : zero;
    const cd accin = (valid && a.has_prev && !a.prev_is_first) ? accp[idx] : zero;
    accb[i] = a.prev_is_first ? psi : accin;  // the step result so far
    x[i] = psi;                               // stage i's argument (not the final stage)
    if constexpr (!RE) {
      e1[i] = (valid && a.update_prev) ? e1p[idx] : zero;
      ec[i] = (valid && want_cur) ? e_cur[idx] : zero;
    }
  }
  // RE: E_{i-1} = exp(i Im L (n + c_{i-1}) h), E_i = exp(i Im L (n + c_i) h) with n the steps
  // taken before this one (one rounding of the distance, where z sums h step by step); the
  // close head's block 0 advances the count in this kernel, so it reads n from n0 (recorded by
  // stage head 1) and its chained next stage 0 is at n + 1
  // (both at one distance, as in a stage that repeats c or a close with c_{s-1} = 1: E_{i-1}
  // serves both, applied to the stage product as soon as that is formed)
  double tp = 0., tc = 0.;
  bool same = false;
  if constexpr (RE) {
    const double n = a.final_stage ? b.st[fib].n0 : (double)b.st[fib].steps;  // uniform
    tp = (n + a.c_prev) * fp.h;
    tc = ((a.final_stage ? n + 1.0 : n) + a.c_cur) * fp.h;
    same = a.has_prev && want_cur && tp == tc;
    if (a.set_n0 && blockIdx.x == 0 && threadIdx.x == 0) b.st[fib].n0 = n;
  }
  if (a.final_stage && a.advance_z && blockIdx.x == 0 && threadIdx.x == 0) {
    b.st[fib].z += fp.h;  // z += h (repeated FP64 addition)
    b.st[fib].steps += 1;
  }
  // stage i's argument psi + sum_j a_ij h N_j over j < i in increasing j: the stashed N_j now,
  // N_{i-1} (the largest j, from this head's transform) last
  const bool prev_term = a.n_a > 0 && a.a_idx[a.n_a - 1] == a.prev_stage;
  if (!a.final_stage) {
#pragma unroll
    for (int e = 0; e < kMaxStages; e++) {
      if (e >= a.n_a - (prev_term ? 1 : 0)) break;
      const int j = a.a_idx[e];
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const cd nj = valid ? stash[(size_t)j * MN + base + i * C::T] : zero;
        x[i] = cadd(x[i], cscale(nj, hc[1 + e]));
      }
    }
  }
  // the Lawson table of stage i-1 moves on by E1 (it is not read again in this step)
  if constexpr (!RE) {
    if (a.has_prev && a.update_prev && valid) {
#pragma unroll
      for (int i = 0; i < 8; i++) store_stream(e_prev + base + i * C::T, cmul_exact(ep[i], e1[i]));
    }
  }
  if (a.has_prev) {
    const double bh = hc[0];
    const double cprev = prev_term ? hc[a.n_a] : 0.;
    line_fft<L, -1, 1, 1>(v, sm, lt, t);
    bool finite = true;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const size_t idx = base + i * C::T;
      cd epi;
      if constexpr (RE) epi = phase(lam[i], tp);
      else epi = ep[i];
      const cd y = cmulc(widen(v[i]), epi);         // sigma_hat * conj(E_{i-1})
      const cd nh = {-y.im * fp.sgn, y.re * fp.sgn};  // * (i sgn) = N_{i-1}
      cd acc = accb[i];
      if (a.b_prev != 0.) acc = cadd(acc, cscale(nh, bh));
      finite = finite && __builtin_isfinite(acc.re) && __builtin_isfinite(acc.im);
      if (valid) {
        if (a.final_stage) store_stream(psi_hat + idx, acc);
        else store_stream(accp + idx, acc);
        if (a.stash_prev) store_stream(stash + (size_t)a.prev_stage * MN + idx, nh);
      }
      if (a.final_stage) x[i] = acc;  // chain_next: stage 0 of the next step from the new psi_hat
      else if (prev_term) x[i] = cadd(x[i], cscale(nh, cprev));
      if constexpr (RE) {
        if (same) x[i] = cmul(epi, x[i]);  // E_i x_i
      }
    }
    // a state that is no longer finite stays so: the fiber is parked (every lane that sees it
    // stores the same values), its later kernels skip it and its z and step count stop here
    if (a.final_stage && valid && !finite) {
      b.st[fib].halted = 1u;
      b.fp[fib].live = 0u;
    }
  }
  if (a.final_stage && !a.chain_next) return;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    cd ex;  // E_i x_i
    if constexpr (RE) ex = same ? x[i] : cmul(phase(lam[i], tc), x[i]);
    else ex = cmul(ec[i], x[i]);
    v[i] = narrow<W>(cscale(ex, a.scale_cur));
  }
  line_fft<L, 1, 1, 1>(v, sm, lt, t);
  if (valid) {
#pragma unroll
    for (int i = 0; i < 8; i++) store_stream(yout + base + i * C::T, v[i]);
  }
}

// ---------------------------------------------------------------------------
// column passes
// ---------------------------------------------------------------------------
struct ColIdx {
  int c, t, gl, cb, lg;
};

// lane -> (column c, line lane t, line gl) and block -> (column block cb, line group lg)
template <int L, int CB>
__device__ __forceinline__ ColIdx col_index(int ncb) {
  using C = ColCfg<L, CB>;
  const int tid = threadIdx.x;
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  return {tid % CB, (tid / CB) % C::T, tid / (CB * C::T), lb % ncb, lb / ncb};
}

// inverse column DFT: in = spectral-side rows (k1), out = time domain, times scale, divided by divisor
template <int L, int CB, class W>
__global__ void __launch_bounds__(256) k_col_inverse(Geometry g, const cpx<W>* __restrict__ coltw,
                                                     const cpx<W>* __restrict__ fourtw, const cpx<W>* in,
                                                     cpx<W>* out, int lines, double scale, double divisor) {
  using C = ColCfg<L, CB>;
  using cw = cpx<W>;
  __shared__ cw smem[C::G * C::SLOTS * CB];
  const ColIdx ix = col_index<L, CB>(g.N2 / CB);
  const int q = ix.lg * C::G + ix.gl;
  const int n2 = ix.cb * CB + ix.c;
  const bool valid = q < lines;
  const size_t base = (size_t)(valid ? q : 0) * g.N + n2;
  cw* sm = smem + ix.gl * C::SLOTS * CB + ix.c;
  LineTw<L, W> lt;
  load_line_tw<L>(lt, coltw, ix.t);
  cw v[8];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const int k1 = ix.t + i * C::T;
    v[i] = valid ? cmulc(in[base + (size_t)k1 * g.N2], fourtw[n2 * k1]) : cw{0, 0};
  }
  line_fft<L, 1, CB, C::PAD>(v, sm, lt, ix.t);
  if (valid) {
    const W sc = (W)scale, dv = (W)divisor;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int n1 = ix.t + i * C::T;
      cw x = cscale(v[i], sc);
      if (divisor != 1.0) x = {x.re / dv, x.im / dv};
      out[base + (size_t)n1 * g.N2] = x;
    }
  }
}

// forward column DFT + four-step twiddle: in = time domain, out = spectral-side rows
template <int L, int CB, class W>
__global__ void __launch_bounds__(256) k_col_forward(Geometry g, const cpx<W>* __restrict__ coltw,
                                                     const cpx<W>* __restrict__ fourtw, const cpx<W>* in,
                                                     cpx<W>* out, int lines) {
  using C = ColCfg<L, CB>;
  using cw = cpx<W>;
  __shared__ cw smem[C::G * C::SLOTS * CB];
  const ColIdx ix = col_index<L, CB>(g.N2 / CB);
  const int q = ix.lg * C::G + ix.gl;
  const int n2 = ix.cb * CB + ix.c;
  const bool valid = q < lines;
  const size_t base = (size_t)(valid ? q : 0) * g.N + n2;
  cw* sm = smem + ix.gl * C::SLOTS * CB + ix.c;
  LineTw<L, W> lt;
  load_line_tw<L>(lt, coltw, ix.t);
  cw v[8], w4[8];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const int n1 = ix.t + i * C::T;
    v[i] = valid ? in[base + (size_t)n1 * g.N2] : cw{0, 0};
    w4[i] = fourtw[n2 * n1];  // index k1 = n1 after the transform
  }
  line_fft<L, -1, CB, C::PAD>(v, sm, lt, ix.t);
  if (valid) {
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int k1 = ix.t + i * C::T;
      out[base + (size_t)k1 * g.N2] = cmul(v[i], w4[i]);
    }
  }
}

// forward column DFT of one real line (the Raman response), FP64 only
template <int L, int CB, class W>
__global__ void __launch_bounds__(256) k_col_forward_real(Geometry g, const cd* __restrict__ coltw,
                                                          const cd* __restrict__ fourtw, const double* __restrict__ in,
                                                          cd* __restrict__ out) {
  using C = ColCfg<L, CB>;
  __shared__ cd smem[C::G * C::SLOTS * CB];
  const ColIdx ix = col_index<L, CB>(g.N2 / CB);
  const bool valid = ix.lg * C::G + ix.gl < 1;
  const int n2 = ix.cb * CB + ix.c;
  cd* sm = smem + ix.gl * C::SLOTS * CB + ix.c;
  LineTw<L> lt;
  load_line_tw<L>(lt, coltw, ix.t);
  cd v[8];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const int n1 = ix.t + i * C::T;
    v[i] = valid ? cd{in[(size_t)n1 * g.N2 + n2], 0.} : cd{0., 0.};
  }
  line_fft<L, -1, CB, C::PAD>(v, sm, lt, ix.t);
  if (valid) {
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int k1 = ix.t + i * C::T;
      out[(size_t)k1 * g.N2 + n2] = cmul(v[i], fourtw[n2 * k1]);
    }
  }
}

// pair products of the time-domain fields U, then forward column DFT
template <int L, int CB, class W>
__global__ void __launch_bounds__(256) k_col_pairs_forward(Geometry g, const cpx<W>* __restrict__ coltw,
                                                           const cpx<W>* __restrict__ fourtw,
                                                           const cpx<W>* __restrict__ U, cpx<W>* __restrict__ out,
                                                           int lines, PairLineArgs pl) {
  using C = ColCfg<L, CB>;
  using cw = cpx<W>;
  __shared__ cw smem[C::G * C::SLOTS * CB];
  const ColIdx ix = col_index<L, CB>(g.N2 / CB);
  const int q = ix.lg * C::G + ix.gl;
  const int n2 = ix.cb * CB + ix.c;
  const bool valid = q < lines;
  const int qq = valid ? q : 0;
  const uint32_t la = pl.line_a[qq], lb = pl.line_b[qq], self = pl.line_self[qq];
  cw* sm = smem + ix.gl * C::SLOTS * CB + ix.c;
  LineTw<L, W> lt;
  load_line_tw<L>(lt, coltw, ix.t);
  cw v[8], w4[8];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const int n1 = ix.t + i * C::T;
    const size_t tt = (size_t)n1 * g.N2 + n2;
    cw x = {0, 0};
    if (valid) {
      const cw ua = U[(size_t)la * g.N + tt];
      if (self) {
        x.re = ua.re * ua.re + ua.im * ua.im;
        if (lb != 0xffffffffu) {
          const cw ub = U[(size_t)lb * g.N + tt];
          x.im = ub.re * ub.re + ub.im * ub.im;
        }
      } else {
        x = cmulc(ua, U[(size_t)lb * g.N + tt]);
      }
    }
    v[i] = x;
    w4[i] = fourtw[n2 * n1];
  }
  line_fft<L, -1, CB, C::PAD>(v, sm, lt, ix.t);
  if (valid) {
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int k1 = ix.t + i * C::T;
      out[(size_t)q * g.N + (size_t)k1 * g.N2 + n2] = cmul(v[i], w4[i]);
    }
  }
}

// ---------------------------------------------------------------------------
// row passes
// ---------------------------------------------------------------------------
template <int L>
__global__ void __launch_bounds__(RowCfg<L>::THREADS) k_row_forward(Geometry g, const cd* __restrict__ rowtw, cd* __restrict__ data,
                                                    int lines) {
  using C = RowCfg<L>;
  __shared__ cd smem[C::G * C::SLOTS];
  const int tid = threadIdx.x;
  const int gl = tid / C::T, t = tid % C::T;
  const int row = blockIdx.x * C::G + gl;
  const bool valid = row < lines * g.N1;
  const size_t base = valid ? (size_t)row * g.N2 : 0;  // rows of consecutive lines are contiguous
  cd* sm = smem + gl * C::SLOTS;
  LineTw<L> lt;
  load_line_tw<L>(lt, rowtw, t);
  cd v[8];
#pragma unroll
  for (int i = 0; i < 8; i++) v[i] = valid ? data[base + t + i * C::T] : cd{0., 0.};
  line_fft<L, -1, 1, 1>(v, sm, lt, t);
  if (valid) {
#pragma unroll
    for (int i = 0; i < 8; i++) data[base + t + i * C::T] = v[i];
  }
}

// optional phase exp(i ImL z) (the reference's exp(L z)), then inverse row DFT
template <int L>
__global__ void __launch_bounds__(RowCfg<L>::THREADS) k_row_inverse_phase(Geometry g, const cd* __restrict__ rowtw,
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
  LineTw<L> lt;
  load_line_tw<L>(lt, rowtw, t);
  cd v[8];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const size_t idx = base + t + i * C::T;
    cd x = valid ? in[idx] : cd{0., 0.};
    if (use_phase && valid) x = cmul(x, phase(imL[idx], z));
    v[i] = x;
  }
  line_fft<L, 1, 1, 1>(v, sm, lt, t);
  if (valid) {
#pragma unroll
    for (int i = 0; i < 8; i++) out[base + t + i * C::T] = v[i];
  }
}

// Raman convolution in the spectral rows: DFT, * H * hscale, inverse DFT (in place)
template <int L, class W>
__global__ void __launch_bounds__(RowCfg<L>::THREADS) k_row_raman(Geometry g, const cpx<W>* __restrict__ rowtw,
                                                  cpx<W>* __restrict__ data, int lines,
                                                  const cpx<W>* __restrict__ hspec, Batch b) {
  using C = RowCfg<L>;
  using cw = cpx<W>;
  __shared__ cw smem[C::G * C::SLOTS];
  const int fib = blockIdx.y;
  if (!b.fp[fib].live) return;  // uniform over the workgroup
  data += fib * b.sl;
  hspec += fib * b.sh;
  const double hscale = b.fp[fib].dt;
  const int tid = threadIdx.x;
  const int gl = tid / C::T, t = tid % C::T;
  const int row = blockIdx.x * C::G + gl;
  const bool valid = row < lines * g.N1;
  const size_t base = valid ? (size_t)row * g.N2 : 0;
  const size_t hbase = valid ? (size_t)(row % g.N1) * g.N2 : 0;
  cw* sm = smem + gl * C::SLOTS;
  LineTw<L, W> lt;
  load_line_tw<L>(lt, rowtw, t);
  cw v[8], hs[8];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    v[i] = valid ? data[base + t + i * C::T] : cw{0, 0};
    hs[i] = hspec[hbase + t + i * C::T];
  }
  line_fft<L, -1, 1, 1>(v, sm, lt, t);
  const W hsc = (W)hscale;
#pragma unroll
  for (int i = 0; i < 8; i++) v[i] = cscale(cmul(v[i], hs[i]), hsc);
  line_fft<L, 1, 1, 1>(v, sm, lt, t);
  if (valid) {
#pragma unroll
    for (int i = 0; i < 8; i++) data[base + t + i * C::T] = v[i];
  }
}

// Raman row pass of the column-stage step: the arithmetic of k_row_raman, with R consecutive
// rows of one line per workgroup (n1 .. n1 + R - 1), its lines read from `in` and stored to
// `out` in column order (sample (n1, n2) of a line at n2 * N1 + n1). The rows are staged in
// LDS as [n2][R + 1] on the way in and out, so that every load and store moves runs of R
// consecutive samples; the column kernels on either side (the pair pass writing the lines,
// the gather reading them) then move each column of a line as one contiguous run instead of
// N1 samples N2 apart. R is the
// largest of 16 (FP32) or 8 (FP64), halved as needed, that keeps the workgroup within 512
// threads and 50 KB of LDS: measured on one MI355X, wider workgroups lose more to occupancy
// than the longer runs gain (FP64 rows of 1024: R = 2 60 us, R = 4 83 us; FP32 rows of 1024:
// R = 4 21.7 us, R = 8 27.4 us; FP32 rows of 256: R = 16 111 us, R = 8 131 us).
template <int L, class W>
constexpr int raman_cols_rows() {
  constexpr int slots = LineSlots<L, 1>::value;
  int r = std::is_same<W, float>::value ? 16 : 8;
  while (r > 1) {
    const int lds = (r * slots > L * (r + 1) ? r * slots : L * (r + 1)) * (int)sizeof(cpx<W>);
    if (r * (L / 8) <= 512 && lds <= 51200) break;
    r /= 2;
  }
  return r;
}
template <int L, class W>
struct RamanColsCfg {
  static constexpr int T = L / 8;                              // lanes per row
  static constexpr int R = raman_cols_rows<L, W>();            // rows per workgroup
  static constexpr int THREADS = R * T;
  static constexpr bool WAVE = T <= 64;                        // each row inside one wavefront
  static constexpr int SLOTS = LineSlots<L, 1>::value;
  static constexpr int FFT = R * SLOTS;                        // transform rows
  static constexpr int STAGE = L * (R + 1);                    // [n2][R + 1] staging
  static constexpr int LDS = FFT > STAGE ? FFT : STAGE;
};

template <int L, class W>
__global__ void __launch_bounds__((RamanColsCfg<L, W>::THREADS)) k_row_raman_cols(Geometry g, const cpx<W>* __restrict__ rowtw,
                                                                            const cpx<W>* __restrict__ in,
                                                                            cpx<W>* __restrict__ out,
                                                                            const cpx<W>* __restrict__ hspec, Batch b,
                                                                            int in_cols) {
  using C = RamanColsCfg<L, W>;
  using cw = cpx<W>;
  __shared__ cw smem[C::LDS];
  const int fib = blockIdx.y;
  if (!b.fp[fib].live) return;  // uniform over the workgroup
  in += fib * b.sl;
  out += fib * b.sl;
  hspec += fib * b.sh;
  const double hscale = b.fp[fib].dt;
  const int tid = threadIdx.x;
  const int r = tid / C::T, t = tid % C::T;
  // R divides N1: the rows share one line. In FP32 neighbouring row groups run on one XCD, so
  // the parts of a cache line that several of them store meet in one L2 (measured: rows of 1024
  // 23.7 -> 21.7 us); FP64 runs of 8 rows already fill 128-byte lines, and there the plain
  // order measured slightly faster (rows of 256: 5.09 vs 5.21 us)
  // Where those FP32 runs fill whole 128-byte lines (R = 16) the blocks take the lines fastest
  // instead, so the blocks on one XCD share row groups and read only those rows of the Raman
  // spectrum (every XCD reading all of it fetched the spectrum 8 times from HBM)
  const int blk = std::is_same<W, float>::value ? xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x;
  int row0 = blk * C::R;
  if constexpr (std::is_same<W, float>::value && C::R * sizeof(cw) >= 128) {
    const int nl = (int)gridDim.x * C::R / g.N1;  // lines of the launch
    if (gridDim.x % 8 == 0) row0 = (blk % nl) * g.N1 + (blk / nl) * C::R;
  }
  const int line = row0 / g.N1, n10 = row0 % g.N1, n1 = n10 + r;
  const size_t cbase = (size_t)line * g.N + n10;  // the R rows' first sample in column order
  const size_t hbase = (size_t)n1 * g.N2;
  cw* sm = smem + r * C::SLOTS;
  cw v[8], hs[8];
  if (in_cols) {  // uniform branch: lines in column order, staged through LDS
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const int e = tid + k * C::THREADS;  // R * L = 8 * THREADS samples
      const int rr = e % C::R, n2 = e / C::R;
      smem[n2 * (C::R + 1) + rr] = in[cbase + (size_t)n2 * g.N1 + rr];
    }
  } else {  // lines in row order: each row read directly
    const size_t rbase = (size_t)line * g.N + (size_t)n1 * g.N2;
#pragma unroll
    for (int i = 0; i < 8; i++) v[i] = in[rbase + t + i * C::T];
  }
#pragma unroll
  for (int i = 0; i < 8; i++) hs[i] = hspec[hbase + t + i * C::T];
  LineTw<L, W> lt;
  load_line_tw<L>(lt, rowtw, t);
  if (in_cols) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 8; i++) v[i] = smem[(t + i * C::T) * (C::R + 1) + r];
    __syncthreads();  // the staging area becomes the rows' transform slots
  }
  line_fft<L, -1, 1, 1, W, C::WAVE>(v, sm, lt, t);
  const W hsc = (W)hscale;
#pragma unroll
  for (int i = 0; i < 8; i++) v[i] = cscale(cmul(v[i], hs[i]), hsc);
  line_fft<L, 1, 1, 1, W, C::WAVE>(v, sm, lt, t);
  __syncthreads();  // every row's transform is done with its LDS row before the staging reuses it
#pragma unroll
  for (int i = 0; i < 8; i++) smem[(t + i * C::T) * (C::R + 1) + r] = v[i];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const int e = tid + k * C::THREADS;  // R * L = 8 * THREADS staged samples
    const int rr = e % C::R, n2 = e / C::R;
    if constexpr (C::R * sizeof(cw) >= 128) store_stream(out + cbase + (size_t)n2 * g.N1 + rr, smem[n2 * (C::R + 1) + rr]);
    else out[cbase + (size_t)n2 * g.N1 + rr] = smem[n2 * (C::R + 1) + rr];
  }
}

// ---------------------------------------------------------------------------
// Kerr + Raman gather over a tile of 64 consecutive samples:
//   X_slot = kv * U_a conj(U_b) + kr * VR_slot        (slot mix, into LDS)
//   sigma_p = sum over the nonzeros (k, slot, q) of Q^(p):  q * X_slot * U_k
// One wavefront per mode p (lanes = samples), so the nonzero loop and its
// metadata loads are uniform; nonzeros are sorted by slot so consecutive ones
// reuse X. Slots are staged through LDS in chunks when M + P exceeds the tile.
// ---------------------------------------------------------------------------
// sum += q_e * X[slot_e - s0] * U[k_e] over the nonzeros [e0, e1) for this lane's sample.
// e0, e1 and s0 are wave-uniform. Xs and Us point at the lane's sample in LDS rows of
// `stride` values. Lane l loads nonzero eb + l; v_readlane broadcasts it, eight at a
// time, so the sixteen LDS reads of a batch are in flight together. The products are
// added in nonzero order.
// Broadcast lane j's coefficient (kept as its bit pattern in one or two ints) to the wave.
struct QBits {
  int lo, hi;
};
__device__ __forceinline__ QBits q_bits(double v) { return {__double2loint(v), __double2hiint(v)}; }
__device__ __forceinline__ QBits q_bits(float v) { return {__float_as_int(v), 0}; }
template <class W>
__device__ __forceinline__ W q_lane(QBits b, int j) {
  if constexpr (sizeof(W) == 8)
    return __hiloint2double(__builtin_amdgcn_readlane(b.hi, j), __builtin_amdgcn_readlane(b.lo, j));
  else
    return __int_as_float(__builtin_amdgcn_readlane(b.lo, j));
}

template <class W>
__device__ __forceinline__ cpx<W> gather_sum(cpx<W> sum, const uint32_t* __restrict__ kslot, const W* __restrict__ q,
                                             uint32_t e0, uint32_t e1, int s0, const cpx<W>* Xs, const cpx<W>* Us,
                                             int stride, int lane) {
  constexpr int B = 8;
  for (uint32_t eb = e0; eb < e1; eb += 64) {
    const uint32_t mine = eb + lane;
    const uint32_t ks = mine < e1 ? kslot[mine] : 0u;
    const QBits qb = q_bits(mine < e1 ? q[mine] : (W)0);
    const int cnt = (int)min(64u, e1 - eb);
    int j = 0;
    for (; j + B <= cnt; j += B) {
      cpx<W> xs[B], us[B];
      W qs[B];
#pragma unroll
      for (int u = 0; u < B; u++) {
        const uint32_t ksj = (uint32_t)__builtin_amdgcn_readlane((int)ks, j + u);
        qs[u] = q_lane<W>(qb, j + u);
        xs[u] = Xs[((int)(ksj >> 8) - s0) * stride];
        us[u] = Us[(int)(ksj & 255u) * stride];
      }
#pragma unroll
      for (int u = 0; u < B; u++) sum = cadd(sum, cscale(cmul(xs[u], us[u]), qs[u]));
    }
    for (; j < cnt; j++) {
      const uint32_t ksj = (uint32_t)__builtin_amdgcn_readlane((int)ks, j);
      const W qj = q_lane<W>(qb, j);
      sum = cadd(sum, cscale(cmul(Xs[((int)(ksj >> 8) - s0) * stride], Us[(int)(ksj & 255u) * stride]), qj));
    }
  }
  return sum;
}

// Slot mixes of a 64-sample tile into LDS: X[s - s0][j] = kv U_a conj(U_b) + kr VR_s[t0 + j]
// for slots [s0, s0 + ns). Every Raman value a lane needs is loaded before any is used, so
// the loads overlap instead of each waiting for the previous one.
template <class W>
__device__ __forceinline__ void slot_mix_tile(const GatherArgs& a, const Geometry& g, const cpx<W>* __restrict__ VR,
                                              const cpx<W>* Ush, cpx<W>* Xsh, int s0, int ns, int t0, int tid,
                                              int NTHR) {
  using cw = cpx<W>;
  constexpr int TS = 64, R = 16;
  const int total = ns * TS;
  const W kv = (W)a.kv, kr = (W)a.kr;
  for (int base = 0; base < total; base += R * NTHR) {
    cw r[R];
#pragma unroll
    for (int i = 0; i < R; i++) {
      const int e = base + tid + i * NTHR;
      r[i] = {0, 0};
      // e / 64 is the same for the whole wavefront (e = base + 64 * wave + lane + i * NTHR)
      const int sl = __builtin_amdgcn_readfirstlane(s0 + e / TS);
      if (a.raman && e < total) r[i] = VR[(size_t)a.slot_line[sl] * g.N + t0 + e % TS];
    }
#pragma unroll
    for (int i = 0; i < R; i++) {
      const int e = base + tid + i * NTHR;
      if (e >= total) break;
      const int sl = __builtin_amdgcn_readfirstlane(s0 + e / TS), j = e % TS;
      const cw V = cmulc(Ush[a.slot_a[sl] * TS + j], Ush[a.slot_b[sl] * TS + j]);
      cw X = cscale(V, kv);
      if (a.raman) {
        cw vr;
        switch (a.slot_kind[sl]) {
          case kDirect: vr = r[i]; break;
          case kConj: vr = cconj(r[i]); break;
          case kRe: vr = {r[i].re, 0}; break;
          default: vr = {r[i].im, 0}; break;
        }
        X = cadd(X, cscale(vr, kr));
      }
      Xsh[e] = X;
    }
  }
}

constexpr int kGatherTS = 64;
constexpr int kGatherMaxPW = 8;  // modes per wavefront

void gather_plan(int M, int P, int& cap, int& chunk) {
  cap = 0;
  chunk = 0;
  const int caps[2] = {4096, 8192};
  for (int c : caps) {
    const int room = c / kGatherTS - M;
    if (room <= 0) continue;
    if (room >= P || c == 8192) {
      cap = c;
      chunk = room < P ? room : P;
      return;
    }
  }
}

template <int CAP, class W>
__global__ void __launch_bounds__(512) k_gather(Geometry g, const cpx<W>* __restrict__ U,
                                                const cpx<W>* __restrict__ VR, cpx<W>* __restrict__ sigma,
                                                GatherArgs a) {
  using cw = cpx<W>;
  __shared__ cw sh[CAP];
  constexpr int TS = kGatherTS;
  const int M = g.M;
  const W* __restrict__ qv = static_cast<const W*>(a.q);
  cw* Ush = sh;            // [M][TS]
  cw* Xsh = sh + M * TS;   // [chunk][TS]
  const int t0 = blockIdx.x * TS;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int lane = tid % TS, nwaves = nthr / TS;
  const int wave = __builtin_amdgcn_readfirstlane(tid / TS);  // wave-uniform for the compiler
#pragma unroll 4
  for (int e = tid; e < M * TS; e += nthr) Ush[e] = U[(size_t)(e / TS) * g.N + t0 + e % TS];
  cw acc[kGatherMaxPW];
#pragma unroll
  for (int r = 0; r < kGatherMaxPW; r++) acc[r] = {0, 0};
  for (int c = 0; c < a.n_chunks; c++) {
    const int s0 = c * a.chunk;
    const int ns = min(a.chunk, a.n_slots - s0);
    __syncthreads();  // Ush loaded / previous chunk consumed
    slot_mix_tile(a, g, VR, Ush, Xsh, s0, ns, t0, tid, nthr);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kGatherMaxPW; r++) {
      const int p = wave + r * nwaves;
      if (p >= M) break;
      const uint32_t* co = a.coff + (size_t)p * (a.n_chunks + 1) + c;
      acc[r] = gather_sum(acc[r], a.kslot, qv, co[0], co[1], s0, Xsh + lane, Ush + lane, TS, lane);
    }
  }
#pragma unroll
  for (int r = 0; r < kGatherMaxPW; r++) {
    const int p = wave + r * nwaves;
    if (p >= M) break;
    sigma[(size_t)p * g.N + t0 + lane] = acc[r];
  }
}

// ---------------------------------------------------------------------------
// Column-stage kernels (Raman on, N1 = 64, M <= 16). One workgroup of 256 (or 512) lanes
// owns one column n2 and does every column-side step of a stage there, with the
// column of each field and Raman line held in LDS between the steps:
//   C1  inverse column DFT of the M fields -> U (HBM, column order, and LDS) -> pair
//       products -> forward column DFT of the Raman lines -> conv (row order, or column
//       order on large grids; the Raman row pass reads them so and writes column order)
//   C2  inverse column DFT of the Raman lines (column order; kept in LDS) -> dense Kerr +
//       Raman gather of the 64 samples, forming each slot mix X_s on the way -> forward
//       column DFT of sigma -> yout
// Column order: sample (n1, n2) of an array at n2 * N1 + n1, so that a column is one
// contiguous run of N1 samples.
// Lanes form 32 line slots of 8 lanes; a round transforms up to 32 lines.
// ---------------------------------------------------------------------------
constexpr int kColStageL = 64;       // column length the column-stage kernels take
constexpr int kColStageMaxM = 16;    // modes kept in registers by the gather
// LDS limits of one chunk, by mode count: a few modes have few lines and small coupling
// blocks, and a small LDS footprint lets more workgroups share a CU (batched fibers)
constexpr int colstage_lines_cap(int M) { return M <= 4 ? 16 : (M <= 6 ? 24 : 40); }  // Raman lines
// coupling coefficients held in LDS (64 KB FP64 would crowd out the lines)
constexpr int colstage_qcap(int M, bool mixed) {
  return M <= 4 ? 1024 : (M <= 6 ? 2048 : ((mixed && M > 10) ? 8192 : 4096));
}
template <int M, class W>
constexpr int kColStageQCapOf = colstage_qcap(M, std::is_same<W, float>::value);
// of those, prefetched into registers before the chunk's barrier (the rest go straight to LDS):
// what a small tensor needs in all (up to 4 modes: 512, up to 6: 1024, parity-compact), 4096
// where the registers allow it (FP32, 7 to 12 modes), else 2048
template <int M, class W>
constexpr int kColStageQPreOf =
    M <= 4 ? 512 : (M <= 6 ? 1024 : ((std::is_same<W, float>::value && M <= 12) ? 4096 : 2048));
constexpr int kColStageSCap = 128;   // slots of one chunk
constexpr int kColStageRS = LineSlots<kColStageL, 1>::value;  // LDS row stride of a line (72)
// LDS rows for the lines of a chunk: every wavefront of a round transforms 8 line slots, so
// the rows cover the chunk's lines rounded up to 8
constexpr int colstage_rows(int M) { return (colstage_lines_cap(M) + 7) / 8 * 8; }

struct ColStageCfg {
  static constexpr int L = kColStageL;
  static constexpr int T = L / 8;                      // 8 lanes per line
  static constexpr int NL = 256 / T;                   // 32 line slots per round
  static constexpr int SLOTS = LineSlots<L, 1>::value;  // 72
  static constexpr int FFT = NL * SLOTS;               // 2304
};

int colstage_max_modes() { return kColStageMaxM; }

// 8-wavefront gather workgroups: up to 12 modes in FP32, 6 in FP64 (beyond, the registers of
// a 512-lane workgroup would spill)
template <int M, class W>
constexpr bool kColStageWide = std::is_same<W, float>::value ? M <= 12 : M <= 6;
bool colstage_wide_ok(int M, bool mixed) { return mixed ? M <= 12 : M <= 6; }
int colstage_lines_per_chunk(int M) { return colstage_lines_cap(M); }
int colstage_coefficients_per_chunk(int M, bool mixed) { return colstage_qcap(M, mixed); }
int colstage_slots_per_chunk() { return kColStageSCap; }

// C1 over CW neighbouring columns per workgroup, the column block's Raman lines dealt to
// `groups` workgroups (each transforms every field of the block, group 0 stores them to U).
// Lanes are ordered column-fastest, tid = (row_job * 8 + t) * CW + c, so every global access
// of a wavefront reads or writes runs of CW consecutive samples of one row: with one column
// per workgroup each lane's sample sits in a cache line of its own and the pass is bound by
// the L2 request rate, not by HBM. The arithmetic of every line is the same for any CW and
// groups, so the results are too.
template <class W, int CW>
__global__ void __launch_bounds__(256) k_colstage_fields_pairs(Geometry g, const cpx<W>* __restrict__ coltw,
                                                               const cpx<W>* __restrict__ fourtw,
                                                               const cpx<W>* __restrict__ yin, cpx<W>* __restrict__ U,
                                                               cpx<W>* __restrict__ conv, int n_lines,
                                                               PairLineArgs pl, double scale, Batch b, int groups,
                                                               int conv_cols) {
  using F = ColStageCfg;
  using cw = cpx<W>;
  constexpr int L = F::L;
  constexpr int JL = F::NL / CW;  // field or line rows per round: 32 jobs = JL rows x CW columns
  static_assert(CW == 1 || CW == 2 || CW == 4, "columns per workgroup");
  extern __shared__ __align__(16) unsigned char c1_lds[];
  cw* sh = reinterpret_cast<cw*>(c1_lds);  // [32 jobs][SLOTS] transform rows, then [M][L][CW] fields
  const int fib = blockIdx.y;
  if (!b.fp[fib].live) return;  // uniform over the workgroup
  yin += fib * b.sf;
  U += fib * b.sf;
  conv += fib * b.sl;
  cw* Ush = sh + F::FFT;
  const int tid = threadIdx.x;
  const int c = tid % CW, t = (tid / CW) % F::T, jl = tid / (F::T * CW);
  const int blk = xcd_block(blockIdx.x, gridDim.x);  // the groups of a column block are neighbours
  const int grp = blk % groups;
  const int n2 = (blk / groups) * CW + c;
  const int lpg = (n_lines + groups - 1) / groups;
  const int lbeg = grp * lpg < n_lines ? grp * lpg : n_lines;
  const int lend = lbeg + lpg < n_lines ? lbeg + lpg : n_lines;
  cw* fsm = sh + (jl * CW + c) * F::SLOTS;
  cw w4[8], v[8];
  // the first round's field samples are requested before the twiddles, and the first round's
  // line descriptors with them (each later round's during the round before)
  {
    const bool valid = jl < g.M;
    const size_t base = (size_t)(valid ? jl : 0) * g.N + n2;
#pragma unroll
    for (int i = 0; i < 8; i++) v[i] = valid ? yin[base + (size_t)(t + i * F::T) * g.N2] : cw{0, 0};
  }
  uint32_t la_n = 0, lb_n = 0, self_n = 0;
  if (lbeg + jl < lend) {
    la_n = pl.line_a[lbeg + jl];
    lb_n = pl.line_b[lbeg + jl];
    self_n = pl.line_self[lbeg + jl];
  }
  LineTw<L, W> lt;
  load_line_tw<L>(lt, coltw, t);
#pragma unroll
  for (int i = 0; i < 8; i++) w4[i] = fourtw[n2 * (t + i * F::T)];
  const W sc = (W)scale;
  const int wave_jl0 = (tid / 64) * (64 / (F::T * CW));  // first row job of this wavefront
  for (int r0 = 0; r0 < g.M; r0 += JL) {
    const int q = r0 + jl;
    const bool valid = q < g.M;
    const size_t base = (size_t)(valid ? q : 0) * g.N + n2;
    if (r0 + wave_jl0 >= g.M) continue;  // no field of this wavefront in the round
    if (r0 > 0) {
#pragma unroll
      for (int i = 0; i < 8; i++) v[i] = valid ? yin[base + (size_t)(t + i * F::T) * g.N2] : cw{0, 0};
    }
#pragma unroll
    for (int i = 0; i < 8; i++) v[i] = cmulc(v[i], w4[i]);
    line_fft<L, 1, 1, 1, W, true>(v, fsm, lt, t);
    if (valid) {
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const int n1 = t + i * F::T;
        const cw x = cscale(v[i], sc);
        if (grp == 0) store_stream64(U + (size_t)q * g.N + (size_t)n2 * L + n1, x);  // column order, for the gather
        Ush[(q * L + n1) * CW + c] = x;
      }
    }
  }
  __syncthreads();
  for (int r0 = lbeg; r0 < lend; r0 += JL) {
    const int q = r0 + jl;
    const bool valid = q < lend;
    if (r0 + wave_jl0 >= lend) continue;  // no line of this wavefront in this round or later
    // cross line: U_a conj(U_b); self line: |U_a|^2 + i |U_b|^2 (b absent: 0)
    const uint32_t la = la_n, lb = lb_n, self = self_n;  // lanes past the last line read line 0's
    if (q + JL < lend) {
      la_n = pl.line_a[q + JL];
      lb_n = pl.line_b[q + JL];
      self_n = pl.line_self[q + JL];
    }
    const bool has_b = lb != 0xffffffffu;
    const int ia = (int)la * L * CW + c, ib = (int)(has_b ? lb : la) * L * CW + c;
    const W bmask = has_b ? (W)1 : (W)0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int n1 = t + i * F::T;
      const cw ua = Ush[ia + n1 * CW], ub = Ush[ib + n1 * CW];
      const cw cross = cmulc(ua, ub);
      const cw selfv = {ua.re * ua.re + ua.im * ua.im, bmask * (ub.re * ub.re + ub.im * ub.im)};
      v[i] = !valid ? cw{0, 0} : (self ? selfv : cross);
    }
    line_fft<L, -1, 1, 1, W, true>(v, fsm, lt, t);
    if (valid) {
      if (conv_cols) {  // column order (uniform branch)
        const size_t base = (size_t)q * g.N + (size_t)n2 * L;
#pragma unroll
        for (int i = 0; i < 8; i++) store_stream64(conv + base + t + i * F::T, cmul(v[i], w4[i]));
      } else {
        const size_t base = (size_t)q * g.N + n2;
#pragma unroll
        for (int i = 0; i < 8; i++) store_stream64(conv + base + (size_t)(t + i * F::T) * g.N2, cmul(v[i], w4[i]));
      }
    }
  }
}

// 4-wavefront FP32 gathers of 7 to 10 modes also come in a variant held to 168 VGPRs, so
// that three workgroups share a CU (their LDS is under 50 KB); it reads the coupling rows in
// 64-byte groups to make room, and spills at most 12 dwords. Measured on one MI355X it pays
// for large grids only: batched M = 10, NT = 16384 x 36 fibers 563 -> 480 us, M = 8 x 16
// fibers 125 -> 102 us; one fiber at NT = 65536 (1024 workgroups) 65 -> 67 us. Beyond 10
// modes (and without parity blocks at 10) the bound spills hundreds of bytes per lane.
template <int M, class W, bool PAR>
constexpr bool kColStageHiOcc = std::is_same<W, float>::value && M >= 7 && (M <= 9 || (M == 10 && PAR));
bool colstage_hi_occupancy_ok(int M, bool mixed, bool parity) { return mixed && M >= 7 && (M <= 9 || (M == 10 && parity)); }

// The slot sums on the matrix cores (MF, FP32): per sample, B[p][k] = sum over slots s of
// Qd[s][p][k] X_s is a real-by-complex matrix product with the rows (p, k) of one coupling parity
// class, the slots of that class as the inner dimension and the samples (real and imaginary parts
// as two products) as the columns; sigma_p = sum_k U_k B[p][k] follows per sample. One
// v_mfma_f32_32x32x2_f32 takes 32 rows, two slots and 32 samples. Rows are ordered so that the
// two lane halves of a 32 x 32 result hold the same modes p: the row of (p, kk) is
// 8 (u / 4) + 4 (kk % 2) + u % 4 with u = p * KHH + kk / 2 (KHH = ceil(QROW / 2), QROW the
// coefficients per row p), so an accumulator register r of tile t holds u = 16 t + r in both
// halves, and only k depends on the half. Tiles: ceil(M * KHH / 16), up to four. With one or two
// tiles each wavefront holds all of them (32 or 64 accumulator registers per lane) and takes every
// other pair of slots; with three or four the wavefront pairs split the tiles (two each) and every
// wavefront takes every pair, forming its slots' X twice over the workgroup.
template <int M, bool PAR>
struct MfmaRows {
  static constexpr int QROW = PAR ? (M + 1) / 2 : M;
  static constexpr int KHH = (QROW + 1) / 2;
  static constexpr int RT = (M * KHH + 15) / 16;
  static constexpr bool TSPLIT = RT > 2;      // the wavefront pairs split the tiles
  static constexpr int TPW = TSPLIT ? 2 : RT;  // tiles per wavefront
};
template <int M, class W, bool PAR>
constexpr bool kColStageMfma = std::is_same<W, float>::value && M >= 4 && MfmaRows<M, PAR>::RT <= 4;
bool colstage_matrix_cores_ok(int M, bool mixed, bool parity) {
  if (!mixed || M < 4 || M > kColStageMaxM) return false;
  const int qrow = parity ? (M + 1) / 2 : M, khh = (qrow + 1) / 2;
  return (M * khh + 15) / 16 <= 4;
}
typedef float mfma_acc_t __attribute__((ext_vector_type(16)));

// NW wavefronts (4, or 8 when M <= 6); each line transform runs inside one wavefront,
// in place in its padded LDS row, and wavefronts without a line skip it. With MF (4 wavefronts,
// FP32) the slot sums run on the matrix cores: wavefront w takes the samples 32 (w % 2) .. + 31
// and, with up to two 32-row tiles, every other pair of its parity class's slots (from pair
// w / 2); with three or four tiles, tiles 2 (w / 2) and 2 (w / 2) + 1 over every pair.
template <int M, int NW, class W, bool PAR, bool HI = false, bool MF = false>
__global__ void __launch_bounds__(NW * 64, HI ? 3 : (MF ? (MfmaRows<M, PAR>::RT == 1 ? 3 : 2) : 1)) k_colstage_raman_gather(Geometry g, const cpx<W>* __restrict__ coltw,
                                                                   const cpx<W>* __restrict__ fourtw,
                                                                   const cpx<W>* __restrict__ U,
                                                                   const cpx<W>* __restrict__ conv,
                                                                   cpx<W>* __restrict__ yout, GatherArgs a,
                                                                   FusedArgs fa, const W* __restrict__ Qd,
                                                                   double scale, Batch b) {
  using cw = cpx<W>;
  const int fib = blockIdx.y;
  const FiberParams fp = b.fp[fib];
  if (!fp.live) return;  // uniform over the workgroup
  U += fib * b.sf;
  yout += fib * b.sf;
  conv += fib * b.sl;
  Qd += fib * b.sq;
  constexpr int L = kColStageL, T = L / 8, RS = kColStageRS;
  constexpr int NT = NW * 64;
  constexpr int JOBS = NT / T;  // line slots per round
  constexpr int KH = (M + 1) / 2;
  constexpr int MM = PAR ? M * KH : M * M;  // coupling coefficients per slot
  constexpr int UR = (M + 7) / 8 * 8;  // field rows, rounded up to whole wavefronts of lines
  constexpr int UIN = (M * L + NT - 1) / NT;
  // Raman line rows of a chunk, or the NW partial sums of M modes after the gather
  constexpr int RCAP = colstage_rows(M) * RS > NW * M * L ? colstage_rows(M) * RS : NW * M * L;
  static_assert(M >= 1 && M <= kColStageMaxM, "modes out of range");
  static_assert(!MF || (NW == 4 && !HI && kColStageMfma<M, W, PAR>), "matrix-core gather: FP32, 4 wavefronts");
  __shared__ cw Ush[UR * RS];   // field rows; sigma rows after the gather
  __shared__ cw Rsh[RCAP];      // Raman line rows of the chunk; partial sums after the gather
  __shared__ W Qsh[kColStageQCapOf<M, W>];  // [slots of the chunk][M][M] coupling blocks
  __shared__ uint32_t Msh[kColStageSCap];  // slot metadata of the chunk
  // MF: the chunk's slots of each parity class in order (metadata, offset in the chunk), and
  // their counts
  __shared__ uint2 Lsh[MF ? 2 * kColStageSCap : 1];
  __shared__ int Nsh[2];
  const int tid = threadIdx.x;
  const int job = tid / T, t = tid % T;
  const int lane = tid % 64;
  const int wave = __builtin_amdgcn_readfirstlane(tid / 64);
  const int n2 = xcd_block(blockIdx.x, gridDim.x);
  // the first chunk's lines, this lane's field samples, coupling blocks and slot metadata are
  // all requested before anything waits on memory
  cw v[8], uin[UIN];
  int l0 = 0, l1 = (int)fa.chunk_line[1], s0 = 0, s1 = (int)fa.chunk_slot[1];
  {
    const bool valid = job < l1;
    const size_t base = (size_t)(valid ? job : 0) * g.N + (size_t)n2 * L;
#pragma unroll
    for (int i = 0; i < 8; i++) v[i] = valid ? conv[base + t + i * T] : cw{0, 0};
#pragma unroll
    for (int r = 0; r < UIN; r++) {
      const int e = tid + r * NT;
      uin[r] = e < M * L ? U[(size_t)(e / L) * g.N + (size_t)n2 * L + e % L] : cw{0, 0};
    }
  }
  constexpr int QIN = kColStageQPreOf<M, W> / NT;
  W qin[QIN];
  uint32_t min_ = 0;
#pragma unroll
  for (int r = 0; r < QIN; r++) {
    const int e = tid + r * NT;
    qin[r] = e < (s1 - s0) * MM ? Qd[(size_t)s0 * MM + e] : (W)0;
  }
  if (tid < s1 - s0) min_ = fa.slot_meta[s0 + tid];
  LineTw<L, W> lt;
  load_line_tw<L>(lt, coltw, t);
  cw w4[8];
#pragma unroll
  for (int i = 0; i < 8; i++) w4[i] = fourtw[n2 * (t + i * T)];
#pragma unroll
  for (int r = 0; r < UIN; r++) {
    const int e = tid + r * NT;
    if (e < M * L) Ush[(e / L) * RS + e % L] = uin[r];
  }
  const W sc = (W)scale, kv = (W)fp.kv, kr = (W)fp.kr;
  cw u[MF ? 1 : M], acc[M];
#pragma unroll
  for (int p = 0; p < M; p++) acc[p] = {0, 0};
  // MF: this lane's sample and lane half, and the coupling offset of its A-operand row in each
  // tile and parity class (-1: a padding row, zero)
  const int hh = lane >> 5;
  const int smp = MF ? 32 * (wave & 1) + (lane & 31) : lane;
  using MR = MfmaRows<M, PAR>;
  const int tg = MR::TSPLIT ? (wave >> 1) : 0;  // this wavefront's tile group (tiles TPW tg .. TPW tg + TPW - 1)
  int aoff[2][MF ? MR::TPW : 1];
  if constexpr (MF) {
#pragma unroll
    for (int cc = 0; cc < 2; cc++)
#pragma unroll
      for (int j = 0; j < MR::TPW; j++) {
        const int t = MR::TPW * tg + j;
        const int R = 32 * t + (lane & 31);
        const int uu = 4 * (R >> 3) + (R & 3);
        const int p = uu / MR::KHH, kk = 2 * (uu % MR::KHH) + ((R >> 2) & 1);
        const int k = PAR ? 2 * kk + ((p + cc) & 1) : kk;
        aoff[cc][j] = (t < MR::RT && p < M && kk < MR::QROW && k < M) ? p * MR::QROW + kk : -1;
      }
  }
  for (int c = 0; c < fa.n_chunks; c++) {
    if (c > 0) {
      __syncthreads();  // Rsh, Qsh and Msh of the previous chunk consumed
      l0 = l1;
      l1 = (int)fa.chunk_line[c + 1];
      s0 = s1;
      s1 = (int)fa.chunk_slot[c + 1];
#pragma unroll
      for (int r = 0; r < QIN; r++) {
        const int e = tid + r * NT;
        qin[r] = e < (s1 - s0) * MM ? Qd[(size_t)s0 * MM + e] : (W)0;
      }
      if (tid < s1 - s0) min_ = fa.slot_meta[s0 + tid];
    }
#pragma unroll
    for (int r = 0; r < QIN; r++) Qsh[tid + r * NT] = qin[r];
    for (int e = QIN * NT + tid; e < (s1 - s0) * MM; e += NT) Qsh[e] = Qd[(size_t)s0 * MM + e];
    if (tid < s1 - s0) Msh[tid] = min_;
    if constexpr (MF) {
      // the class lists: wavefronts 0 and 1 hold the chunk's slots 0..63 and 64..127 (at most
      // kColStageSCap = 128); wavefront 1 also reads the metadata of slots 0..63 for their counts
      if (tid < 128) {
        const int ns = s1 - s0, ln = tid % 64;
        const bool has = tid < ns;
        const int cls = PAR ? (int)((min_ >> 18) & 1u) : 0;
        const uint64_t own1 = __ballot(has && cls == 1), own0 = __ballot(has && cls == 0);
        uint64_t lo1 = own1, lo0 = own0;  // class masks of slots 0..63
        if (tid >= 64) {
          const bool has_lo = ln < ns;
          const uint32_t m = has_lo ? fa.slot_meta[s0 + ln] : 0u;
          const int cl = PAR ? (int)((m >> 18) & 1u) : 0;
          lo1 = __ballot(has_lo && cl == 1);
          lo0 = __ballot(has_lo && cl == 0);
        }
        const uint64_t below = ln ? (~0ull >> (64 - ln)) : 0ull;
        const int pos = cls ? (tid >= 64 ? __popcll(lo1) : 0) + __popcll(own1 & below)
                            : (tid >= 64 ? __popcll(lo0) : 0) + __popcll(own0 & below);
        if (has) Lsh[cls * kColStageSCap + pos] = make_uint2(min_, (uint32_t)tid);
        if (tid == 64) {
          Nsh[0] = __popcll(lo0) + __popcll(own0);
          Nsh[1] = __popcll(lo1) + __popcll(own1);
        }
      }
    }
    for (int r0 = l0; r0 < l1; r0 += JOBS) {
      const int q = r0 + job;
      const bool valid = q < l1;
      if (r0 + wave * (64 / T) >= l1) continue;  // no line of this wavefront in the round
      if (c > 0 || r0 > l0) {
        const size_t base = (size_t)(valid ? q : 0) * g.N + (size_t)n2 * L;
#pragma unroll
        for (int i = 0; i < 8; i++) v[i] = valid ? conv[base + t + i * T] : cw{0, 0};
      }
#pragma unroll
      for (int i = 0; i < 8; i++) v[i] = cmulc(v[i], w4[i]);
      cw* row = Rsh + (q - l0) * RS;
      line_fft<L, 1, 1, 1, W, true>(v, row, lt, t);
      if (valid) {
#pragma unroll
        for (int i = 0; i < 8; i++) row[t + i * T] = cscale(v[i], sc);
      }
    }
    __syncthreads();
    if constexpr (!MF) {
      if (c == 0) {
#pragma unroll
        for (int k = 0; k < M; k++) u[k] = Ush[k * RS + lane];
      }
    }
    // slot s mixes kv U_a conj(U_b) with its Raman line read as is, conjugated, or by real or
    // imaginary part; everything but the sample is wave-uniform and read from LDS by broadcast.
    // With parity-structured blocks each wavefront passes over its slots once per parity, so
    // the parity is a compile-time constant of the loop body (no per-slot selects).
    auto slot_pass = [&](auto par_c) {
      constexpr int par = decltype(par_c)::value;
      // each slot's metadata is read from LDS one slot ahead of its use
      uint32_t meta_next = s0 + wave < s1 ? Msh[wave] : 0u;
      for (int sl = s0 + wave; sl < s1; sl += NW) {
        const uint32_t meta = __builtin_amdgcn_readfirstlane(meta_next);  // wave-uniform
        meta_next = sl + NW < s1 ? Msh[sl + NW - s0] : 0u;
        if (PAR && (int)((meta >> 18) & 1u) != par) continue;
        const int pa = meta & 255u, pb = (meta >> 8) & 255u, kind = (meta >> 16) & 3u, line = (int)(meta >> 19);
        const W* qb = Qsh + (sl - s0) * MM;
        const cw r = Rsh[(line - l0) * RS + lane];
        const cw V = cmulc(Ush[pa * RS + lane], Ush[pb * RS + lane]);
        cw vr;
        switch (kind) {
          case kDirect: vr = r; break;
          case kConj: vr = cconj(r); break;
          case kRe: vr = {r.re, 0}; break;
          default: vr = {r.im, 0}; break;
        }
        const cw X = cadd(cscale(V, kv), cscale(vr, kr));
        cw y[M];
#pragma unroll
        for (int k = 0; k < M; k++) y[k] = cmul(X, u[k]);
        // the block's rows in groups of about 224 bytes (128 for FP32 beyond 10 modes, whose
        // registers are fuller; 64 in the 168-VGPR variant): each group's LDS reads are issued together (the scheduling
        // barrier keeps them ahead of their uses), so one LDS latency is exposed per group
        // rather than one per pair of coefficients
        constexpr int QROW = PAR ? KH : M;  // coefficients per row p
        constexpr int QGB = HI ? 64 : ((std::is_same<W, float>::value && M > 10) ? 128 : 224);
        constexpr int PG = (int)(QGB / sizeof(W)) / QROW > 0 ? (int)(QGB / sizeof(W)) / QROW : 1;
#pragma unroll
        for (int p0 = 0; p0 < M; p0 += PG) {
          W qv[PG * QROW];
#pragma unroll
          for (int e = 0; e < PG * QROW; e++) qv[e] = p0 * QROW + e < M * QROW ? qb[p0 * QROW + e] : (W)0;
          __builtin_amdgcn_sched_barrier(0);
#pragma unroll
          for (int pp = 0; pp < PG; pp++) {
            const int p = p0 + pp;
            if (p >= M) break;
            cw tsum = acc[p];
            if constexpr (PAR) {
              // only k with (p + k) % 2 == the slot's parity: q[p][kk] belongs to k = 2 kk + ((p + par) % 2)
#pragma unroll
              for (int kk = 0; kk < KH; kk++) {
                const int k = 2 * kk + ((p + par) & 1);
                if (k < M) tsum = cadd(tsum, cscale(y[k], qv[pp * QROW + kk]));
              }
            } else {
#pragma unroll
              for (int k = 0; k < M; k++) tsum = cadd(tsum, cscale(y[k], qv[pp * QROW + k]));
            }
            acc[p] = tsum;
          }
        }
      }
    };
    // the matrix-core form: per parity class, the pairs of its slots in the chunk through the
    // 32 x 32 x 2 products, then sigma_p += U_k B[p][k] for this lane's sample
    auto mf_pass = [&](auto cls_c) {
      constexpr int cc = decltype(cls_c)::value;
      const int n = Nsh[cc];
      const int npairs = (n + 1) >> 1;
      constexpr int TPW = MR::TPW;
      constexpr int PSTEP = MR::TSPLIT ? 1 : 2;  // pairs between a wavefront's consecutive pairs
      mfma_acc_t ar[TPW], ai[TPW];
#pragma unroll
      for (int t = 0; t < TPW; t++) {
#pragma unroll
        for (int r = 0; r < 16; r++) {
          ar[t][r] = 0.f;
          ai[t][r] = 0.f;
        }
      }
      // the operands of pair pi (this lane's slot 2 pi + hh: X at its sample and its A-operand
      // coefficients; zeros past the class's last slot), built one pair ahead of its products so
      // that their LDS reads overlap the matrix-core work of the pair before; the slot's
      // descriptor (metadata, offset in the chunk) is read one pair before that
      auto descriptor = [&](int pi) { return Lsh[cc * kColStageSCap + (2 * pi + hh < n ? 2 * pi + hh : 0)]; };
      auto operands = [&](int pi, uint2 d, W& xre, W& xim, W (&qa)[TPW]) {
        // past the last slot: slot 0's data (finite) times zero, so that no branch splits the reads
        const W hm = 2 * pi + hh < n ? (W)1 : (W)0;
        const uint32_t meta = d.x;
        const int x = (int)d.y;
        const int pa = meta & 255u, pb = (meta >> 8) & 255u, kind = (meta >> 16) & 3u, line = (int)(meta >> 19);
        const cw r = Rsh[(line - l0) * RS + smp];
        const cw ua = Ush[pa * RS + smp], ub = Ush[pb * RS + smp];
#pragma unroll
        for (int t = 0; t < TPW; t++) {
          const int o = aoff[cc][t];
          qa[t] = Qsh[x * MM + (o >= 0 ? o : 0)] * (o >= 0 ? hm : (W)0);
        }
        const cw V = cmulc(ua, ub);
        const W vre = kind == kIm ? r.im : r.re;
        const W vim = kind == kDirect ? r.im : (kind == kConj ? -r.im : (W)0);
        xre = (kv * V.re + kr * vre) * hm;
        xim = (kv * V.im + kr * vim) * hm;
      };
      int pi = MR::TSPLIT ? 0 : (wave >> 1);
      W xre, xim, qa[TPW];
      uint2 dn = {0u, 0u};
      if (n > 0) {
        operands(pi, descriptor(pi), xre, xim, qa);
        dn = descriptor(pi + PSTEP);
      }
#pragma nounroll
      for (; pi < npairs; pi += PSTEP) {  // wave-uniform
        W nre, nim, nqa[TPW];
        const uint2 d2 = dn;
        dn = descriptor(pi + 2 * PSTEP);
        operands(pi + PSTEP, d2, nre, nim, nqa);
#pragma unroll
        for (int t = 0; t < TPW; t++) {
          if (MR::TSPLIT && TPW * tg + t >= MR::RT) continue;  // wave-uniform: a tile of the group absent
          ar[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(qa[t], xre, ar[t], 0, 0, 0);
          ai[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(qa[t], xim, ai[t], 0, 0, 0);
        }
        xre = nre;
        xim = nim;
#pragma unroll
        for (int t = 0; t < TPW; t++) qa[t] = nqa[t];
      }
      // accumulator register r of tile t holds u = 16 t + r: mode p = u / KHH, and k from the
      // half; the tile group is a compile-time constant of each epilogue instance
      auto epilogue = [&](auto tg_c) {
        constexpr int TG = decltype(tg_c)::value;
#pragma unroll
        for (int j = 0; j < TPW; j++) {
          const int t = TPW * TG + j;
          if (t >= MR::RT) break;
#pragma unroll
          for (int r = 0; r < 16; r++) {
            const int uu = 16 * t + r;
            const int p = uu / MR::KHH;
            if (p < M) {
              const int kk0 = 2 * (uu % MR::KHH);
              const int kA = PAR ? 2 * kk0 + ((p + cc) & 1) : kk0;
              const int kB = PAR ? kA + 2 : kA + 1;
              // U_k of this lane's half from its LDS row (a padding row's products are zero: any
              // field row serves)
              const int row = kA >= M ? 0 : (kB < M ? kA + (kB - kA) * hh : kA);
              const cw uk = Ush[row * RS + smp];
              acc[p] = cadd(acc[p], cmul(uk, cw{ar[j][r], ai[j][r]}));
            }
          }
        }
      };
      if constexpr (!MR::TSPLIT) {
        epilogue(std::integral_constant<int, 0>{});
      } else {
        if (tg == 0) epilogue(std::integral_constant<int, 0>{});
        else epilogue(std::integral_constant<int, 1>{});
      }
      // the sums are taken here, so that the accumulators of one class are free before the next
      // class's products start (otherwise the compiler sinks them past those and spills)
#pragma unroll
      for (int p = 0; p < M; p++) asm volatile("" : "+v"(acc[p].re), "+v"(acc[p].im));
    };
    if constexpr (MF) {
      mf_pass(std::integral_constant<int, 0>{});
      if constexpr (PAR) mf_pass(std::integral_constant<int, 1>{});
    } else {
      slot_pass(std::integral_constant<int, 0>{});
      if constexpr (PAR) slot_pass(std::integral_constant<int, 1>{});
    }
  }
  __syncthreads();  // every lane done with Ush and Rsh
  // NW partial sums of every mode and sample (MF: one per pair parity, or tile group, and lane half)
  const int part = MF ? 2 * (wave >> 1) + hh : wave;
#pragma unroll
  for (int p = 0; p < M; p++) Rsh[(part * M + p) * L + smp] = acc[p];
  __syncthreads();
  for (int e = tid; e < M * L; e += NT) {
    const int p = e / L, j = e % L;
    cw tsum = Rsh[p * L + j];
#pragma unroll
    for (int w = 1; w < NW; w++) tsum = cadd(tsum, Rsh[(w * M + p) * L + j]);
    Ush[p * RS + j] = tsum;
  }
  __syncthreads();
  if (wave * (64 / T) < M) {
    const int q = job;
    const bool valid = q < M;
    cw* row = Ush + q * RS;
#pragma unroll
    for (int i = 0; i < 8; i++) v[i] = valid ? row[t + i * T] : cw{0, 0};
    line_fft<L, -1, 1, 1, W, true>(v, row, lt, t);
    if (valid) {
      const size_t base = (size_t)q * g.N + n2;
#pragma unroll
      for (int i = 0; i < 8; i++) store_stream64(yout + base + (size_t)(t + i * T) * g.N2, cmul(v[i], w4[i]));
    }
  }
}

// ---------------------------------------------------------------------------
// Dense-block Kerr + Raman gather for M <= 16 modes, 64 samples per workgroup:
//   sigma_p = sum over slots s of  sum_k Qd[s][p][k] * X_s * U_k
// Each lane keeps its sample's U_k (all k) and sigma_p (all p) in registers; X_s
// comes from LDS (one read per slot) and the M x M block Qd[s] (zeros included)
// from scalar loads. The four wavefronts take every fourth slot and their
// partial sums are added in wavefront order at the end.
// ---------------------------------------------------------------------------
constexpr int kDenseCap = 4096;  // complex LDS values (64 KB)

int dense_gather_chunk(int M) {
  if (M < 1 || M > 16) return 0;
  const int room = kDenseCap / 64 - M;
  return room >= 4 * M ? room : 0;
}

template <int M, class W>
__global__ void __launch_bounds__(256) k_gather_dense(Geometry g, const cpx<W>* __restrict__ U,
                                                      const cpx<W>* __restrict__ VR, cpx<W>* __restrict__ sigma,
                                                      GatherArgs a, const W* __restrict__ Qd, int chunk) {
  using cw = cpx<W>;
  __shared__ cw sh[kDenseCap];
  constexpr int TS = 64;
  cw* Ush = sh;           // [M][TS]
  cw* Xsh = sh + M * TS;  // [chunk][TS]; reused for the partial sums
  const int t0 = blockIdx.x * TS;
  const int tid = threadIdx.x, lane = tid % TS;
  const int wave = __builtin_amdgcn_readfirstlane(tid / TS);
#pragma unroll 4
  for (int e = tid; e < M * TS; e += 256) Ush[e] = U[(size_t)(e / TS) * g.N + t0 + e % TS];
  __syncthreads();
  cw u[M], acc[M];
#pragma unroll
  for (int k = 0; k < M; k++) {
    u[k] = Ush[k * TS + lane];
    acc[k] = {0, 0};
  }
  for (int s0 = 0; s0 < a.n_slots; s0 += chunk) {
    const int ns = min(chunk, a.n_slots - s0);
    __syncthreads();  // previous chunk consumed
    slot_mix_tile(a, g, VR, Ush, Xsh, s0, ns, t0, tid, 256);
    __syncthreads();
    for (int sl = s0 + wave; sl < s0 + ns; sl += 4) {
      const cw X = Xsh[(sl - s0) * TS + lane];
      const W* qb = Qd + (size_t)sl * M * M;
      cw y[M];
#pragma unroll
      for (int k = 0; k < M; k++) y[k] = cmul(X, u[k]);
#pragma unroll
      for (int p = 0; p < M; p++) {
        cw t = acc[p];
#pragma unroll
        for (int k = 0; k < M; k++) t = cadd(t, cscale(y[k], qb[p * M + k]));
        acc[p] = t;
      }
    }
  }
  __syncthreads();  // X consumed: the region takes the partial sums
#pragma unroll
  for (int p = 0; p < M; p++) Xsh[(wave * M + p) * TS + lane] = acc[p];
  __syncthreads();
  for (int e = tid; e < M * TS; e += 256) {
    const int p = e / TS, j = e % TS;
    cw t = Xsh[p * TS + j];
#pragma unroll
    for (int w = 1; w < 4; w++) t = cadd(t, Xsh[(w * M + p) * TS + j]);
    sigma[(size_t)p * g.N + t0 + j] = t;
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

// Per-mode photon number of a spectral state in four-step order (S[k1*N2 + k2] holds bin
// k = k1 + N1*k2): out[p] = scale * sum_k |S_p[k]|^2 / (omega0 - dalpha * k_signed), k_signed = k
// for k <= N/2, else k - N (the reference's frequency grid, src/MultimodePropagation.cpp:171-175).
// One workgroup per mode, fixed reduction order; FP64 throughout.
__global__ void __launch_bounds__(256) k_photon_number(Geometry g, const cd* __restrict__ S, double* __restrict__ out,
                                                       double omega0, double dalpha, double scale) {
  __shared__ double red[256];
  const int p = blockIdx.x;
  double s = 0.;
  for (int idx = threadIdx.x; idx < g.N; idx += 256) {
    const int k1 = idx / g.N2, k2 = idx - k1 * g.N2;
    const int k = k1 + g.N1 * k2;
    const int ks = k <= g.N / 2 ? k : k - g.N;
    const cd x = S[(size_t)p * g.N + idx];
    s += (x.re * x.re + x.im * x.im) / (omega0 - dalpha * (double)ks);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[p] = red[0] * scale;
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
    case 512: return FN<512>(__VA_ARGS__);                 \
    case 1024: return FN<1024>(__VA_ARGS__);               \
    default: return hipErrorInvalidValue;                  \
  }

// column lengths: N1 <= 64 for every supported nt (the split keeps rows within 1024)
#define MMP_DISPATCH_COL_LW(Lv, FN, W, ...)                \
  switch (Lv) {                                            \
    case 8: return FN<8, W>(__VA_ARGS__);                  \
    case 16: return FN<16, W>(__VA_ARGS__);                \
    case 32: return FN<32, W>(__VA_ARGS__);                \
    case 64: return FN<64, W>(__VA_ARGS__);                \
    default: return hipErrorInvalidValue;                  \
  }

#define MMP_DISPATCH_LW(Lv, FN, W, ...)                    \
  switch (Lv) {                                            \
    case 8: return FN<8, W>(__VA_ARGS__);                  \
    case 16: return FN<16, W>(__VA_ARGS__);                \
    case 32: return FN<32, W>(__VA_ARGS__);                \
    case 64: return FN<64, W>(__VA_ARGS__);                \
    case 128: return FN<128, W>(__VA_ARGS__);              \
    case 256: return FN<256, W>(__VA_ARGS__);              \
    case 512: return FN<512, W>(__VA_ARGS__);              \
    case 1024: return FN<1024, W>(__VA_ARGS__);            \
    default: return hipErrorInvalidValue;                  \
  }

template <int L, class W>
static hipError_t row_head_impl(const Geometry& g, const Twiddles& tw, const HeadArgs& a, const Batch& b, int F,
                                hipStream_t s) {
  using C = RowCfg<L>;
  const int rows = g.M * g.N1;
  dim3 grid((rows + C::G - 1) / C::G, F);
  // the table-free form only for FP32 work arrays (mixed precision)
  if (a.recompute_e && !std::is_same<W, float>::value) return hipErrorInvalidValue;
  if (a.recompute_e)
    MMP_LAUNCH((k_row_head<L, W, std::is_same<W, float>::value>), grid, dim3(C::THREADS), 0, s, g,
               TwOf<W>::row(tw), a, b);
  else
    MMP_LAUNCH((k_row_head<L, W, false>), grid, dim3(C::THREADS), 0, s, g, TwOf<W>::row(tw), a, b);
  return hipGetLastError();
}
template <class W>
hipError_t launch_row_head(const Geometry& g, const Twiddles& tw, const HeadArgs& a, const Batch& b, int F,
                           hipStream_t s) {
  if (F < 1 || F > 65535) return hipErrorInvalidValue;
  MMP_DISPATCH_LW(g.N2, row_head_impl, W, g, tw, a, b, F, s);
}

// Column-block width: 8 when that already gives >= 256 workgroups, else 4, else 2.
template <int L>
static int pick_cb(const Geometry& g, int lines) {
  const int cbs[3] = {8, 4, 2};
  for (int cb : cbs) {
    const int T = L / 8;
    const int G = (cb * T >= 64) ? 1 : 64 / (cb * T);
    const int blocks = (g.N2 / cb) * ((lines + G - 1) / G);
    if (blocks >= 256 || cb == 2) return cb;
  }
  return 2;
}

template <int L, int CB>
static dim3 col_grid(const Geometry& g, int lines) {
  using C = ColCfg<L, CB>;
  return dim3((g.N2 / CB) * ((lines + C::G - 1) / C::G));
}

#define MMP_COL_LAUNCH(KERNEL, LINES, ...)                                                                      \
  switch (pick_cb<L>(g, LINES)) {                                                                               \
    case 8: MMP_LAUNCH((KERNEL<L, 8, W>), (col_grid<L, 8>(g, LINES)), (dim3(ColCfg<L, 8>::THREADS)), 0, s, \
                               __VA_ARGS__); break;                                                            \
    case 4: MMP_LAUNCH((KERNEL<L, 4, W>), (col_grid<L, 4>(g, LINES)), (dim3(ColCfg<L, 4>::THREADS)), 0, s, \
                               __VA_ARGS__); break;                                                            \
    default: MMP_LAUNCH((KERNEL<L, 2, W>), (col_grid<L, 2>(g, LINES)), (dim3(ColCfg<L, 2>::THREADS)), 0,  \
                                s, __VA_ARGS__); break;                                                        \
  }                                                                                                             \
  return hipGetLastError();

template <int L, class W>
static hipError_t col_inverse_impl(const Geometry& g, const Twiddles& tw, const cpx<W>* in, cpx<W>* out, int lines,
                                   double scale, double divisor, hipStream_t s) {
  MMP_COL_LAUNCH(k_col_inverse, lines, g, TwOf<W>::col(tw), TwOf<W>::four(tw), in, out, lines, scale, divisor)
}
template <class W>
hipError_t launch_col_inverse(const Geometry& g, const Twiddles& tw, const cpx<W>* in, cpx<W>* out, int lines,
                              double scale, double divisor, hipStream_t s) {
  MMP_DISPATCH_COL_LW(g.N1, col_inverse_impl, W, g, tw, in, out, lines, scale, divisor, s);
}

template <int L, class W>
static hipError_t col_forward_impl(const Geometry& g, const Twiddles& tw, const cpx<W>* in, cpx<W>* out, int lines,
                                   hipStream_t s) {
  MMP_COL_LAUNCH(k_col_forward, lines, g, TwOf<W>::col(tw), TwOf<W>::four(tw), in, out, lines)
}
template <class W>
hipError_t launch_col_forward(const Geometry& g, const Twiddles& tw, const cpx<W>* in, cpx<W>* out, int lines,
                              hipStream_t s) {
  MMP_DISPATCH_COL_LW(g.N1, col_forward_impl, W, g, tw, in, out, lines, s);
}

template <int L, class W>
static hipError_t col_forward_real_impl(const Geometry& g, const Twiddles& tw, const double* in, cd* out,
                                        hipStream_t s) {
  MMP_COL_LAUNCH(k_col_forward_real, 1, g, tw.coltw, tw.fourtw, in, out)
}
hipError_t launch_col_forward_real(const Geometry& g, const Twiddles& tw, const double* in, cd* out,
                                   hipStream_t s) {
  MMP_DISPATCH_COL_LW(g.N1, col_forward_real_impl, double, g, tw, in, out, s);
}

template <int L, class W>
static hipError_t col_pairs_forward_impl(const Geometry& g, const Twiddles& tw, const cpx<W>* U, cpx<W>* out,
                                         int lines, const PairLineArgs& pl, hipStream_t s) {
  MMP_COL_LAUNCH(k_col_pairs_forward, lines, g, TwOf<W>::col(tw), TwOf<W>::four(tw), U, out, lines, pl)
}
template <class W>
hipError_t launch_col_pairs_forward(const Geometry& g, const Twiddles& tw, const cpx<W>* U, cpx<W>* out, int lines,
                                    const PairLineArgs& pl, hipStream_t s) {
  MMP_DISPATCH_COL_LW(g.N1, col_pairs_forward_impl, W, g, tw, U, out, lines, pl, s);
}

template <int L>
static hipError_t row_forward_impl(const Geometry& g, const Twiddles& tw, cd* data, int lines, hipStream_t s) {
  using C = RowCfg<L>;
  const int rows = lines * g.N1;
  MMP_LAUNCH(k_row_forward<L>, dim3((rows + C::G - 1) / C::G), dim3(C::THREADS), 0, s, g, tw.rowtw, data,
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
  MMP_LAUNCH(k_row_inverse_phase<L>, dim3((rows + C::G - 1) / C::G), dim3(C::THREADS), 0, s, g, tw.rowtw,
                     in, out, imL, st, use_phase);
  return hipGetLastError();
}
hipError_t launch_row_inverse_phase(const Geometry& g, const Twiddles& tw, const cd* in, cd* out,
                                    const double* imL, const StepState* st, int use_phase, hipStream_t s) {
  MMP_DISPATCH_L(g.N2, row_inverse_phase_impl, g, tw, in, out, imL, st, use_phase, s);
}

template <int L, class W>
static hipError_t row_raman_impl(const Geometry& g, const Twiddles& tw, cpx<W>* data, int lines, const cpx<W>* hspec,
                                 const Batch& b, int F, hipStream_t s) {
  using C = RowCfg<L>;
  const int rows = lines * g.N1;
  MMP_LAUNCH((k_row_raman<L, W>), dim3((rows + C::G - 1) / C::G, F), dim3(C::THREADS), 0, s, g,
                     TwOf<W>::row(tw), data, lines, hspec, b);
  return hipGetLastError();
}
template <class W>
hipError_t launch_row_raman(const Geometry& g, const Twiddles& tw, cpx<W>* data, int lines, const cpx<W>* hspec,
                            const Batch& b, int F, hipStream_t s) {
  if (F < 1 || F > 65535) return hipErrorInvalidValue;
  MMP_DISPATCH_LW(g.N2, row_raman_impl, W, g, tw, data, lines, hspec, b, F, s);
}

template <int L, class W>
static hipError_t row_raman_cols_impl(const Geometry& g, const Twiddles& tw, const cpx<W>* in, cpx<W>* out, int lines,
                                      const cpx<W>* hspec, const Batch& b, int F, int in_cols, hipStream_t s) {
  using C = RamanColsCfg<L, W>;
  if (g.N1 % C::R) return hipErrorInvalidValue;
  MMP_LAUNCH((k_row_raman_cols<L, W>), dim3(lines * g.N1 / C::R, F), dim3(C::THREADS), 0, s, g, TwOf<W>::row(tw), in,
             out, hspec, b, in_cols);
  return hipGetLastError();
}
template <class W>
hipError_t launch_row_raman_cols(const Geometry& g, const Twiddles& tw, const cpx<W>* in, cpx<W>* out, int lines,
                                 const cpx<W>* hspec, const Batch& b, int F, int in_cols, hipStream_t s) {
  if (F < 1 || F > 65535 || lines < 1 || in == out) return hipErrorInvalidValue;
  MMP_DISPATCH_LW(g.N2, row_raman_cols_impl, W, g, tw, in, out, lines, hspec, b, F, in_cols, s);
}

template <class W>
hipError_t launch_gather(const Geometry& g, const cpx<W>* U, const cpx<W>* VR, cpx<W>* sigma, const GatherArgs& a,
                         hipStream_t s) {
  if (g.N % kGatherTS) return hipErrorInvalidValue;
  int waves = g.M < 8 ? g.M : 8;
  if ((g.M + waves - 1) / waves > kGatherMaxPW) return hipErrorInvalidValue;
  dim3 grid(g.N / kGatherTS), block(kGatherTS * waves);
  if (a.cap == 4096) MMP_LAUNCH((k_gather<4096, W>), grid, block, 0, s, g, U, VR, sigma, a);
  else if (a.cap == 8192) MMP_LAUNCH((k_gather<8192, W>), grid, block, 0, s, g, U, VR, sigma, a);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

size_t colstage_pairs_lds_bytes(int M, int columns, bool mixed) {
  return (size_t)(ColStageCfg::FFT + M * kColStageL * columns) * (mixed ? sizeof(cf) : sizeof(cd));
}

// Measured on one MI355X (M = 2..15, NT = 4096..65536, 1 to 36 fibers): four columns per
// workgroup took the pass from 0.6-1.5 TB/s to 2-3.4 TB/s; eight gained nothing over four
// (twice the LDS, lower occupancy). Line groups fill the CUs when the column blocks alone are
// too few, but each group transforms every field again, so a group keeps at least 4 lines.
void colstage_pairs_plan(int M, int N2, int n_lines, int F, bool mixed, int& columns, int& groups) {
  columns = 1;
  for (int cw : {4, 2})
    if (N2 % cw == 0 && colstage_pairs_lds_bytes(M, cw, mixed) <= kColStagePairsLds) {
      columns = cw;
      break;
    }
  groups = 1;
  while ((long)(N2 / columns) * F * groups < 512 && groups * 2 * 4 <= n_lines) groups *= 2;
}

template <class W>
hipError_t launch_colstage_fields_pairs(const Geometry& g, const Twiddles& tw, const cpx<W>* yin, cpx<W>* U,
                                        cpx<W>* conv, int n_lines, const PairLineArgs& pl, double scale,
                                        const Batch& b, int F, int columns, int groups, int conv_cols, hipStream_t s) {
  if (g.N1 != kColStageL || g.M < 1 || g.M > kColStageMaxM || F < 1 || F > 65535 || n_lines < 1) return hipErrorInvalidValue;
  if (columns < 1 || g.N2 % columns || groups < 1 || groups > n_lines) return hipErrorInvalidValue;
  const size_t lds = colstage_pairs_lds_bytes(g.M, columns, std::is_same<W, float>::value);
  if (lds > kColStagePairsLds) return hipErrorInvalidValue;
  const dim3 grid((unsigned)(g.N2 / columns * groups), F);
  switch (columns) {
    case 1: MMP_LAUNCH((k_colstage_fields_pairs<W, 1>), grid, dim3(256), lds, s, g, TwOf<W>::col(tw),
                       TwOf<W>::four(tw), yin, U, conv, n_lines, pl, scale, b, groups, conv_cols); break;
    case 2: MMP_LAUNCH((k_colstage_fields_pairs<W, 2>), grid, dim3(256), lds, s, g, TwOf<W>::col(tw),
                       TwOf<W>::four(tw), yin, U, conv, n_lines, pl, scale, b, groups, conv_cols); break;
    case 4: MMP_LAUNCH((k_colstage_fields_pairs<W, 4>), grid, dim3(256), lds, s, g, TwOf<W>::col(tw),
                       TwOf<W>::four(tw), yin, U, conv, n_lines, pl, scale, b, groups, conv_cols); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

#define MMP_COLSTAGE_LAUNCH(MV, PARV)                                                                             \
  if constexpr (kColStageMfma<MV, W, PARV>)                                                                       \
    if (fa.waves == 4 && fa.matrix_cores) {                                                                       \
      MMP_LAUNCH((k_colstage_raman_gather<MV, 4, W, PARV, false, true>), dim3(g.N2, F), dim3(256), 0, s, g,      \
                 TwOf<W>::col(tw), TwOf<W>::four(tw), U, conv, yout, ga, fa, Qd, scale, b);                        \
      break;                                                                                                      \
    }                                                                                                             \
  if constexpr (kColStageHiOcc<MV, W, PARV>)                                                                      \
    if (fa.waves == 4 && fa.hi_occupancy) {                                                                       \
      MMP_LAUNCH((k_colstage_raman_gather<MV, 4, W, PARV, true>), dim3(g.N2, F), dim3(256), 0, s, g,             \
                 TwOf<W>::col(tw), TwOf<W>::four(tw), U, conv, yout, ga, fa, Qd, scale, b);                        \
      break;                                                                                                      \
    }                                                                                                             \
  if constexpr (kColStageWide<MV, W>)                                                                             \
    if (fa.waves == 8) {                                                                                          \
      MMP_LAUNCH((k_colstage_raman_gather<MV, 8, W, PARV>), dim3(g.N2, F), dim3(512), 0, s, g, TwOf<W>::col(tw), \
                 TwOf<W>::four(tw), U, conv, yout, ga, fa, Qd, scale, b);                                         \
      break;                                                                                                      \
    }                                                                                                             \
  MMP_LAUNCH((k_colstage_raman_gather<MV, 4, W, PARV>), dim3(g.N2, F), dim3(256), 0, s, g, TwOf<W>::col(tw),     \
             TwOf<W>::four(tw), U, conv, yout, ga, fa, Qd, scale, b);
#define MMP_COLSTAGE_CASE(MV)                                                                                    \
  case MV:                                                                                                        \
    if (fa.parity) {                                                                                              \
      MMP_COLSTAGE_LAUNCH(MV, true)                                                                               \
    } else {                                                                                                      \
      MMP_COLSTAGE_LAUNCH(MV, false)                                                                              \
    }                                                                                                             \
    break;

template <class W>
hipError_t launch_colstage_raman_gather(const Geometry& g, const Twiddles& tw, const cpx<W>* U, const cpx<W>* conv,
                                        cpx<W>* yout, const GatherArgs& ga, const FusedArgs& fa, const W* Qd,
                                        double scale, const Batch& b, int F, hipStream_t s) {
  if (g.N1 != kColStageL || F < 1 || F > 65535) return hipErrorInvalidValue;
  const int block = fa.parity ? g.M * ((g.M + 1) / 2) : g.M * g.M;
  if (fa.n_chunks < 1 || fa.max_chunk_lines < 1 || fa.max_chunk_lines > colstage_lines_per_chunk(g.M) ||
      fa.max_chunk_slots < 1 || fa.max_chunk_slots > colstage_slots_per_chunk() ||
      fa.max_chunk_slots * block > colstage_coefficients_per_chunk(g.M, std::is_same<W, float>::value))
    return hipErrorInvalidValue;
  switch (g.M) {
    MMP_COLSTAGE_CASE(1) MMP_COLSTAGE_CASE(2) MMP_COLSTAGE_CASE(3) MMP_COLSTAGE_CASE(4) MMP_COLSTAGE_CASE(5)
    MMP_COLSTAGE_CASE(6) MMP_COLSTAGE_CASE(7) MMP_COLSTAGE_CASE(8) MMP_COLSTAGE_CASE(9) MMP_COLSTAGE_CASE(10)
    MMP_COLSTAGE_CASE(11) MMP_COLSTAGE_CASE(12) MMP_COLSTAGE_CASE(13) MMP_COLSTAGE_CASE(14) MMP_COLSTAGE_CASE(15)
    MMP_COLSTAGE_CASE(16)
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

#define MMP_DENSE_CASE(MV)                                                                                     \
  case MV:                                                                                                      \
    MMP_LAUNCH((k_gather_dense<MV, W>), dim3(g.N / 64), dim3(256), 0, s, g, U, VR, sigma, a, Qd, chunk); \
    break;

template <class W>
hipError_t launch_gather_dense(const Geometry& g, const cpx<W>* U, const cpx<W>* VR, cpx<W>* sigma,
                               const GatherArgs& a, const W* Qd, hipStream_t s) {
  const int chunk = dense_gather_chunk(g.M);
  if (!chunk || g.N % 64) return hipErrorInvalidValue;
  switch (g.M) {
    MMP_DENSE_CASE(1) MMP_DENSE_CASE(2) MMP_DENSE_CASE(3) MMP_DENSE_CASE(4) MMP_DENSE_CASE(5) MMP_DENSE_CASE(6)
    MMP_DENSE_CASE(7) MMP_DENSE_CASE(8) MMP_DENSE_CASE(9) MMP_DENSE_CASE(10) MMP_DENSE_CASE(11) MMP_DENSE_CASE(12)
    MMP_DENSE_CASE(13) MMP_DENSE_CASE(14) MMP_DENSE_CASE(15) MMP_DENSE_CASE(16)
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

__global__ void __launch_bounds__(256) k_narrow(const cd* __restrict__ in, cf* __restrict__ out, size_t n) {
  for (size_t j = (size_t)blockIdx.x * 256 + threadIdx.x; j < n; j += (size_t)gridDim.x * 256)
    out[j] = narrow<float>(in[j]);
}

}  // namespace mmp

// Device memory of the no-allocation entry point (mmp_solve_static): every device array of one
// solve is carved from this array, so the call needs no hipMalloc. One copy per device, held
// until the process ends; its host-side name is not a device address (the address comes from
// hipGetSymbolAddress, once per device, in static_scratch_resolve).
static __device__ __align__(256) unsigned char scratch[mmp::kStaticScratchBytes];

namespace mmp {

namespace {
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kStaticScratchDevices] = {};
}  // namespace

hipError_t static_scratch_resolve(int device) {
  if (device < 0 || device >= kStaticScratchDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[device]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch_ptr[device] = p;
  return hipSuccess;
}

void* static_scratch(int device) {
  if (device < 0 || device >= kStaticScratchDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  return g_scratch_ptr[device];
}

hipError_t launch_narrow(const cd* in, cf* out, size_t n, hipStream_t s) {
  const size_t blocks = (n + 255) / 256;
  MMP_LAUNCH(k_narrow, dim3((unsigned)(blocks < 4096 ? (blocks ? blocks : 1) : 4096)), dim3(256), 0, s, in,
                     out, n);
  return hipGetLastError();
}

hipError_t launch_energy(const Geometry& g, const cd* fields, double* out, double dt, hipStream_t s) {
  MMP_LAUNCH(k_energy, dim3(g.M), dim3(256), 0, s, g, fields, out, dt);
  return hipGetLastError();
}

hipError_t launch_photon_number(const Geometry& g, const cd* spectra, double* out, double omega0, double dalpha,
                                double scale, hipStream_t s) {
  if (g.N1 * g.N2 != g.N || g.M < 1) return hipErrorInvalidValue;
  MMP_LAUNCH(k_photon_number, dim3(g.M), dim3(256), 0, s, g, spectra, out, omega0, dalpha, scale);
  return hipGetLastError();
}

#define MMP_INSTANTIATE_W(W)                                                                                        \
  template hipError_t launch_row_head<W>(const Geometry&, const Twiddles&, const HeadArgs&, const Batch&, int,    \
                                         hipStream_t);                                                             \
  template hipError_t launch_col_inverse<W>(const Geometry&, const Twiddles&, const cpx<W>*, cpx<W>*, int, double, \
                                            double, hipStream_t);                                                 \
  template hipError_t launch_col_forward<W>(const Geometry&, const Twiddles&, const cpx<W>*, cpx<W>*, int,         \
                                            hipStream_t);                                                         \
  template hipError_t launch_col_pairs_forward<W>(const Geometry&, const Twiddles&, const cpx<W>*, cpx<W>*, int,   \
                                                  const PairLineArgs&, hipStream_t);                              \
  template hipError_t launch_row_raman<W>(const Geometry&, const Twiddles&, cpx<W>*, int, const cpx<W>*,           \
                                          const Batch&, int, hipStream_t);                                         \
  template hipError_t launch_gather<W>(const Geometry&, const cpx<W>*, const cpx<W>*, cpx<W>*, const GatherArgs&,  \
                                       hipStream_t);                                                              \
  template hipError_t launch_gather_dense<W>(const Geometry&, const cpx<W>*, const cpx<W>*, cpx<W>*,               \
                                             const GatherArgs&, const W*, hipStream_t);                            \
  template hipError_t launch_row_raman_cols<W>(const Geometry&, const Twiddles&, const cpx<W>*, cpx<W>*, int,         \
                                               const cpx<W>*, const Batch&, int, int, hipStream_t);                 \
  template hipError_t launch_colstage_fields_pairs<W>(const Geometry&, const Twiddles&, const cpx<W>*, cpx<W>*,     \
                                                      cpx<W>*, int, const PairLineArgs&, double, const Batch&, int, \
                                                      int, int, int, hipStream_t);                                 \
  template hipError_t launch_colstage_raman_gather<W>(const Geometry&, const Twiddles&, const cpx<W>*,             \
                                                      const cpx<W>*, cpx<W>*, const GatherArgs&, const FusedArgs&, \
                                                      const W*, double, const Batch&, int, hipStream_t);
MMP_INSTANTIATE_W(double)
MMP_INSTANTIATE_W(float)

}  // namespace mmp
