// kernels.h -- launch interface of the gfx950 propagation kernels.
//
// Data layout in HBM (complex FP64 = 16 B, NT = N1 * N2, row-major [N1][N2]):
//   time domain   x[q][n1*N2 + n2]  (natural sample order t = n1*N2 + n2)
//   spectral      S[q][k1*N2 + k2]  holds X[k1 + N1*k2]  (four-step order)
// A forward transform is a column pass (length-N1 DFT down each column n2,
// then twiddle W_N^(n2*k1)) followed by a row pass (length-N2 DFT along each
// row k1); the inverse runs the row pass first. Spectral state never leaves
// the four-step order, so no transposes are needed. Inside one column-stage
// step, the fields U and the Raman lines after their row pass (on large grids
// also before it) are kept in column order instead, x[q][n2*N1 + n1], so that
// the column kernels read and write whole columns contiguously.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft_line.h"

namespace mmp {

constexpr int kMaxStages = 11;

// Events the next launches record at their dispatch's start and end (both null: none).
struct LaunchEvents {
  hipEvent_t start = nullptr;
  hipEvent_t stop = nullptr;
};
extern thread_local LaunchEvents g_launch_events;

// Conv-line kinds: how a pair slot reads its Raman-convolved product.
enum SlotKind : int { kDirect = 0, kConj = 1, kRe = 2, kIm = 3 };

struct Geometry {
  int N, N1, N2, M;
};

// Twiddle tables on the device, FP64 and (mixed precision only) the same values rounded to FP32.
struct Twiddles {
  const cd* rowtw;   // W_N2^j, j < N2
  const cd* coltw;   // W_N1^j, j < N1
  const cd* fourtw;  // W_N^e, e < N
  const cf* rowtw32;
  const cf* coltw32;
  const cf* fourtw32;
};

template <class W>
struct TwOf;
template <>
struct TwOf<double> {
  static const cd* row(const Twiddles& t) { return t.rowtw; }
  static const cd* col(const Twiddles& t) { return t.coltw; }
  static const cd* four(const Twiddles& t) { return t.fourtw; }
};
template <>
struct TwOf<float> {
  static const cf* row(const Twiddles& t) { return t.rowtw32; }
  static const cf* col(const Twiddles& t) { return t.coltw32; }
  static const cf* four(const Twiddles& t) { return t.fourtw32; }
};

// Per-step device state updated on the stream (z by repeated FP64 addition).
// halted: set by the step-closing kernel when the fiber's new spectral state is not finite; it
// then also parks the fiber (FiberParams::live = 0), so later kernels skip it.
struct StepState {
  double z;
  uint32_t steps;
  uint32_t halted;
  double n0;  // steps at the start of the step, for the step-closing head (written by stage head 1)
};

// Per-fiber scalars of a batch of fibers that share M, NT, the tableau and the pair structure.
struct FiberParams {
  double h;    // the step h (dimensionless delta z)
  double sgn;  // sign of gamma
  double dt;   // delta t, the Raman convolution's quadrature weight
  double kv;   // slot mix: 1 - fR (Raman on) or 1 (off)
  double kr;   // slot mix: fR T0 (Raman on) or 0 (off)
  uint32_t live;  // 0: the fiber is parked; the batched kernels leave its arrays and z untouched
  uint32_t pad;
};

// Fiber f of a batch (blockIdx.y in the batched kernels) finds its arrays at base + f * stride.
// Per fiber, per stage head i: kHeadCoef doubles, [0] = b_{i-1} h, [1 + e] = a_{i, a_idx[e]} h.
constexpr int kHeadCoef = kMaxStages + 1;

struct Batch {
  FiberParams* fp;        // [F] (the step-closing kernel parks a fiber whose state is not finite)
  const double* coef;     // [F][nheads][kHeadCoef] stage-head coefficients times the fiber's h
  size_t nheads;          // s + 1
  StepState* st;          // [F]
  size_t sf;              // M x NT field arrays (complex values)
  size_t sl;              // Raman line arrays (lines x NT)
  size_t se;              // Lawson table sets (tables x M x NT)
  size_t ss;              // stage stash (s x M x NT)
  size_t sq;              // dense coupling blocks (P x M x M coefficients)
  size_t sh;              // Raman spectra (NT)
};

// Arguments of the row "stage head" kernel for stage i of the tableau
// (i == s finishes the step). Lawson tables E_u = exp(c_u h L) * E1^n (one per
// distinct c_u, four-step order) are advanced by the head that last reads them
// in a step: E_u <- E_u * E1, as src/MultimodePropagation_CPU.cpp:37-44 does.
// The stage head works in FP64 on the spectral state; its row DFTs run in the work
// precision W of the step (ysig and yout are cpx<W> arrays).
struct HeadArgs {
  const void* ysig;    // column-forwarded sigma of stage i-1 (unused when has_prev == 0)
  cd* psi_hat;         // spectral state
  cd* acc;             // spectral RK accumulator
  cd* stash;           // [s][M][N] spectral N_j kept for later stages
  void* yout;          // row-inverse output for stage i (unused when final)
  cd* e_prev;          // Lawson table of stage i-1 (conjugated for N_{i-1})
  const cd* e_cur;     // Lawson table of stage i
  const cd* e1;        // exp(h L)
  int head;            // i (selects the fiber's b_{i-1} h and a_ij h in Batch::coef)
  double b_prev;       // b_{i-1} (0 when b_{i-1} == 0): whether N_{i-1} enters the step result
  double scale_cur;    // factor folded into the stage-i product (1)
  int has_prev;        // 1 when i > 0
  int prev_stage;      // i - 1
  int prev_is_first;   // 1 when i - 1 == 0 (acc starts from psi_hat)
  int stash_prev;      // 1 when N_{i-1} is read by a later stage
  int final_stage;     // 1 when i == s: psi_hat = acc, no row-inverse output
  int update_prev;     // 1 when this head is the last reader of e_prev in the step
  int chain_next;      // with final_stage: also build stage 0 of the next step (E_0 in e_cur, no a terms)
  int advance_z;       // with final_stage: z += h, steps += 1 on the fiber's StepState
  int n_a;             // number of nonzero a_ij, j < i, in increasing j
  int a_idx[kMaxStages];    // j (N_{i-1} comes from registers, earlier N_j from stash)
  double a_coef[kMaxStages];  // a_ij (the fiber's a_ij h come from Batch::coef)
  // Lawson factors computed from the FP64 phase instead of read from the tables (mixed precision):
  // E_u = exp(i Im L (n + c_u) h) with n the fiber's step count at the start of the step
  // (StepState::steps, or StepState::n0 in the step-closing head, whose own block 0 advances
  // the count), the chained next stage 0 at n + 1. No table is read or written.
  int recompute_e;
  int set_n0;          // stage head 1: StepState::n0 = steps
  double c_prev;       // c_{i-1}
  double c_cur;        // c_i (c_0 for the chained next stage 0)
  const double* imL;   // Im L in four-step order ([F][M][N], stride Batch::sf)
};

struct GatherArgs {
  const uint32_t* coff;    // [M][n_chunks+1] nonzero ranges of Q^(p) per slot chunk (sorted by slot)
  const uint32_t* kslot;   // [nnz] k | slot << 8
  const void* q;           // [nnz] Q / max|Q| in the work precision
  const uint32_t* slot_a;  // [P]
  const uint32_t* slot_b;  // [P]
  const uint32_t* slot_line;  // [P] Raman conv line
  const uint32_t* slot_kind;  // [P]
  int n_slots;             // P
  int chunk;               // slots per LDS chunk
  int n_chunks;
  int cap;                 // LDS capacity (complex values) of the kernel instance
  int raman;
  double kv, kr;           // slot mix: (1 - fR, fR T0) with Raman on, (1, 0) off
};

// LDS plan of the gather for M modes and P slots: capacity and slots per chunk (0 if none fits).
void gather_plan(int M, int P, int& cap, int& chunk);

struct PairLineArgs {
  const uint32_t* line_a;  // [n_lines]
  const uint32_t* line_b;  // [n_lines]; for self lines the second mode or 0xffffffff
  const uint32_t* line_self;  // [n_lines] 1 = packed self pairs, 0 = cross pair a, b
};

// Column-stage kernels (one workgroup per column, Raman on, N1 = 64, M <= 16): Raman lines
// are processed in chunks; chunk c covers lines [chunk_line[c], chunk_line[c+1]) and their
// slots [chunk_slot[c], chunk_slot[c+1]) (slots are numbered by line).
struct FusedArgs {
  const uint32_t* chunk_line;  // [n_chunks + 1]
  const uint32_t* chunk_slot;  // [n_chunks + 1]
  const uint32_t* slot_meta;   // [P] a | b << 8 | kind << 16 | parity << 18 | line << 19 of each slot
  int n_chunks, max_chunk_lines, max_chunk_slots;
  int parity;                  // coupling blocks hold only the entries (p, k) with (p + k) % 2 == the
                               // slot's parity: M x ceil(M/2) per slot, k = 2 kk + ((p + parity) % 2)
  int waves;                   // wavefronts of the gather workgroup: 4, or 8 (colstage_wide_ok)
  int hi_occupancy;            // with 4 wavefronts: the three-workgroups-per-CU variant where it
                               // exists (FP32, 7 to 10 modes); same arithmetic, same results
  int matrix_cores;            // with 4 wavefronts: the slot sums on the matrix cores where that
                               // variant exists (colstage_matrix_cores_ok); FP32 rounding differs
};

// Limits of one column-stage chunk: Raman lines, slots, and coupling coefficients
// (slots x M x M) held in LDS; and the largest mode count.
int colstage_max_modes();
// Whether the 8-wavefront gather workgroup exists for M modes in this precision.
bool colstage_wide_ok(int M, bool mixed);
// Whether the three-workgroups-per-CU 4-wavefront gather exists (FusedArgs::hi_occupancy).
bool colstage_hi_occupancy_ok(int M, bool mixed, bool parity);
// Whether the matrix-core gather exists (FusedArgs::matrix_cores): FP32, 4 to 16 modes with
// parity-structured coupling blocks, 4 to 10 without.
bool colstage_matrix_cores_ok(int M, bool mixed, bool parity);
int colstage_lines_per_chunk(int M);
int colstage_slots_per_chunk();
int colstage_coefficients_per_chunk(int M, bool mixed);
constexpr size_t kColStagePairsLds = 65536;
size_t colstage_pairs_lds_bytes(int M, int columns, bool mixed);
// C1 layout for F fibers: four columns per workgroup where the LDS allows (else two, else
// one), then line groups (powers of two, at least 4 lines each) until there are 512 workgroups.
void colstage_pairs_plan(int M, int N2, int n_lines, int F, bool mixed, int& columns, int& groups);

// Launchers (return hipSuccess or the launch error). Those templated on W run in the
// work precision W (double, or float for the mixed-precision step); the others are FP64.
// The stage kernels (row head, row Raman, column-stage C1 and C2) run F fibers of a batch
// in one launch (grid y = fiber); F = 1 for a single solver.
template <class W>
hipError_t launch_row_head(const Geometry& g, const Twiddles& tw, const HeadArgs& a, const Batch& b, int F,
                           hipStream_t s);
template <class W>
hipError_t launch_col_inverse(const Geometry& g, const Twiddles& tw, const cpx<W>* in, cpx<W>* out, int lines,
                              double scale, double divisor, hipStream_t s);
template <class W>
hipError_t launch_col_forward(const Geometry& g, const Twiddles& tw, const cpx<W>* in, cpx<W>* out, int lines,
                              hipStream_t s);
hipError_t launch_col_forward_real(const Geometry& g, const Twiddles& tw, const double* in, cd* out,
                                   hipStream_t s);
template <class W>
hipError_t launch_col_pairs_forward(const Geometry& g, const Twiddles& tw, const cpx<W>* U, cpx<W>* out, int lines,
                                    const PairLineArgs& pl, hipStream_t s);
hipError_t launch_row_forward(const Geometry& g, const Twiddles& tw, cd* data, int lines, hipStream_t s);
hipError_t launch_row_inverse_phase(const Geometry& g, const Twiddles& tw, const cd* in, cd* out,
                                    const double* imL, const StepState* st, int use_phase, hipStream_t s);
template <class W>
hipError_t launch_row_raman(const Geometry& g, const Twiddles& tw, cpx<W>* data, int lines, const cpx<W>* hspec,
                            const Batch& b, int F, hipStream_t s);
// The column-stage step's Raman row pass: the transform of k_row_raman on the lines of `in`
// (row order, or column order when in_cols: line q's sample (n1, n2) at q * N + n2 * N1 + n1),
// written to `out` (another buffer) in column order.
template <class W>
hipError_t launch_row_raman_cols(const Geometry& g, const Twiddles& tw, const cpx<W>* in, cpx<W>* out, int lines,
                                 const cpx<W>* hspec, const Batch& b, int F, int in_cols, hipStream_t s);
template <class W>
hipError_t launch_gather(const Geometry& g, const cpx<W>* U, const cpx<W>* VR, cpx<W>* sigma, const GatherArgs& a,
                         hipStream_t s);
// C1 takes `columns` (1, 2 or 4) neighbouring columns per workgroup and deals each column
// block's Raman lines to `groups` workgroups; the LDS it then needs is
// colstage_pairs_lds_bytes, at most kColStagePairsLds. Results do not depend on either.
// conv_cols: store the Raman lines in column order (for launch_row_raman_cols with in_cols).
template <class W>
hipError_t launch_colstage_fields_pairs(const Geometry& g, const Twiddles& tw, const cpx<W>* yin, cpx<W>* U,
                                        cpx<W>* conv, int n_lines, const PairLineArgs& pl, double scale,
                                        const Batch& b, int F, int columns, int groups, int conv_cols,
                                        hipStream_t s);
template <class W>
hipError_t launch_colstage_raman_gather(const Geometry& g, const Twiddles& tw, const cpx<W>* U, const cpx<W>* conv,
                                        cpx<W>* yout, const GatherArgs& ga, const FusedArgs& fa, const W* Qd,
                                        double scale, const Batch& b, int F, hipStream_t s);
// Dense-block gather (M <= 16): Qd is [P][M][M] with Qd[s][p][k] = sum of Q/max|Q| over the
// nonzeros (k, slot s) of Q^(p). dense_gather_chunk(M) is 0 when M is out of range.
int dense_gather_chunk(int M);
template <class W>
hipError_t launch_gather_dense(const Geometry& g, const cpx<W>* U, const cpx<W>* VR, cpx<W>* sigma,
                               const GatherArgs& a, const W* Qd, hipStream_t s);
// The no-allocation entry point's device scratch: a file-scope __device__ array of
// kStaticScratchBytes per device (kernels.hip). static_scratch_resolve looks up its address on
// `device` (which must be current) once; static_scratch returns it (nullptr before that).
constexpr size_t kStaticScratchBytes = (size_t)64 << 20;
constexpr int kStaticScratchDevices = 64;
hipError_t static_scratch_resolve(int device);
void* static_scratch(int device);

// out[j] = in[j] rounded to FP32, j < n
hipError_t launch_narrow(const cd* in, cf* out, size_t n, hipStream_t s);
hipError_t launch_energy(const Geometry& g, const cd* fields, double* out, double dt, hipStream_t s);
// out[p] = scale * sum_k |spectra_p[k]|^2 / (omega0 - dalpha * k_signed) over the M spectra in
// four-step order (k_signed in (-N/2, N/2], the reference's frequency grid).
hipError_t launch_photon_number(const Geometry& g, const cd* spectra, double* out, double omega0, double dalpha,
                                double scale, hipStream_t s);

}  // namespace mmp
