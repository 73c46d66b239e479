// setup.h -- host-side FP64 preparation of one propagation problem: the
// reference constructor's dimensionless scaling, Butcher tableau, linear
// operator and pair table, plus the tables the device kernels consume.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/mmprop.h"

namespace mmp {

struct Tableau {
  int s = 0;
  double a[11][11] = {};
  double b[11] = {};
  double c[11] = {};
};

// setRungeKuttaCoefficients (src/MultimodePropagation.cpp:15-27) with the
// tableaux of src/RungeKutta.cpp:3-72,125-175.
Tableau make_tableau(uint32_t method_order);

struct Problem {
  // sizes
  uint32_t M = 0, N = 0, N1 = 0, N2 = 0, n_beta = 0, nnz = 0, P = 0;
  Tableau rk;
  // dimensionless scalars (src/MultimodePropagation.cpp:31-56)
  double T0 = 0, sign_gamma = 1, max_beta2 = 0, max_q = 0, ld = 0;
  double t_final = 0, delta_t = 0, z_final = 0, delta_z = 0, conversion_factor = 0;
  uint32_t n_steps = 0, default_steps = 0;
  int raman = 0;
  double fR = 0;
  std::vector<double> beta_tilde;   // [M][n_beta]
  std::vector<double> psi0;         // [M][N] complex, conv * A0 (time domain)
  std::vector<double> imL_fourstep; // [M][N] Im L_p(k), at S[k1*N2+k2] <-> k = k1 + N1*k2
  // Lawson tables (src/MultimodePropagation.cpp:190-199), four-step order, complex:
  // one table exp(c_u h L) per distinct c value, and E1 = exp(h L).
  std::vector<double> c_unique;       // distinct c_i in first-occurrence order
  std::vector<int> stage_table;       // [s] index into c_unique
  std::vector<int> table_last_head;   // [n_tables] head (stage i + 1) that last reads the table in a step
  std::vector<double> etab0;          // [n_tables][M][N] complex
  std::vector<double> e1;             // [M][N] complex
  std::vector<double> raman_response;  // [N]
  // pair slots (src/MultimodePropagation.cpp:139-155)
  std::vector<uint32_t> pair_codes;  // [P] first-occurrence order of l + m*M
  std::vector<uint32_t> slot_a, slot_b, slot_line, slot_kind;  // [P]
  std::vector<uint32_t> pair_a_ref, pair_b_ref;  // [P] decoded pairs in the reference's slot order
  // Raman convolution lines
  std::vector<uint32_t> line_a, line_b, line_self;  // [n_lines]
  std::vector<uint32_t> line_slot0;                 // [n_lines + 1] first device slot of each line
  // parity structure of the coupling: every nonzero (p, k, slot s) has (p + k) % 2 == slot_parity[s]
  // (GRIN tensors: Q^(p)_klm != 0 only for k + l - m - p even), so each slot's M x M block is
  // half zeros in a fixed checkerboard
  bool parity = false;
  std::vector<uint8_t> slot_parity;  // [P]
  // column-stage kernels: chunks of Raman lines (0 chunks: not eligible), their line and slot
  // bounds, and the largest chunk's line and slot counts
  bool mixed = false;  // FP32 work arrays (the column-stage LDS limits depend on it)
  int fused_chunks = 0, fused_max_lines = 0, fused_max_slots = 0;
  std::vector<uint32_t> fused_line;                 // [fused_chunks + 1]
  std::vector<uint32_t> fused_slot;                 // [fused_chunks + 1]
  // Kerr gather: nonzeros of each Q^(p) sorted by (slot, k), with per-chunk ranges
  std::vector<uint32_t> off, k, slot;  // [M+1], [nnz], [nnz]
  int gather_cap = 0, gather_chunk = 0, gather_chunks = 0;
  std::vector<uint32_t> coff;          // [M][gather_chunks + 1]
  std::vector<double> q;               // [nnz] Q / max|Q|
  double kv = 1, kr = 0;               // slot mix: X = kv*V + kr*VR
  // which stage results must be kept for later stages
  std::vector<int> stash_needed;       // [s]
};

// Validates and prepares. On failure returns a status and fills err.
mmp_status prepare_problem(const mmp_params* in, const mmp_options& opt, Problem& out, std::string& err);

// Four-step split NT = N1 * N2: N1 = min(2^floor(log2(NT)/2), 64), raised to NT/1024 when the
// rows would exceed 1024 samples; false if NT is unsupported.
bool split_length(uint32_t nt, uint32_t& n1, uint32_t& n2);

// Chunks of Raman lines (at most max_lines each, within the column-stage LDS limits) for the
// column-stage kernels: fills P.fused_*; false (and no chunks) when those kernels do not apply.
bool plan_colstage_chunks(Problem& P, int max_lines);

// Sets the calling thread's mmp_last_error() text and returns st.
mmp_status record_error(mmp_status st, const std::string& msg);

// exp(-2*pi*i*j/n), j < n, evaluated in long double and rounded once.
std::vector<double> twiddle_table(uint32_t n);

}  // namespace mmp
