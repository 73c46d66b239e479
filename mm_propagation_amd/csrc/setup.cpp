// setup.cpp -- host-side FP64 preparation (see setup.h).
//
// The scalar and operator arithmetic follows the reference's expressions and
// operation order (src/MultimodePropagation.cpp:31-199) so that the prepared
// psi0 and Im L are bit-identical to the reference's; that file is built with
// FMA contraction off for this reason.
#include "setup.h"

#include <cmath>

// Keep a*b+c as two roundings, as the reference's x86-64 build does.
#pragma clang fp contract(off)
#include <algorithm>
#include <complex>
#include <cstring>

#include "kernels.h"

// glibc's cexp: the routine std::exp(std::complex<double>) resolves to in the reference's
// g++ build (ComplexArray.cpp exp()), called directly so the Lawson tables match it bit for bit.
extern "C" double _Complex cexp(double _Complex z);

namespace mmp {

static std::complex<double> exp_c99(std::complex<double> z) {
  double _Complex c;
  __real__ c = z.real();
  __imag__ c = z.imag();
  const double _Complex r = cexp(c);
  return {__real__ r, __imag__ r};
}

Tableau make_tableau(uint32_t method_order) {
  Tableau t;
  const double r = std::sqrt(21.);
  if (method_order == 8) {
    // 11 stages, order 8 (src/RungeKutta.cpp:125-175); the reference's 12th c entry is never read.
    t.s = 11;
    auto& a = t.a;
    a[1][0] = 1 / 2.;
    a[2][0] = 1 / 4.; a[2][1] = 1 / 4.;
    a[3][0] = 1 / 7.; a[3][1] = (-7. + 3. * r) / 98.; a[3][2] = (21. - 5. * r) / 49.;
    a[4][0] = (11. - r) / 84.; a[4][2] = (18. - 4. * r) / 63.; a[4][3] = (21. + r) / 252.;
    a[5][0] = (5. - r) / 48.; a[5][2] = (9. - r) / 36.; a[5][3] = (-231. - 14. * r) / 360.; a[5][4] = (63. + 7. * r) / 80.;
    a[6][0] = (10. + r) / 42.; a[6][2] = (-432. - 92. * r) / 315.; a[6][3] = (633. + 145. * r) / 90.;
    a[6][4] = (-504. - 115. * r) / 70.; a[6][5] = (63. + 13. * r) / 35.;
    a[7][0] = 1 / 14.; a[7][4] = (14. + 3. * r) / 126.; a[7][5] = (13. + 3. * r) / 63.; a[7][6] = 1 / 9.;
    a[8][0] = 1 / 32.; a[8][4] = (91. + 21. * r) / 576.; a[8][5] = 11 / 72.; a[8][6] = (-385. + 75. * r) / 1152.;
    a[8][7] = (63. - 13. * r) / 128.;
    a[9][0] = 1 / 14.; a[9][4] = 1 / 9.; a[9][5] = (-733. + 147. * r) / 2205.; a[9][6] = (515. - 111. * r) / 504.;
    a[9][7] = (-51. + 11. * r) / 56.; a[9][8] = (132. - 28. * r) / 245.;
    a[10][4] = (-42. - 7. * r) / 18.; a[10][5] = (-18. - 28. * r) / 45.; a[10][6] = (-273. + 53. * r) / 72.;
    a[10][7] = (301. - 53. * r) / 72.; a[10][8] = (28. + 28. * r) / 45.; a[10][9] = (49. + 7. * r) / 18.;
    const double b[11] = {1 / 20., 0., 0., 0., 0., 0., 0., 49 / 180., 16 / 45., 49 / 180., 1 / 20.};
    const double c[11] = {0., 1 / 2., 1 / 2., (7. - r) / 14., (7. - r) / 14., 1 / 2., (7. + r) / 14.,
                          (7. + r) / 14., 1 / 2., (7. - r) / 14., 1.};
    std::memcpy(t.b, b, sizeof b);
    std::memcpy(t.c, c, sizeof c);
  } else if (method_order == 6) {
    // 7 stages, order 6, nu = 1 (src/RungeKutta.cpp:24-72 called with nu = 1)
    const double nu = 1.;
    t.s = 7;
    auto& a = t.a;
    a[1][0] = nu;
    a[2][0] = (4. * nu - 1.) / (8. * nu); a[2][1] = 1. / (8. * nu);
    a[3][0] = (10. * nu - 2.) / (27. * nu); a[3][1] = 2. / (27. * nu); a[3][2] = (8. * nu) / (27. * nu);
    a[4][0] = (56. - 77. * nu + (8. - 17. * nu) * r) / (392. * nu); a[4][1] = (-8. * (7. + r)) / (392. * nu);
    a[4][2] = (48. * (7. + r) * nu) / (392. * nu); a[4][3] = (-3. * (21. + r) * nu) / (392. * nu);
    a[5][0] = (-5. * (287. * nu - 56. - (59. * nu - 8.) * r)) / (1960. * nu); a[5][1] = (-40. * (7. - r)) / (1960. * nu);
    a[5][2] = (320. * r * nu) / (1960. * nu); a[5][3] = (3. * (21. - 121. * r) * nu) / (1960. * nu);
    a[5][4] = (392. * (6. - r) * nu) / (1960. * nu);
    a[6][0] = (15. * ((30. * nu - 8.) - 7. * nu * r)) / (180. * nu); a[6][1] = 120. / (180. * nu);
    a[6][2] = (-40. * (5. + 7. * r) * nu) / (180. * nu); a[6][3] = (63. * (2. + 3. * r) * nu) / (180. * nu);
    a[6][4] = (-14. * (49. - 9. * r) * nu) / (180. * nu); a[6][5] = (70. * (7. + r) * nu) / (180. * nu);
    const double b[7] = {9 / 180., 0., 64 / 180., 0., 49 / 180., 49 / 180., 9 / 180.};
    const double c[7] = {0., nu, 1 / 2., 2 / 3., (7. + r) / 14., (7. - r) / 14., 1.};
    std::memcpy(t.b, b, sizeof b);
    std::memcpy(t.c, c, sizeof c);
  } else {
    // classic RK4 (src/RungeKutta.cpp:3-21)
    t.s = 4;
    t.a[1][0] = 1 / 2.;
    t.a[2][1] = 1 / 2.;
    t.a[3][2] = 1.;
    const double b[4] = {1 / 6., 1 / 3., 1 / 3., 1 / 6.};
    const double c[4] = {0., 1 / 2., 1 / 2., 1.};
    std::memcpy(t.b, b, sizeof b);
    std::memcpy(t.c, c, sizeof c);
  }
  return t;
}

bool split_length(uint32_t nt, uint32_t& n1, uint32_t& n2) {
  if (nt < 64 || nt > 65536 || (nt & (nt - 1)) != 0) return false;
  uint32_t lg = 0;
  while ((1u << lg) < nt) lg++;
  // columns of 64 samples while rows stay within 1024 (more columns = more parallel column
  // workgroups, and a 64-sample column of every field and Raman line fits in LDS)
  n1 = std::min(1u << (lg / 2), 64u);
  if (nt / n1 > 1024) n1 = nt / 1024;
  n2 = nt / n1;
  return true;
}

std::vector<double> twiddle_table(uint32_t n) {
  std::vector<double> w(2 * (size_t)n);
  const long double two_pi = 6.283185307179586476925286766559005768L;
  for (uint32_t j = 0; j < n; j++) {
    const long double ang = -two_pi * (long double)j / (long double)n;
    w[2 * j] = (double)cosl(ang);
    w[2 * j + 1] = (double)sinl(ang);
  }
  return w;
}

static int factorial_int(int a) {
  int r = 1;
  for (int i = 1; i <= a; i++) r *= i;
  return r;
}

static bool all_finite(const double* x, size_t n) {
  for (size_t i = 0; i < n; i++)
    if (!std::isfinite(x[i])) return false;
  return true;
}

mmp_status prepare_problem(const mmp_params* in, const mmp_options& opt, Problem& P, std::string& err) {
  if (!in) { err = "params is NULL"; return MMP_ERR_INVALID_ARGUMENT; }
  const uint32_t M = in->n_modes, nt = in->nt, nb = in->n_beta;
  if (M == 0) { err = "n_modes must be >= 1"; return MMP_ERR_INVALID_ARGUMENT; }
  if (M > 64) { err = "n_modes above 64 is not supported"; return MMP_ERR_UNSUPPORTED; }
  if (nt == 0) { err = "nt must be >= 1"; return MMP_ERR_INVALID_ARGUMENT; }
  if (nb < 3) {
    err = "dispersion_coefficients needs at least 3 orders per mode (beta_2 is read)";
    return MMP_ERR_INVALID_ARGUMENT;
  }
  if (!in->dispersion_coefficients || !in->coupling_offsets || !in->initial_fields) {
    err = "dispersion_coefficients, coupling_offsets and initial_fields are required";
    return MMP_ERR_INVALID_ARGUMENT;
  }
  if (in->n_steps == 0) { err = "n_steps must be >= 1"; return MMP_ERR_INVALID_ARGUMENT; }
  if (in->raman && !in->raman_response) { err = "raman_response is required with Raman on"; return MMP_ERR_INVALID_ARGUMENT; }
  P.mixed = opt.precision == MMP_PRECISION_MIXED;
  if (opt.precision != MMP_PRECISION_FP64 && opt.precision != MMP_PRECISION_MIXED) {
    err = "unknown precision";
    return MMP_ERR_INVALID_ARGUMENT;
  }
  if (opt.pair_decode != MMP_PAIRS_REFERENCE_CPU && opt.pair_decode != MMP_PAIRS_PHYSICAL) {
    err = "unknown pair_decode";
    return MMP_ERR_INVALID_ARGUMENT;
  }
  if ((opt.flags & MMP_FLAG_MATRIX_CORES) && (opt.flags & MMP_FLAG_NO_MATRIX_CORES)) {
    err = "MMP_FLAG_MATRIX_CORES and MMP_FLAG_NO_MATRIX_CORES exclude each other";
    return MMP_ERR_INVALID_ARGUMENT;
  }
  if (in->coupling_offsets[0] != 0) { err = "coupling_offsets[0] must be 0"; return MMP_ERR_INVALID_ARGUMENT; }
  for (uint32_t p = 0; p < M; p++)
    if (in->coupling_offsets[p + 1] < in->coupling_offsets[p]) {
      err = "coupling_offsets must be non-decreasing";
      return MMP_ERR_INVALID_ARGUMENT;
    }
  const uint32_t nnz = in->coupling_offsets[M];
  if (nnz && (!in->coupling_klm || !in->coupling_values)) { err = "coupling arrays missing"; return MMP_ERR_INVALID_ARGUMENT; }
  for (size_t i = 0; i < 3 * (size_t)nnz; i++)
    if (in->coupling_klm[i] >= M) { err = "coupling index (k, l or m) >= n_modes"; return MMP_ERR_INVALID_ARGUMENT; }
  if (!all_finite(in->dispersion_coefficients, (size_t)M * nb) || (nnz && !all_finite(in->coupling_values, nnz)) ||
      !all_finite(in->initial_fields, 2 * (size_t)M * nt) || (in->raman && !all_finite(in->raman_response, nt)) ||
      !std::isfinite(in->fiber_length) || !std::isfinite(in->time_window) || !std::isfinite(in->pulse_width) ||
      !std::isfinite(in->nonlinearity_const) || !std::isfinite(in->raman_proportion)) {
    err = "non-finite input";
    return MMP_ERR_NONFINITE;
  }
  if (!split_length(nt, P.N1, P.N2)) {
    err = "nt must be a power of two in [64, 65536]";
    return MMP_ERR_UNSUPPORTED;
  }
  P.M = M;
  P.N = nt;
  P.n_beta = nb;
  P.nnz = nnz;
  P.n_steps = in->n_steps;
  P.rk = make_tableau(in->method_order);
  P.raman = in->raman ? 1 : 0;
  P.fR = in->raman_proportion;

  // setDimensionlessVariables (src/MultimodePropagation.cpp:31-56)
  P.T0 = in->pulse_width;
  const double gamma = in->nonlinearity_const;
  P.sign_gamma = gamma >= 0 ? 1 : -1;
  double max_beta2 = 0.;
  for (uint32_t p = 0; p < M; p++) {
    const double v = std::abs(in->dispersion_coefficients[(size_t)p * nb + 2]);
    if (v > max_beta2) max_beta2 = v;
  }
  double max_q = 0.;
  for (uint32_t i = 0; i < nnz; i++) {
    const double v = std::abs(in->coupling_values[i]);
    if (v > max_q) max_q = v;
  }
  if (max_beta2 == 0.) { err = "max |beta_2| is zero (the dispersion length is undefined)"; return MMP_ERR_INVALID_ARGUMENT; }
  if (max_q == 0.) { err = "max |Q| is zero (the coupling normalisation is undefined)"; return MMP_ERR_INVALID_ARGUMENT; }
  if (P.T0 <= 0. || in->time_window <= 0. || in->fiber_length <= 0.) {
    err = "pulse_width, time_window and fiber_length must be positive";
    return MMP_ERR_INVALID_ARGUMENT;
  }
  P.max_beta2 = max_beta2;
  P.max_q = max_q;
  const double LD = (P.T0 * P.T0) / max_beta2;
  P.ld = LD;
  P.beta_tilde.resize((size_t)M * nb);
  std::vector<double> cf(nb);
  for (uint32_t i = 0; i < nb; i++) cf[i] = 1. / std::pow(P.T0, i) * LD;
  for (uint32_t p = 0; p < M; p++)
    for (uint32_t i = 0; i < nb; i++)
      P.beta_tilde[(size_t)p * nb + i] = in->dispersion_coefficients[(size_t)p * nb + i] * cf[i];
  P.t_final = in->time_window / P.T0;
  P.delta_t = (double)P.t_final / nt;
  P.z_final = in->fiber_length / LD;
  P.delta_z = P.z_final / in->n_steps;
  P.conversion_factor = std::sqrt(std::abs(gamma) * LD) * std::sqrt(max_q);
  P.default_steps = (uint32_t)std::floor(P.z_final / P.delta_z);

  // psi = A0 * conversion_factor (src/MultimodePropagation.cpp:59-63)
  P.psi0.resize(2 * (size_t)M * nt);
  for (size_t j = 0; j < 2 * (size_t)M * nt; j++) P.psi0[j] = in->initial_fields[j] * P.conversion_factor;

  // Distinct stage nodes c_i, one Lawson table each.
  for (int i = 0; i < P.rk.s; i++) {
    int u = 0;
    while (u < (int)P.c_unique.size() && P.c_unique[u] != P.rk.c[i]) u++;
    if (u == (int)P.c_unique.size()) P.c_unique.push_back(P.rk.c[i]);
    P.stage_table.push_back(u);
  }
  P.table_last_head.assign(P.c_unique.size(), 0);
  for (int i = 0; i < P.rk.s; i++) P.table_last_head[P.stage_table[i]] = i + 1;
  const size_t n_tab = P.c_unique.size();

  // L_p(alpha_j) (src/MultimodePropagation.cpp:165-187) and the Lawson tables
  // E1 = exp(L h), E_u = exp(L (c_u h)) (:190-199), stored in four-step order.
  P.imL_fourstep.resize((size_t)M * nt);
  P.etab0.resize(2 * n_tab * (size_t)M * nt);
  P.e1.resize(2 * (size_t)M * nt);
  const std::complex<double> iu(0., 1.);
  for (uint32_t p = 0; p < M; p++) {
    const double* bt = &P.beta_tilde[(size_t)p * nb];
    for (uint32_t j = 0; j < nt; j++) {
      const int kk = (j <= nt / 2) ? (int)j : (int)j - (int)nt;
      const double alpha = (2 * M_PI / P.t_final) * kk;
      std::complex<double> L = iu * (bt[0] - bt[1] * alpha - bt[2] / 2 * (alpha * alpha));
      for (uint32_t o = 3; o < nb; o++) L += iu * (std::pow(-1, o) * bt[o] / factorial_int((int)o) * std::pow(alpha, o));
      // natural index j = k1 + N1*k2  ->  four-step slot k1*N2 + k2
      const uint32_t k1 = j % P.N1, k2 = j / P.N1;
      const size_t slot4 = (size_t)p * nt + (size_t)k1 * P.N2 + k2;
      P.imL_fourstep[slot4] = L.imag();
      const std::complex<double> e1 = exp_c99(L * P.delta_z);
      P.e1[2 * slot4] = e1.real();
      P.e1[2 * slot4 + 1] = e1.imag();
      for (size_t u = 0; u < n_tab; u++) {
        const std::complex<double> eu = exp_c99(L * (P.c_unique[u] * P.delta_z));
        P.etab0[2 * (u * M * nt + slot4)] = eu.real();
        P.etab0[2 * (u * M * nt + slot4) + 1] = eu.imag();
      }
    }
  }

  if (P.raman) P.raman_response.assign(in->raman_response, in->raman_response + nt);

  // unique pair codes in first-occurrence order (src/MultimodePropagation.cpp:139-155)
  P.slot.resize(nnz);
  for (uint32_t e = 0; e < nnz; e++) {
    const uint32_t lm = in->coupling_klm[3 * e + 1] + in->coupling_klm[3 * e + 2] * M;
    uint32_t j = 0;
    while (j < P.pair_codes.size() && P.pair_codes[j] != lm) j++;
    if (j == P.pair_codes.size()) P.pair_codes.push_back(lm);
    P.slot[e] = j;  // src/MultimodePropagation_CPU.cpp:149-150 lookup
  }
  P.P = (uint32_t)P.pair_codes.size();
  // decoded pair of each slot: REFERENCE_CPU decodes the slot counter
  // (src/MultimodePropagation_CPU.cpp:128-129, 198-199), PHYSICAL the stored code.
  P.slot_a.resize(P.P);
  P.slot_b.resize(P.P);
  for (uint32_t j = 0; j < P.P; j++) {
    const uint32_t code = (opt.pair_decode == MMP_PAIRS_PHYSICAL) ? P.pair_codes[j] : j;
    P.slot_a[j] = code % M;
    P.slot_b[j] = code / M;
  }
  // Raman convolution lines: one per unordered cross pair (the reversed pair is its
  // conjugate since h_R is real) and one per two self pairs (|U_a|^2 + i |U_a'|^2).
  P.slot_line.assign(P.P, 0);
  P.slot_kind.assign(P.P, 0);
  if (P.raman) {
    std::vector<int> cross_line((size_t)M * M, -1);
    std::vector<uint32_t> self_modes;
    std::vector<int> self_pos(M, -1);
    for (uint32_t j = 0; j < P.P; j++) {
      const uint32_t a = P.slot_a[j], b = P.slot_b[j];
      if (a == b) {
        if (self_pos[a] < 0) {
          self_pos[a] = (int)self_modes.size();
          self_modes.push_back(a);
        }
      } else {
        const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
        if (cross_line[lo * M + hi] < 0) {
          cross_line[lo * M + hi] = (int)P.line_a.size();
          P.line_a.push_back(lo);
          P.line_b.push_back(hi);
          P.line_self.push_back(0);
        }
        P.slot_line[j] = (uint32_t)cross_line[lo * M + hi];
        P.slot_kind[j] = (a == lo) ? 0u : 1u;  // direct / conjugate
      }
    }
    const uint32_t first_self_line = (uint32_t)P.line_a.size();
    for (size_t s = 0; s < self_modes.size(); s += 2) {
      P.line_a.push_back(self_modes[s]);
      P.line_b.push_back(s + 1 < self_modes.size() ? self_modes[s + 1] : 0xffffffffu);
      P.line_self.push_back(1);
    }
    for (uint32_t j = 0; j < P.P; j++) {
      const uint32_t a = P.slot_a[j];
      if (a != P.slot_b[j]) continue;
      const int pos = self_pos[a];
      P.slot_line[j] = first_self_line + (uint32_t)pos / 2;
      P.slot_kind[j] = (pos % 2 == 0) ? 2u : 3u;  // real / imaginary part
    }
  }
  P.pair_a_ref = P.slot_a;
  P.pair_b_ref = P.slot_b;
  // Internal slot order: slots grouped by Raman line (a line's one or two slots are
  // consecutive), so a chunk of lines owns a contiguous slot range on the device.
  {
    std::vector<uint32_t> order(P.P);
    for (uint32_t j = 0; j < P.P; j++) order[j] = j;
    if (P.raman)
      std::stable_sort(order.begin(), order.end(),
                       [&](uint32_t x, uint32_t y) { return P.slot_line[x] < P.slot_line[y]; });
    std::vector<uint32_t> newid(P.P);
    for (uint32_t j = 0; j < P.P; j++) newid[order[j]] = j;
    auto permute = [&](std::vector<uint32_t>& v) {
      std::vector<uint32_t> w(v.size());
      for (uint32_t j = 0; j < P.P; j++) w[j] = v[order[j]];
      v.swap(w);
    };
    permute(P.slot_a);
    permute(P.slot_b);
    permute(P.slot_line);
    permute(P.slot_kind);
    for (uint32_t e = 0; e < nnz; e++) P.slot[e] = newid[P.slot[e]];
    const size_t nl = P.line_a.size();
    P.line_slot0.assign(nl + 1, P.P);
    for (uint32_t j = P.P; j-- > 0;)
      if (P.raman) P.line_slot0[P.slot_line[j]] = j;
    if (nl) P.line_slot0[nl] = P.P;
  }
  // Kerr gather: W = Q * (kv * V + kr * VR) * U_k with (kv, kr) = (1 - fR, fR T0) when Raman is
  // on (src/MultimodePropagation_CPU.cpp:223-226) and (1, 0) when off (:152-154).
  // Within each Q^(p) the nonzeros are sorted by (slot, k) so the device reuses a slot's mix
  // across its k's; this reorders the sum over nonzeros (the reference sums in insertion order).
  P.off.assign(in->coupling_offsets, in->coupling_offsets + M + 1);
  P.k.resize(nnz);
  P.q.resize(nnz);
  std::vector<uint32_t> slot_in = P.slot;
  for (uint32_t p = 0; p < M; p++) {
    std::vector<uint32_t> idx;
    for (uint32_t e = P.off[p]; e < P.off[p + 1]; e++) idx.push_back(e);
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t x, uint32_t y) {
      if (slot_in[x] != slot_in[y]) return slot_in[x] < slot_in[y];
      return in->coupling_klm[3 * x] < in->coupling_klm[3 * y];
    });
    for (size_t r = 0; r < idx.size(); r++) {
      const uint32_t e = P.off[p] + (uint32_t)r, src = idx[r];
      P.k[e] = in->coupling_klm[3 * src];
      P.slot[e] = slot_in[src];
      P.q[e] = in->coupling_values[src] / max_q;
    }
  }
  // parity structure (zero coupling values do not count)
  P.slot_parity.assign(P.P, 0);
  {
    std::vector<int> par(P.P, -1);
    bool ok = true;
    for (uint32_t p = 0; p < M && ok; p++)
      for (uint32_t e = P.off[p]; e < P.off[p + 1]; e++) {
        if (P.q[e] == 0.) continue;
        const int v = (int)((p + P.k[e]) & 1u);
        int& ps = par[P.slot[e]];
        if (ps < 0) ps = v;
        else if (ps != v) {
          ok = false;
          break;
        }
      }
    P.parity = ok;
    for (uint32_t j = 0; j < P.P; j++) P.slot_parity[j] = (uint8_t)(par[j] > 0 ? 1 : 0);
  }
  gather_plan((int)M, (int)P.P, P.gather_cap, P.gather_chunk);
  if (!P.gather_cap || P.gather_chunk <= 0) {
    err = "the Kerr gather tile does not fit (n_modes too large)";
    return MMP_ERR_UNSUPPORTED;
  }
  P.gather_chunks = ((int)P.P + P.gather_chunk - 1) / P.gather_chunk;
  P.coff.assign((size_t)M * (P.gather_chunks + 1), 0);
  for (uint32_t p = 0; p < M; p++) {
    uint32_t e = P.off[p];
    for (int c = 0; c <= P.gather_chunks; c++) {
      const uint32_t lim = (uint32_t)std::min<int>((int)P.P, c * P.gather_chunk);
      while (e < P.off[p + 1] && P.slot[e] < lim) e++;
      P.coff[(size_t)p * (P.gather_chunks + 1) + c] = (c == P.gather_chunks) ? P.off[p + 1] : e;
    }
  }
  P.kv = P.raman ? 1 - P.fR : 1.;
  P.kr = P.raman ? P.fR * P.T0 : 0.;
  plan_colstage_chunks(P, colstage_lines_per_chunk((int)M));
  // a stage result N_j needs a stash slot when a stage i > j + 1 reads it
  P.stash_needed.assign(P.rk.s, 0);
  for (int i = 0; i < P.rk.s; i++)
    for (int j = 0; j + 1 < i; j++)
      if (P.rk.a[i][j] != 0.) P.stash_needed[j] = 1;
  return MMP_OK;
}

// Column-stage kernels (Raman on, N1 = 64, M <= 16, dense coupling blocks): Raman lines are
// processed in chunks that respect the LDS limits on lines, slots and coupling blocks;
// chunk c covers lines [fused_line[c], fused_line[c+1]) and slots [fused_slot[c], fused_slot[c+1]).
bool plan_colstage_chunks(Problem& P, int max_lines) {
  const uint32_t M = P.M;
  P.fused_chunks = 0;
  P.fused_max_lines = 0;
  P.fused_max_slots = 0;
  P.fused_line.clear();
  P.fused_slot.clear();
  const int nl = (int)P.line_a.size();
  if (max_lines < 1 || max_lines > colstage_lines_per_chunk((int)M)) return false;
  if (P.raman && P.N1 == 64 && (int)M <= colstage_max_modes() && nl > 0 &&
      P.line_slot0.size() == (size_t)nl + 1 && P.line_slot0[0] == 0 && P.line_slot0[nl] == P.P) {
    const int block = P.parity ? (int)(M * ((M + 1) / 2)) : (int)(M * M);  // coefficients per slot
    const int max_slots =
        std::min(colstage_slots_per_chunk(), colstage_coefficients_per_chunk((int)M, P.mixed) / block);
    bool ok = max_slots >= 2;
    P.fused_line.push_back(0);
    P.fused_slot.push_back(0);
    int l = 0;
    while (ok && l < nl) {
      int lines = 0, slots = 0;
      while (l < nl && lines < max_lines) {
        const int ns = (int)(P.line_slot0[l + 1] - P.line_slot0[l]);
        if (slots + ns > max_slots) break;
        slots += ns;
        lines++;
        l++;
      }
      if (lines == 0) ok = false;
      P.fused_line.push_back((uint32_t)l);
      P.fused_slot.push_back(P.line_slot0[l]);
      P.fused_max_lines = std::max(P.fused_max_lines, lines);
      P.fused_max_slots = std::max(P.fused_max_slots, slots);
    }
    if (ok) {
      P.fused_chunks = (int)P.fused_line.size() - 1;
      return true;
    }
    P.fused_line.clear();
    P.fused_slot.clear();
    P.fused_max_lines = P.fused_max_slots = 0;
  }
  return false;
}

}  // namespace mmp
