// ref_setup_probe.cpp -- test harness around the reference's own solver setup.
//
// TEST INFRASTRUCTURE ONLY. Built by oracle/Makefile together with the
// reference's unmodified, FFTW-free sources (src/MultimodePropagation.cpp,
// src/RungeKutta.cpp, src/ComplexArray.cpp, src/ComplexArraysContainer.cpp,
// src/Sparse3DArray.cpp, compiled where they lie under the reference tree)
// into oracle/_ref/libref_setup.so. It subclasses the reference's abstract
// MultimodePropagation (include/MultimodePropagation.hpp:36-94) so the
// constructor, initLawsonRK, computeLinearOperators and
// initExponentialOperatorsMatrix run exactly as shipped, and copies their
// products out through a C interface. The stepping code
// (src/MultimodePropagation_CPU.cpp) needs fftw3.h, which this image lacks,
// so it is not built.
#include "MultimodePropagation.hpp"

#include <cstring>

#include "../include/mmprop.h"

namespace {

class SetupProbe : public MultimodePropagation {
 public:
  explicit SetupProbe(const SimulationParameters& in) : MultimodePropagation(in) {}
  void computeLawsonRK(const unsigned int) override { initLawsonRK(); }
  ComplexArraysContainer getResult() override { return ComplexArraysContainer(); }
  void setFFTPlans() override {}

  void init() { initLawsonRK(); }
  unsigned stages() const { return RK.s; }
  unsigned pairs() const { return static_cast<unsigned>(unique_pairs.size()); }
  unsigned defaultSteps() const { return static_cast<unsigned>(floor(z_final / delta_z)); }

  void scalars(double* out) const {
    out[0] = t_final;
    out[1] = delta_t;
    out[2] = z_final;
    out[3] = delta_z;
    out[4] = z;
    out[5] = conversion_factor;
    out[6] = sign_gamma;
    out[7] = T0;
  }
  void tableau(double* a, double* b, double* c) const {
    const unsigned s = RK.s;
    for (unsigned i = 0; i < s; i++) {
      for (unsigned j = 0; j < s; j++) a[i * s + j] = (j < i) ? RK.a[i][j] : 0.;
      b[i] = RK.b[i];
      c[i] = RK.c[i];
    }
  }
  void betaTilde(double* out) const {
    unsigned nb = beta_tilde[0].size();
    for (unsigned p = 0; p < M; p++)
      for (unsigned i = 0; i < nb; i++) out[p * nb + i] = beta_tilde[p][i];
  }
  void qTilde(double* out) const {
    unsigned e = 0;
    for (unsigned p = 0; p < M; p++)
      for (unsigned i = 0; i < Q_tilde[p].getNumberOfNonZeroValues(); i++) out[e++] = Q_tilde[p].getNonZeroValue(i).value;
  }
  void uniquePairs(unsigned* out) const {
    for (size_t j = 0; j < unique_pairs.size(); j++) out[j] = unique_pairs[j];
  }
  static void copyContainer(ComplexArraysContainer& c, unsigned M, unsigned nt, double* out) {
    for (unsigned p = 0; p < M; p++)
      for (unsigned j = 0; j < nt; j++) {
        out[2 * (p * nt + j)] = c[p][j].real();
        out[2 * (p * nt + j) + 1] = c[p][j].imag();
      }
  }
  void psiOut(double* out) { copyContainer(psi, M, nt, out); }
  void lOut(double* out) { copyContainer(L, M, nt, out); }
  void e1Out(double* out) { copyContainer(E1, M, nt, out); }
  void eOut(unsigned i, double* out) { copyContainer(E[i], M, nt, out); }
};

SimulationParameters toParameters(const mmp_params* in) {
  SimulationParameters sp;
  sp.n_modes = in->n_modes;
  sp.fiber_length = in->fiber_length;
  for (unsigned p = 0; p < in->n_modes; p++) {
    doubleArray row(in->n_beta);
    for (unsigned i = 0; i < in->n_beta; i++) row[i] = in->dispersion_coefficients[p * in->n_beta + i];
    sp.dispersion_coefficients.push_back(row);
  }
  for (unsigned p = 0; p < in->n_modes; p++) {
    Sparse3DArray q;
    for (unsigned e = in->coupling_offsets[p]; e < in->coupling_offsets[p + 1]; e++) {
      NonZeroValue nzv;
      nzv.index[0] = in->coupling_klm[3 * e];
      nzv.index[1] = in->coupling_klm[3 * e + 1];
      nzv.index[2] = in->coupling_klm[3 * e + 2];
      nzv.value = in->coupling_values[e];
      q.addNonZeroValue(nzv);
    }
    sp.coupling_coefficients.push_back(q);
  }
  sp.nt = in->nt;
  sp.time_window = in->time_window;
  sp.pulse_width = in->pulse_width;
  sp.initial_fields = ComplexArraysContainer(in->nt, in->n_modes);
  for (unsigned p = 0; p < in->n_modes; p++)
    for (unsigned j = 0; j < in->nt; j++)
      sp.initial_fields[p][j] = Complex(in->initial_fields[2 * (p * in->nt + j)], in->initial_fields[2 * (p * in->nt + j) + 1]);
  sp.n_steps = in->n_steps;
  sp.method_order = in->method_order;
  sp.nonlinearity_const = in->nonlinearity_const;
  sp.raman_proportion = in->raman_proportion;
  if (in->raman_response) sp.raman_response.assign(in->raman_response, in->raman_response + in->nt);
  return sp;
}

}  // namespace

extern "C" {

// Runs the reference constructor and initLawsonRK on `in` and copies out:
// counts[0..3] = stages, pairs, default step count, n_beta;
// scalars[8] = t_final, delta_t, z_final, delta_z, z, conversion_factor, sign_gamma, T0;
// a[s*s], b[s], c[s]; beta_tilde[M*n_beta]; q_tilde[nnz]; pairs[<= nnz];
// psi, L, E1 [M][NT] complex; E [s][M][NT] complex. Any output pointer may be NULL.
int ref_probe_setup(const mmp_params* in, unsigned* counts, double* scalars, double* a, double* b, double* c,
                    double* beta_tilde, double* q_tilde, unsigned* pairs, double* psi, double* L, double* E1,
                    double* E) {
  SimulationParameters sp = toParameters(in);
  SetupProbe probe(sp);
  probe.init();
  const unsigned M = in->n_modes, nt = in->nt;
  if (counts) {
    counts[0] = probe.stages();
    counts[1] = probe.pairs();
    counts[2] = probe.defaultSteps();
    counts[3] = in->n_beta;
  }
  if (scalars) probe.scalars(scalars);
  if (a && b && c) probe.tableau(a, b, c);
  if (beta_tilde) probe.betaTilde(beta_tilde);
  if (q_tilde) probe.qTilde(q_tilde);
  if (pairs) probe.uniquePairs(pairs);
  if (psi) probe.psiOut(psi);
  if (L) probe.lOut(L);
  if (E1) probe.e1Out(E1);
  if (E)
    for (unsigned i = 0; i < probe.stages(); i++) probe.eOut(i, E + 2 * (size_t)i * M * nt);
  return 0;
}

}  // extern "C"
