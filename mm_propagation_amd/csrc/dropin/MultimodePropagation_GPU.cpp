// MultimodePropagation_GPU.cpp -- the drop-in solver classes over the C ABI (include/mmprop.h).
#include "MultimodePropagation_GPU.hpp"

#include <cstdint>
#include <vector>

#include "../../../include/mmprop.h"

namespace {

void check(mmp_status st, const char* where) {
  if (st != MMP_OK)
    throw MmpropError((int)st, std::string(where) + ": " + mmp_status_string(st) + " (" + mmp_last_error() + ")");
}

// Flat views of a SimulationParameters (the C ABI's mmp_params keeps pointers into these).
struct FlatParams {
  std::vector<double> beta, values, fields, response;
  std::vector<uint32_t> offsets, klm;
  mmp_params c{};

  FlatParams(const SimulationParameters& in, bool raman) {
    const unsigned int M = in.n_modes, nt = in.nt;
    if (in.dispersion_coefficients.size() != M || in.coupling_coefficients.size() != M ||
        in.initial_fields.getNumberOfArrays() != M || in.initial_fields.getArraySize() != nt)
      throw MmpropError(MMP_ERR_INVALID_ARGUMENT, "SimulationParameters sizes do not match n_modes and nt");
    const unsigned int nb = M ? (unsigned int)in.dispersion_coefficients[0].size() : 0;
    for (unsigned int p = 0; p < M; p++) {
      if (in.dispersion_coefficients[p].size() != nb)
        throw MmpropError(MMP_ERR_INVALID_ARGUMENT, "every mode needs the same number of dispersion orders");
      beta.insert(beta.end(), in.dispersion_coefficients[p].begin(), in.dispersion_coefficients[p].end());
    }
    offsets.push_back(0);
    for (unsigned int p = 0; p < M; p++) {
      const Sparse3DArray& q = in.coupling_coefficients[p];
      for (unsigned int e = 0; e < q.getNumberOfNonZeroValues(); e++) {
        const NonZeroValue v = q.getNonZeroValue(e);
        klm.insert(klm.end(), {v.index[0], v.index[1], v.index[2]});
        values.push_back(v.value);
      }
      offsets.push_back((uint32_t)values.size());
    }
    fields.resize(2 * (size_t)M * nt);
    for (unsigned int p = 0; p < M; p++)
      for (unsigned int j = 0; j < nt; j++) {
        const std::complex<double> z = in.initial_fields.getValue(p, j);
        fields[2 * ((size_t)p * nt + j)] = z.real();
        fields[2 * ((size_t)p * nt + j) + 1] = z.imag();
      }
    if (raman) {
      if (in.raman_response.size() != nt)
        throw MmpropError(MMP_ERR_INVALID_ARGUMENT, "raman_response must have nt samples");
      response = in.raman_response;
    }
    c.n_modes = M;
    c.fiber_length = in.fiber_length;
    c.n_beta = nb;
    c.dispersion_coefficients = beta.data();
    c.coupling_offsets = offsets.data();
    c.coupling_klm = klm.empty() ? nullptr : klm.data();
    c.coupling_values = values.empty() ? nullptr : values.data();
    c.nt = nt;
    c.time_window = in.time_window;
    c.pulse_width = in.pulse_width;
    c.initial_fields = fields.data();
    c.n_steps = in.n_steps;
    c.method_order = in.method_order;
    c.nonlinearity_const = in.nonlinearity_const;
    c.raman_proportion = in.raman_proportion;
    c.raman_response = raman ? response.data() : nullptr;
    c.raman = raman ? 1u : 0u;
  }
};

}  // namespace

MultimodePropagationGPU::MultimodePropagationGPU(const SimulationParameters& in, bool raman,
                                                 const MultimodePropagationGPUOptions& opt)
    : ctx_(nullptr), solver_(nullptr), M_(in.n_modes), nt_(in.nt) {
  FlatParams fp(in, raman);
  check(mmp_ctx_create(opt.device, &ctx_), "mmp_ctx_create");
  mmp_options o{};
  o.precision = MMP_PRECISION_FP64;
  o.pair_decode = opt.pair_decode;
  o.flags = opt.flags;
  const mmp_status st = mmp_solver_create(ctx_, &fp.c, &o, &solver_);
  if (st != MMP_OK) {
    mmp_ctx_destroy(ctx_);
    ctx_ = nullptr;
    check(st, "mmp_solver_create");
  }
}

MultimodePropagationGPU::~MultimodePropagationGPU() {
  mmp_solver_destroy(solver_);
  mmp_ctx_destroy(ctx_);
}

void MultimodePropagationGPU::computeLawsonRK(const unsigned int nz) {
  check(mmp_compute_lawson_rk(solver_, nz), "mmp_compute_lawson_rk");
}

void MultimodePropagationGPU::computeLawsonRK() {
  check(mmp_compute_lawson_rk_default(solver_), "mmp_compute_lawson_rk_default");
}

ComplexArraysContainerCPU MultimodePropagationGPU::getResult() {
  std::vector<double> buf(2 * (size_t)M_ * nt_);
  check(mmp_get_result(solver_, buf.data()), "mmp_get_result");
  ComplexArraysContainerCPU out(nt_, M_);
  for (unsigned int p = 0; p < M_; p++)
    for (unsigned int j = 0; j < nt_; j++)
      out[p][j] = std::complex<double>(buf[2 * ((size_t)p * nt_ + j)], buf[2 * ((size_t)p * nt_ + j) + 1]);
  return out;
}

void MultimodePropagationGPU::advance(const unsigned int n) { check(mmp_advance(solver_, n), "mmp_advance"); }

std::vector<double> MultimodePropagationGPU::getEnergy() {
  std::vector<double> e(M_);
  check(mmp_energy(solver_, e.data()), "mmp_energy");
  return e;
}

std::vector<MultimodePropagationGPU::Snapshot> MultimodePropagationGPU::propagateWithSnapshots(
    unsigned int n_steps, unsigned int n_save, bool resumable) {
  const size_t per = 2 * (size_t)M_ * nt_;
  std::vector<double> buf(per * (n_save ? n_save : 1)), z(n_save ? n_save : 1);
  check(mmp_propagate_snapshots(solver_, n_steps, n_save,
                                resumable ? MMP_SNAPSHOTS_RESUMABLE : MMP_SNAPSHOTS_REFERENCE_RESTART, buf.data(),
                                z.data()),
        "mmp_propagate_snapshots");
  std::vector<Snapshot> out;
  for (unsigned int i = 0; i < n_save; i++) {
    Snapshot sn{ComplexArraysContainerCPU(nt_, M_), z[i]};
    for (unsigned int p = 0; p < M_; p++)
      for (unsigned int j = 0; j < nt_; j++) {
        const size_t k = i * per + 2 * ((size_t)p * nt_ + j);
        sn.fields[p][j] = std::complex<double>(buf[k], buf[k + 1]);
      }
    out.push_back(sn);
  }
  return out;
}

double MultimodePropagationGPU::getZ() {
  double z = 0.;
  check(mmp_get_state(solver_, nullptr, &z), "mmp_get_state");
  return z;
}
