// MultimodePropagation_GPU.cpp -- the drop-in solver classes over the C ABI (include/mmprop.h).
#include "MultimodePropagation_GPU.hpp"

#include <cstdint>
#include <vector>

#include "../../../include/mmprop.h"
#include "FlatParams.hpp"

using mmprop_detail::check;
using mmprop_detail::FlatParams;

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

std::vector<double> MultimodePropagationGPU::getPhotonNumber(double wavelength) {
  std::vector<double> n(M_);
  check(mmp_photon_number(solver_, wavelength, n.data()), "mmp_photon_number");
  return n;
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
