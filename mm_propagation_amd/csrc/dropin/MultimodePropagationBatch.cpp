// MultimodePropagationBatch.cpp -- the batch solver and request server classes over the C ABI.
#include "MultimodePropagationBatch.hpp"

#include <memory>

#include "../../../include/mmprop.h"
#include "FlatParams.hpp"

using mmprop_detail::check;
using mmprop_detail::FlatParams;

namespace {

ComplexArraysContainerCPU unpack(const std::vector<double>& buf, unsigned int M, unsigned int nt) {
  ComplexArraysContainerCPU out(nt, M);
  for (unsigned int p = 0; p < M; p++)
    for (unsigned int j = 0; j < nt; j++)
      out[p][j] = std::complex<double>(buf[2 * ((size_t)p * nt + j)], buf[2 * ((size_t)p * nt + j) + 1]);
  return out;
}

}  // namespace

MultimodePropagationBatchGPU::MultimodePropagationBatchGPU(const std::vector<SimulationParameters>& fibers,
                                                           const MultimodePropagationBatchOptions& opt)
    : ctx_(nullptr), solver_(nullptr), n_((unsigned int)fibers.size()), M_(0), nt_(0) {
  if (fibers.empty()) throw MmpropError(MMP_ERR_INVALID_ARGUMENT, "a batch needs at least one fiber");
  M_ = fibers[0].n_modes;
  nt_ = fibers[0].nt;
  std::vector<std::unique_ptr<FlatParams>> flat;
  std::vector<mmp_params> c;
  for (const SimulationParameters& f : fibers) {
    flat.emplace_back(new FlatParams(f, true));
    c.push_back(flat.back()->c);
  }
  check(mmp_ctx_create(opt.device, &ctx_), "mmp_ctx_create");
  mmp_options o{};
  o.precision = opt.precision;
  o.pair_decode = opt.pair_decode;
  o.flags = opt.flags;
  const mmp_status st = mmp_batch_create(ctx_, n_, c.data(), &o, &solver_);
  if (st != MMP_OK) {
    mmp_ctx_destroy(ctx_);
    ctx_ = nullptr;
    check(st, "mmp_batch_create");
  }
}

MultimodePropagationBatchGPU::~MultimodePropagationBatchGPU() {
  mmp_solver_destroy(solver_);
  mmp_ctx_destroy(ctx_);
}

void MultimodePropagationBatchGPU::computeLawsonRK(unsigned int nz) {
  check(mmp_compute_lawson_rk(solver_, nz), "mmp_compute_lawson_rk");
}

void MultimodePropagationBatchGPU::advance(unsigned int n) { check(mmp_advance(solver_, n), "mmp_advance"); }

ComplexArraysContainerCPU MultimodePropagationBatchGPU::getResult(unsigned int fiber) {
  std::vector<double> buf(2 * (size_t)M_ * nt_);
  check(mmp_fiber_result(solver_, fiber, buf.data()), "mmp_fiber_result");
  return unpack(buf, M_, nt_);
}

std::vector<double> MultimodePropagationBatchGPU::getEnergy(unsigned int fiber) {
  std::vector<double> e(M_);
  check(mmp_fiber_energy(solver_, fiber, e.data()), "mmp_fiber_energy");
  return e;
}

double MultimodePropagationBatchGPU::getZ(unsigned int fiber) {
  double z = 0.;
  check(mmp_fiber_state(solver_, fiber, nullptr, &z), "mmp_fiber_state");
  return z;
}

void MultimodePropagationBatchGPU::load(unsigned int fiber, const SimulationParameters& in) {
  FlatParams fp(in, true);
  check(mmp_fiber_load(solver_, fiber, &fp.c), "mmp_fiber_load");
}

void MultimodePropagationBatchGPU::setLive(unsigned int fiber, bool live) {
  check(mmp_fiber_set_live(solver_, fiber, live ? 1u : 0u), "mmp_fiber_set_live");
}

PropagationServer::PropagationServer(const Options& opt) : server_(nullptr) {
  mmp_serve_options o{};
  o.device = opt.device;
  o.slots = opt.slots;
  o.max_chunk = opt.max_chunk;
  o.precision = opt.precision;
  o.pair_decode = opt.pair_decode;
  o.flags = opt.flags;
  o.hbm_budget = opt.hbm_budget;
  check(mmp_serve_create(&o, &server_), "mmp_serve_create");
}

PropagationServer::~PropagationServer() { mmp_serve_destroy(server_); }

uint64_t PropagationServer::submit(const SimulationParameters& in, unsigned int n_steps, bool raman) {
  FlatParams fp(in, raman);  // the server keeps its own copy
  uint64_t id = 0;
  check(mmp_serve_submit(server_, &fp.c, n_steps, &id), "mmp_serve_submit");
  if (shapes_.size() <= id) shapes_.resize(id + 1);
  shapes_[id] = {in.n_modes, in.nt};
  return id;
}

unsigned int PropagationServer::poll() {
  uint32_t active = 0;
  check(mmp_serve_poll(server_, &active), "mmp_serve_poll");
  return active;
}

void PropagationServer::drain() { check(mmp_serve_drain(server_), "mmp_serve_drain"); }

PropagationServer::Record PropagationServer::record(uint64_t id) {
  mmp_serve_record r{};
  check(mmp_serve_record_get(server_, id, &r), "mmp_serve_record_get");
  return Record{r.state, r.batched != 0, r.group, r.slot, r.t_submit, r.t_admit, r.t_finish};
}

ComplexArraysContainerCPU PropagationServer::getResult(uint64_t id) {
  if (id >= shapes_.size()) throw MmpropError(MMP_ERR_INVALID_ARGUMENT, "unknown request id");
  const unsigned int M = shapes_[id].first, nt = shapes_[id].second;
  std::vector<double> buf(2 * (size_t)M * nt);
  check(mmp_serve_result(server_, id, buf.data()), "mmp_serve_result");
  return unpack(buf, M, nt);
}
