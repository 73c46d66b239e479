// MultimodePropagationBatch.hpp -- C++ faces of the batch solver and the request server of
// include/mmprop.h (no counterpart in MM_Propagation, which solves one fiber per object).
// Same conventions as MultimodePropagation_GPU.hpp: SimulationParameters in,
// ComplexArraysContainerCPU out, MmpropError on any failure.
#ifndef MMPROP_MULTIMODE_PROPAGATION_BATCH_HPP
#define MMPROP_MULTIMODE_PROPAGATION_BATCH_HPP

#include <cstdint>
#include <vector>

#include "MultimodePropagation_GPU.hpp"

struct mmp_server;

struct MultimodePropagationBatchOptions {
  unsigned int precision = 0;    // MMP_PRECISION_FP64 (0) or MMP_PRECISION_MIXED (1)
  unsigned int pair_decode = 1;  // MMP_PAIRS_PHYSICAL (1), as MultimodePropagationGPU_*, or MMP_PAIRS_REFERENCE_CPU (0)
  unsigned int flags = 0;        // MMP_FLAG_* (e.g. MMP_FLAG_AUTOTUNE)
  int device = 0;
};

// Fibers that share n_modes, nt, method_order and the coupling pattern, advanced together
// (Raman on, nt >= 4096, n_modes <= 16). Every fiber's result equals the same fiber solved by
// MultimodePropagationGPU_RamanON with the same options.
class MultimodePropagationBatchGPU {
 public:
  explicit MultimodePropagationBatchGPU(const std::vector<SimulationParameters>& fibers,
                                        const MultimodePropagationBatchOptions& opt = MultimodePropagationBatchOptions());
  ~MultimodePropagationBatchGPU();
  MultimodePropagationBatchGPU(const MultimodePropagationBatchGPU&) = delete;
  MultimodePropagationBatchGPU& operator=(const MultimodePropagationBatchGPU&) = delete;

  unsigned int size() const { return n_; }
  void computeLawsonRK(unsigned int nz);  // every fiber restarts at z = 0 and takes nz steps
  void advance(unsigned int n);           // live fibers take n more steps
  ComplexArraysContainerCPU getResult(unsigned int fiber);
  std::vector<double> getEnergy(unsigned int fiber);
  double getZ(unsigned int fiber);
  void load(unsigned int fiber, const SimulationParameters& in);  // a new problem of the same structure, z = 0
  void setLive(unsigned int fiber, bool live);                     // parked fibers are not advanced

 private:
  mmp_ctx* ctx_;
  mmp_solver* solver_;
  unsigned int n_, M_, nt_;
};

struct PropagationServerOptions {
  int device = 0;
  unsigned int slots = 16;       // fiber slots per batch group
  unsigned int max_chunk = 256;  // steps between scheduling points
  unsigned int precision = 1;    // MMP_PRECISION_MIXED
  unsigned int pair_decode = 1;  // MMP_PAIRS_PHYSICAL
  unsigned int flags = 0;
  double hbm_budget = 64e9;      // bytes
};

// Continuous batching of solves of mixed n_modes, nt and step counts on one device.
class PropagationServer {
 public:
  using Options = PropagationServerOptions;
  enum State { Queued = 0, Running = 1, Done = 2, Nonfinite = 3 };  // MMP_REQ_* of include/mmprop.h
  struct Record {
    unsigned int state;           // State
    bool batched;
    unsigned int group, slot;
    double t_submit, t_admit, t_finish;  // seconds since the server was made
  };

  explicit PropagationServer(const Options& opt = Options());
  ~PropagationServer();
  PropagationServer(const PropagationServer&) = delete;
  PropagationServer& operator=(const PropagationServer&) = delete;

  uint64_t submit(const SimulationParameters& in, unsigned int n_steps, bool raman = true);
  unsigned int poll();  // one scheduling pass; returns queued + running
  void drain();         // until nothing is queued or running
  Record record(uint64_t id);
  ComplexArraysContainerCPU getResult(uint64_t id);  // MmpropError until done, or when not finite

 private:
  mmp_server* server_;
  std::vector<std::pair<unsigned int, unsigned int>> shapes_;  // (n_modes, nt) per request
};

#endif
