// MultimodePropagation_GPU.hpp -- the drop-in solver classes for gfx950
// (names, constructors and methods of include/MultimodePropagation_GPU.cuh:10-85 of
// MM_Propagation), implemented over the C ABI of include/mmprop.h.
//
// Errors are reported by throwing MmpropError (the reference has no error reporting);
// destruction releases the device memory (the destructor is virtual through the base).
#ifndef MMPROP_MULTIMODE_PROPAGATION_GPU_HPP
#define MMPROP_MULTIMODE_PROPAGATION_GPU_HPP

#include <stdexcept>
#include <string>

#include "MultimodePropagation.hpp"

struct mmp_ctx;
struct mmp_solver;

class MmpropError : public std::runtime_error {
 public:
  MmpropError(int status, const std::string& what) : std::runtime_error(what), status_(status) {}
  int status() const { return status_; }

 private:
  int status_;
};

// Choices with no counterpart in the reference. The default pair decode is the one of the class
// these replace: MultimodePropagationGPU_* reads slot j's modes from the stored pair code
// (l = pair % M, m = pair / M; src/MultimodePropagation_GPU.cu:188-190,278-280). The reference
// CPU class's counter decode (j % M, j / M; src/MultimodePropagation_CPU.cpp:126-133) is
// pair_decode = 0.
struct MultimodePropagationGPUOptions {
  unsigned int pair_decode = 1;  // MMP_PAIRS_PHYSICAL (1) or MMP_PAIRS_REFERENCE_CPU (0)
  unsigned int flags = 0;        // MMP_FLAG_* of include/mmprop.h
  int device = 0;
};

class MultimodePropagationGPU : public MultimodePropagation {
 public:
  ~MultimodePropagationGPU() override;
  MultimodePropagationGPU(const MultimodePropagationGPU&) = delete;
  MultimodePropagationGPU& operator=(const MultimodePropagationGPU&) = delete;

  void computeLawsonRK(const unsigned int nz) override;
  void computeLawsonRK() override;
  ComplexArraysContainerCPU getResult() override;

  // Extensions: resumable stepping, per-mode energy (W ps), per-mode photon number for a carrier
  // wavelength in metres, current z.
  void advance(const unsigned int n);
  std::vector<double> getEnergy();
  std::vector<double> getPhotonNumber(double wavelength);
  double getZ();

  // The MEX gateway's n_save loop (matlab/mm_propagation.cpp:61-101): n_steps in n_save chunks
  // (the last takes the remainder), the physical field and the distance Z (m) after each.
  // resumable = false restarts z and the Lawson tables per chunk as the gateway does; true runs
  // one solve paused at each snapshot.
  struct Snapshot {
    ComplexArraysContainerCPU fields;
    double z;
  };
  std::vector<Snapshot> propagateWithSnapshots(unsigned int n_steps, unsigned int n_save, bool resumable = false);

 protected:
  MultimodePropagationGPU(const SimulationParameters& in, bool raman, const MultimodePropagationGPUOptions& opt);

 private:
  mmp_ctx* ctx_;
  mmp_solver* solver_;
  unsigned int M_, nt_;
};

class MultimodePropagationGPU_RamanOFF : public MultimodePropagationGPU {
 public:
  explicit MultimodePropagationGPU_RamanOFF(const SimulationParameters& in,
                                            const MultimodePropagationGPUOptions& opt = MultimodePropagationGPUOptions())
      : MultimodePropagationGPU(in, false, opt) {}
};

class MultimodePropagationGPU_RamanON : public MultimodePropagationGPU {
 public:
  explicit MultimodePropagationGPU_RamanON(const SimulationParameters& in,
                                           const MultimodePropagationGPUOptions& opt = MultimodePropagationGPUOptions())
      : MultimodePropagationGPU(in, true, opt) {}
};

#endif
