// dropin_propagate.cpp -- drives the drop-in C++ API the way the reference's
// simulations/main.cpp drives its solver (SimulationParameters, Sparse3DArray,
// MultimodePropagation*, computeLawsonRK(n), getResult()) on the 2-mode Gaussian
// workload, then writes the fields as raw interleaved doubles [M][NT][2].
//   usage: dropin_propagate <nt> <steps> <out.bin> [raman_off]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>

#include "MultimodePropagation_GPU.hpp"

int main(int argc, char** argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s nt steps out.bin [raman_off]\n", argv[0]);
    return 2;
  }
  const unsigned int nt = (unsigned int)std::atoi(argv[1]);
  const unsigned int steps = (unsigned int)std::atoi(argv[2]);
  const bool raman = !(argc > 4 && std::string(argv[4]) == "raman_off");
  const double SR = 6.4341e12, T0 = 0.1, tw = 10.0;

  SimulationParameters in;
  in.n_modes = 2;
  in.fiber_length = 0.5;
  in.dispersion_coefficients.push_back({0., 0., 0.02});
  in.dispersion_coefficients.push_back({-2.0e3, 1.0e-3, 0.02});
  Sparse3DArray Q0, Q1;
  NonZeroValue nzv;
  nzv = {{0, 0, 0}, SR}; Q0.addNonZeroValue(nzv);
  nzv = {{0, 0, 1}, SR / 3}; Q0.addNonZeroValue(nzv);
  nzv = {{0, 1, 1}, SR / 2}; Q0.addNonZeroValue(nzv);
  nzv = {{1, 1, 1}, -SR / 3}; Q1.addNonZeroValue(nzv);
  nzv = {{0, 0, 1}, SR / 3}; Q1.addNonZeroValue(nzv);
  nzv = {{0, 1, 1}, SR / 2}; Q1.addNonZeroValue(nzv);
  in.coupling_coefficients.push_back(Q0);
  in.coupling_coefficients.push_back(Q1);
  in.nt = nt;
  in.time_window = tw;
  in.pulse_width = T0;
  in.initial_fields = ComplexArraysContainer(nt, 2);
  const double dt = tw / nt, P[2] = {100., 10.};
  for (unsigned int p = 0; p < 2; p++)
    for (unsigned int j = 0; j < nt; j++) {
      const double t = ((double)j - (double)(nt / 2)) * dt;
      in.initial_fields[p][j] = std::sqrt(P[p]) * std::exp(-t * t / (2.0 * T0 * T0));
    }
  in.n_steps = 5000;
  in.method_order = 4;
  in.nonlinearity_const = 1.4e-13;
  in.raman_proportion = 0.18;
  in.raman_response = std::vector<double>(nt);
  const double tau1 = 12.2e-3, tau2 = 32e-3;
  const double a = (1. / (tau1 * tau1) + 1. / (tau2 * tau2)) * tau1;
  for (unsigned int i = 0; i < nt; i++) {
    const double t = i * dt;
    in.raman_response[i] = a * (std::exp(-t / tau2) * std::sin(t / tau1));
  }

  try {
    MultimodePropagation* solver = raman ? static_cast<MultimodePropagation*>(new MultimodePropagationGPU_RamanON(in))
                                         : static_cast<MultimodePropagation*>(new MultimodePropagationGPU_RamanOFF(in));
    solver->computeLawsonRK(steps);
    ComplexArraysContainer out(nt, 2);
    out = solver->getResult();
    delete solver;  // virtual destructor releases the device solver
    std::ofstream f(argv[3], std::ios::binary);
    for (unsigned int p = 0; p < 2; p++)
      for (unsigned int j = 0; j < nt; j++) {
        const double v[2] = {out[p][j].real(), out[p][j].imag()};
        f.write(reinterpret_cast<const char*>(v), sizeof v);
      }
  } catch (const MmpropError& e) {
    std::fprintf(stderr, "error %d: %s\n", e.status(), e.what());
    return 1;
  }
  return 0;
}
