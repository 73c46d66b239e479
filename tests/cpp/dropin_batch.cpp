// dropin_batch.cpp -- the batch solver and the request server through the C++ drop-in API:
// three 2-mode Gaussian fibers (different powers and lengths) solved as one batch and as
// requests of a server, each compared bit for bit with the same fiber solved alone by
// MultimodePropagationGPU_RamanON.   usage: dropin_batch   (exit 0 when every result matches)
#include <cmath>
#include <cstdio>

#include "MultimodePropagationBatch.hpp"

static SimulationParameters fiber(unsigned int nt, double scale, double length) {
  const double SR = 6.4341e12, T0 = 0.1, tw = 10.0;
  SimulationParameters in;
  in.n_modes = 2;
  in.fiber_length = length;
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
  const double dt = tw / nt, P[2] = {100. * scale, 10. * scale};
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
  return in;
}

static bool same(const ComplexArraysContainerCPU& a, const ComplexArraysContainerCPU& b, unsigned int M,
                 unsigned int nt) {
  for (unsigned int p = 0; p < M; p++)
    for (unsigned int j = 0; j < nt; j++)
      if (a[p][j] != b[p][j]) return false;
  return true;
}

static ComplexArraysContainerCPU alone(const SimulationParameters& in, unsigned int steps) {
  MultimodePropagationGPU_RamanON s(in);  // FP64, the GPU class's pair decode (by pair value)
  s.computeLawsonRK(steps);
  return s.getResult();
}

int main() {
  const unsigned int nt = 4096;
  std::vector<SimulationParameters> fibers = {fiber(nt, 1.0, 0.5), fiber(nt, 0.5, 0.4), fiber(nt, 2.0, 0.3)};
  const unsigned int steps[3] = {40, 24, 56};
  int bad = 0;
  try {
    MultimodePropagationBatchGPU batch(fibers);
    batch.computeLawsonRK(40);
    for (unsigned int f = 0; f < 3; f++)
      if (!same(batch.getResult(f), alone(fibers[f], 40), 2, nt)) {
        std::fprintf(stderr, "batch fiber %u differs from its solve alone\n", f);
        bad++;
      }
    PropagationServer::Options o;
    o.slots = 2;  // the third request waits for a free slot
    o.max_chunk = 16;
    o.precision = 0;
    o.pair_decode = 1;
    PropagationServer server(o);
    uint64_t id[3];
    for (unsigned int f = 0; f < 3; f++) id[f] = server.submit(fibers[f], steps[f]);
    server.drain();
    for (unsigned int f = 0; f < 3; f++) {
      const PropagationServer::Record r = server.record(id[f]);
      if (r.state != PropagationServer::Done || !r.batched || !(r.t_finish >= r.t_admit && r.t_admit >= r.t_submit)) {
        std::fprintf(stderr, "request %u: state %u batched %d\n", f, r.state, (int)r.batched);
        bad++;
      }
      if (!same(server.getResult(id[f]), alone(fibers[f], steps[f]), 2, nt)) {
        std::fprintf(stderr, "request %u differs from its solve alone\n", f);
        bad++;
      }
    }
  } catch (const MmpropError& e) {
    std::fprintf(stderr, "error %d: %s\n", e.status(), e.what());
    return 1;
  }
  if (!bad) std::printf("batch and server results equal the solves alone\n");
  return bad ? 1 : 0;
}
