// MultimodePropagation.hpp -- solver input record and abstract solver interface
// (include/MultimodePropagation.hpp:16-94 of MM_Propagation). SimulationParameters keeps
// the reference's fields and units; InputParameters is the README's name for it.
#ifndef MMPROP_MULTIMODE_PROPAGATION_HPP
#define MMPROP_MULTIMODE_PROPAGATION_HPP

#include <vector>

#include "ComplexArraysContainer.hpp"
#include "RungeKutta.hpp"
#include "Sparse3DArray.hpp"

struct SimulationParameters {
  unsigned int n_modes;
  double fiber_length;
  std::vector<doubleArray> dispersion_coefficients;
  std::vector<Sparse3DArray> coupling_coefficients;

  unsigned int nt;
  double time_window;
  double pulse_width;
  ComplexArraysContainer initial_fields;

  unsigned int n_steps;
  unsigned int method_order;
  double nonlinearity_const;
  double raman_proportion;
  doubleArray raman_response;
};

typedef SimulationParameters InputParameters;

class MultimodePropagation {
 public:
  virtual ~MultimodePropagation();

  // Restart at z = 0 (fresh Lawson operators, current field kept) and take nz steps.
  virtual void computeLawsonRK(const unsigned int nz) = 0;
  // computeLawsonRK(floor(z_final / delta_z)).
  virtual void computeLawsonRK() = 0;
  // Physical field A = F^-1[exp(L z) F psi] / conversion_factor.
  virtual ComplexArraysContainer getResult() = 0;
};

#endif
