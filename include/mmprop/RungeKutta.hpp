// RungeKutta.hpp -- Butcher tableaux (include/RungeKutta.hpp:7-22 of MM_Propagation).
// Butcher4, Butcher7(nu) and Butcher11 return the tableaux the solver uses (read from the
// C ABI's mmp_tableau); Butcher7(lambda, mu) is the second 7-stage family of the
// reference's API, which the solver never selects.
#ifndef MMPROP_RUNGE_KUTTA_HPP
#define MMPROP_RUNGE_KUTTA_HPP

#include <vector>

typedef std::vector<double> doubleArray;

struct RungeKuttaCoefficients {
  unsigned int s;
  std::vector<doubleArray> a;  // row i holds a_ij, j < i (row 0 holds one zero)
  doubleArray b;
  doubleArray c;
};

void Butcher4(RungeKuttaCoefficients& RK);
void Butcher7(RungeKuttaCoefficients& RK, const double nu);
void Butcher7(RungeKuttaCoefficients& RK, const double lambda, const double mu);
void Butcher11(RungeKuttaCoefficients& RK);

#endif
