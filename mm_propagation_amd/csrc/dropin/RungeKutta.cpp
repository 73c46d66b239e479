// RungeKutta.cpp -- Butcher tableaux of the drop-in API. The three tableaux the solver uses
// come from the library (mmp_tableau) so there is one definition of their coefficients.
#include "RungeKutta.hpp"

#include <cmath>

#include "../../../include/mmprop.h"

static void from_library(RungeKuttaCoefficients& RK, uint32_t order) {
  uint32_t s = 0;
  double a[121], b[11], c[11];
  mmp_tableau(order, &s, a, b, c);
  RK.s = s;
  RK.a.clear();
  RK.b.assign(b, b + s);
  RK.c.assign(c, c + s);
  RK.a.push_back(doubleArray(1, 0.));
  for (uint32_t i = 1; i < s; i++) RK.a.push_back(doubleArray(a + i * 11, a + i * 11 + i));
}

void Butcher4(RungeKuttaCoefficients& RK) { from_library(RK, 4); }
void Butcher11(RungeKuttaCoefficients& RK) { from_library(RK, 8); }

void Butcher7(RungeKuttaCoefficients& RK, const double nu) {
  if (nu == 1.) {
    from_library(RK, 6);
    return;
  }
  // the nu-family of 7-stage order-6 methods for other nu (src/RungeKutta.cpp:24-72)
  const double r = std::sqrt(21.);
  RK.s = 7;
  RK.a = {{0.},
          {nu},
          {(4. * nu - 1.) / (8. * nu), 1. / (8. * nu)},
          {(10. * nu - 2.) / (27. * nu), 2. / (27. * nu), (8. * nu) / (27. * nu)},
          {(56. - 77. * nu + (8. - 17. * nu) * r) / (392. * nu), (-8. * (7. + r)) / (392. * nu),
           (48. * (7. + r) * nu) / (392. * nu), (-3. * (21. + r) * nu) / (392. * nu)},
          {(-5. * (287. * nu - 56. - (59. * nu - 8.) * r)) / (1960. * nu), (-40. * (7. - r)) / (1960. * nu),
           (320. * r * nu) / (1960. * nu), (3. * (21. - 121. * r) * nu) / (1960. * nu), (392. * (6. - r) * nu) / (1960. * nu)},
          {(15. * ((30. * nu - 8.) - 7. * nu * r)) / (180. * nu), 120. / (180. * nu), (-40. * (5. + 7. * r) * nu) / (180. * nu),
           (63. * (2. + 3. * r) * nu) / (180. * nu), (-14. * (49. - 9. * r) * nu) / (180. * nu), (70. * (7. + r) * nu) / (180. * nu)}};
  RK.b = {9 / 180., 0., 64 / 180., 0., 49 / 180., 49 / 180., 9 / 180.};
  RK.c = {0., nu, 1 / 2., 2 / 3., (7. + r) / 14., (7. - r) / 14., 1.};
}

void Butcher7(RungeKuttaCoefficients& RK, const double lambda, const double mu) {
  // the (lambda, mu) family of 7-stage methods of the reference's API (src/RungeKutta.cpp:75-122)
  RK.s = 7;
  RK.a = {{0.},
          {mu},
          {2 / 3. - 2. / (9. * mu), 2. / (9. * mu)},
          {5 / 12. - 1. / (9 * mu), 1. / (9 * mu), -1 / 12.},
          {17 / 16. - 3. / (8. * mu), 3. / (8. * mu), -3 / 16., -3 / 8.},
          {17 / 16. - 3. / (8 * mu) + 1. / (4. * lambda), 3. / (8. * mu), -3 / 16. - 3. / (4. * lambda),
           -3 / 8. - 3. / (2. * lambda), 2. / lambda},
          {-27 / 44. + 3. / (11. * mu), -3. / (11. * mu), 63 / 44., 18 / 11., 4. * ((lambda - 4.) / 11.), -(4. * lambda) / 11.}};
  RK.b = {1 / 120., 0., 27 / 40., 27 / 40., (lambda - 8.) / 15., -lambda / 15., 11 / 120.};
  RK.c = {0., mu, 2 / 3., 1 / 3., 1 / 2., 1 / 2., 1.};
}
