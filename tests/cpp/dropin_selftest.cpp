// dropin_selftest.cpp -- host-only checks of the drop-in containers and tableaux (no device):
// element-wise arithmetic, conj/exp/abs/norm/sum, the container layout, insertion order of
// Sparse3DArray, and the Butcher tableaux read through the C ABI. Exit status 0 = pass.
#include <cmath>
#include <cstdio>

#include <sstream>

#include "FieldOutput.hpp"
#include "MultimodePropagation.hpp"

static int failures = 0;
#define EXPECT(c)                                                        \
  do {                                                                   \
    if (!(c)) {                                                          \
      std::fprintf(stderr, "%s:%d: expectation failed: %s\n", __FILE__, __LINE__, #c); \
      failures++;                                                        \
    }                                                                    \
  } while (0)

int main() {
  ComplexArray a(3, Complex(1., 2.)), b(3);
  b[0] = Complex(0., 1.);
  b[1] = Complex(2., 0.);
  b[2] = Complex(-1., -1.);
  ComplexArray c = a * b;
  EXPECT(c[0] == Complex(-2., 1.));
  EXPECT(c[1] == Complex(2., 4.));
  EXPECT(c[2] == Complex(1., -3.));
  ComplexArray d = a + b;
  d -= b;
  EXPECT(d[2] == Complex(1., 2.));
  d *= 2.;
  d /= Complex(0., 1.);
  EXPECT(d[0] == Complex(4., -2.));
  ComplexArray e = exp(ComplexArray(2, Complex(0., M_PI)));
  EXPECT(std::abs(e[1] + 1.) < 1e-15);
  EXPECT(conj(a)[1] == Complex(1., -2.));
  EXPECT(std::abs(norm(a) - std::sqrt(15.)) < 1e-15);
  EXPECT(std::abs(sum(abs(b)) - (1. + 2. + std::sqrt(2.))) < 1e-15);

  ComplexArraysContainer box(4, 3);
  EXPECT(box.getNumberOfArrays() == 3 && box.getArraySize() == 4);
  box[2][3] = Complex(5., 6.);
  EXPECT(box.getValue(2, 3) == Complex(5., 6.));
  ComplexArraysContainer other(4, Complex(1., 1.), 3);
  box = other;
  EXPECT(box.getValue(2, 3) == Complex(1., 1.));

  Sparse3DArray q;
  NonZeroValue v = {{0, 1, 1}, 2.5};
  q.addNonZeroValue(v);
  v = {{1, 0, 0}, -1.};
  q.addNonZeroValue(v);
  q.resetValue(1, 3.);
  EXPECT(q.getNumberOfNonZeroValues() == 2 && q.getNonZeroValue(0).index[2] == 1 && q.getNonZeroValue(1).value == 3.);

  RungeKuttaCoefficients rk4, rk6, rk8, rk6b;
  Butcher4(rk4);
  Butcher7(rk6, 1.);
  Butcher11(rk8);
  Butcher7(rk6b, 2., 0.5);
  EXPECT(rk4.s == 4 && rk6.s == 7 && rk8.s == 11 && rk6b.s == 7);
  for (const RungeKuttaCoefficients* r : {&rk4, &rk6, &rk8}) {
    double sb = 0.;
    for (double x : r->b) sb += x;
    EXPECT(std::abs(sb - 1.) < 1e-15);
    for (unsigned int i = 1; i < r->s; i++) {
      double row = 0.;
      for (double x : r->a[i]) row += x;
      EXPECT(std::abs(row - r->c[i]) < 1e-14);
    }
  }
  EXPECT(rk4.a[1][0] == 0.5 && rk4.b[0] == 1 / 6.);
  // text output of simulations/main.cpp:101-103 for two modes
  ComplexArraysContainer fin(3, 2), fout(3, 2);
  fin[0][1] = Complex(3., 4.);
  fin[1][1] = Complex(0., 2.);
  fout[0][1] = Complex(-1., 0.);
  fout[1][1] = Complex(0.5, 0.);
  std::ostringstream os;
  writeFieldMagnitudes(os, fin, fout);
  EXPECT(os.str() == "0 ; 0 ; 0 ; 0\n5 ; 2 ; 1 ; 0.5\n0 ; 0 ; 0 ; 0\n");
  if (failures) std::fprintf(stderr, "%d failure(s)\n", failures);
  else std::printf("dropin selftest passed\n");
  return failures ? 1 : 0;
}
