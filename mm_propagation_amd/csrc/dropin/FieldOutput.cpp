// FieldOutput.cpp -- see include/mmprop/FieldOutput.hpp.
#include "FieldOutput.hpp"

#include <complex>
#include <stdexcept>

void writeFieldMagnitudes(std::ostream& os, const ComplexArraysContainer& initial,
                          const ComplexArraysContainer& result) {
  const unsigned int n = initial.getArraySize();
  if (result.getArraySize() != n) throw std::invalid_argument("writeFieldMagnitudes: sizes differ");
  const unsigned int mi = initial.getNumberOfArrays(), mr = result.getNumberOfArrays();
  for (unsigned int j = 0; j < n; j++) {
    bool first = true;
    for (unsigned int p = 0; p < mi; p++, first = false) os << (first ? "" : " ; ") << std::abs(initial[p][j]);
    for (unsigned int p = 0; p < mr; p++, first = false) os << (first ? "" : " ; ") << std::abs(result[p][j]);
    os << std::endl;
  }
}
