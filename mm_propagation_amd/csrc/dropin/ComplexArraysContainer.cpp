// ComplexArraysContainer.cpp -- M host ComplexArrays (interface of the reference's
// src/ComplexArraysContainer.cpp).
#include "ComplexArraysContainer.hpp"

ComplexArraysContainer::ComplexArraysContainer(unsigned int arraySize, unsigned int numArrays)
    : M(numArrays), n(arraySize), arrays(numArrays, ComplexArray(arraySize)) {}

ComplexArraysContainer::ComplexArraysContainer(unsigned int arraySize, std::complex<double> value,
                                               unsigned int numArrays)
    : M(numArrays), n(arraySize), arrays(numArrays, ComplexArray(arraySize, value)) {}

unsigned int ComplexArraysContainer::getNumberOfArrays() const { return M; }
unsigned int ComplexArraysContainer::getArraySize() const { return n; }
std::complex<double> ComplexArraysContainer::getValue(const unsigned int arrayIndex, const unsigned int i) const {
  return arrays[arrayIndex][i];
}
ComplexArray& ComplexArraysContainer::operator[](unsigned int i) { return arrays[i]; }
ComplexArray& ComplexArraysContainer::operator()(unsigned int i) { return arrays[i]; }
const ComplexArray& ComplexArraysContainer::operator[](unsigned int i) const { return arrays[i]; }
std::vector<ComplexArray> ComplexArraysContainer::operator()() const { return arrays; }
ComplexArraysContainer::operator std::vector<ComplexArray>() const { return arrays; }
