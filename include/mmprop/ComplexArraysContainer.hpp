// ComplexArraysContainer.hpp -- M host ComplexArrays of NT values each, with the reference's
// interface (include/ComplexArraysContainer.hpp:6-28 of MM_Propagation).
#ifndef MMPROP_COMPLEX_ARRAYS_CONTAINER_HPP
#define MMPROP_COMPLEX_ARRAYS_CONTAINER_HPP

#include "ComplexArray.hpp"

class ComplexArraysContainer {
 private:
  unsigned int M;
  unsigned int n;
  std::vector<ComplexArray> arrays;

 public:
  ComplexArraysContainer() : M(0), n(0) {}
  ComplexArraysContainer(unsigned int arraySize, unsigned int numArrays);
  ComplexArraysContainer(unsigned int arraySize, std::complex<double> value, unsigned int numArrays);

  unsigned int getNumberOfArrays() const;
  unsigned int getArraySize() const;
  std::complex<double> getValue(const unsigned int arrayIndex, const unsigned int i) const;

  ComplexArray& operator[](unsigned int i);
  ComplexArray& operator()(unsigned int i);
  const ComplexArray& operator[](unsigned int i) const;

  std::vector<ComplexArray> operator()() const;
  operator std::vector<ComplexArray>() const;
};

typedef ComplexArraysContainer ComplexArraysContainerCPU;

#endif
