// Sparse3DArray.cpp -- coupling coefficients in insertion order (interface of the reference's
// src/Sparse3DArray.cpp).
#include "Sparse3DArray.hpp"

Sparse3DArray::Sparse3DArray() {}
NonZeroValue Sparse3DArray::getNonZeroValue(const unsigned int i) const { return Q[i]; }
unsigned int Sparse3DArray::getNumberOfNonZeroValues() const { return (unsigned int)Q.size(); }
void Sparse3DArray::addNonZeroValue(const NonZeroValue nzv) { Q.push_back(nzv); }
void Sparse3DArray::resetValue(const unsigned int i, const double Q_klm) { Q[i].value = Q_klm; }
