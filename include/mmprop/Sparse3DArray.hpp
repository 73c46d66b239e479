// Sparse3DArray.hpp -- COO list of the coupling coefficients Q^(p)_{k,l,m} kept in insertion
// order (include/Sparse3DArray.hpp:7-30 of MM_Propagation). The order fixes the pair slots.
#ifndef MMPROP_SPARSE_3D_ARRAY_HPP
#define MMPROP_SPARSE_3D_ARRAY_HPP

#include <vector>

struct NonZeroValue {
  unsigned int index[3];
  double value;
};

class Sparse3DArray {
 public:
  Sparse3DArray();

  NonZeroValue getNonZeroValue(const unsigned int i) const;
  unsigned int getNumberOfNonZeroValues() const;

  void addNonZeroValue(const NonZeroValue nzv);
  void resetValue(const unsigned int i, const double Q_klm);

 private:
  std::vector<NonZeroValue> Q;
};

#endif
