// ComplexArray.hpp -- host array of std::complex<double> with the reference's interface
// (include/ComplexArray.hpp:12-60 of MM_Propagation): element access, in-place and binary
// arithmetic, and the free functions sum, norm, abs, conj and exp.
#ifndef MMPROP_COMPLEX_ARRAY_HPP
#define MMPROP_COMPLEX_ARRAY_HPP

#include <complex>
#include <vector>

typedef std::complex<double> Complex;

class ComplexArray {
 public:
  ComplexArray();
  explicit ComplexArray(unsigned int size);
  ComplexArray(unsigned int size, Complex a);
  ComplexArray(const ComplexArray& a);
  ~ComplexArray();

  unsigned int getSize() const;

  Complex& operator[](unsigned int i) const;
  Complex& operator()(unsigned int i) const;
  Complex* operator()() const;
  operator Complex*() const;

  // Copies min(size, a.size) leading elements (the reference copies this array's size).
  ComplexArray& operator=(const ComplexArray& a);
  ComplexArray& operator=(const ComplexArray* a);
  ComplexArray& operator+=(const ComplexArray& a);
  ComplexArray& operator+=(const Complex a);
  ComplexArray& operator-=(const ComplexArray& a);
  ComplexArray& operator*=(const ComplexArray& a);
  ComplexArray& operator*=(const Complex a);
  ComplexArray& operator*=(const double a);
  ComplexArray& operator/=(const ComplexArray& a);
  ComplexArray& operator/=(const Complex a);
  ComplexArray& operator/=(const double a);

 private:
  unsigned int n;
  Complex* data;
};

ComplexArray operator+(ComplexArray const& a, ComplexArray const& b);
ComplexArray operator-(ComplexArray const& a, ComplexArray const& b);
ComplexArray operator*(ComplexArray const& a, ComplexArray const& b);
ComplexArray operator*(ComplexArray const& a, Complex const& b);
ComplexArray operator*(ComplexArray const& a, double const& b);
ComplexArray operator/(ComplexArray const& a, ComplexArray const& b);
ComplexArray operator/(ComplexArray const& a, Complex const& b);
ComplexArray operator/(ComplexArray const& a, double const& b);

double sum(const std::vector<double> A);
double norm(const ComplexArray A);
std::vector<double> abs(const ComplexArray A);
ComplexArray conj(const ComplexArray A);
ComplexArray exp(const ComplexArray A);

#endif
