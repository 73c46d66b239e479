// ComplexArray.cpp -- host complex arrays of the drop-in API (semantics of the reference's
// src/ComplexArray.cpp: element-wise std::complex<double> arithmetic, exp via std::exp).
#include "ComplexArray.hpp"

#include <algorithm>
#include <cmath>

ComplexArray::ComplexArray() : n(0), data(nullptr) {}

ComplexArray::ComplexArray(unsigned int size) : n(size), data(new Complex[size]()) {}

ComplexArray::ComplexArray(unsigned int size, Complex a) : n(size), data(new Complex[size]) {
  std::fill(data, data + n, a);
}

ComplexArray::ComplexArray(const ComplexArray& a) : n(a.n), data(new Complex[a.n]) { std::copy(a.data, a.data + n, data); }

ComplexArray::~ComplexArray() { delete[] data; }

unsigned int ComplexArray::getSize() const { return n; }
Complex& ComplexArray::operator[](unsigned int i) const { return data[i]; }
Complex& ComplexArray::operator()(unsigned int i) const { return data[i]; }
Complex* ComplexArray::operator()() const { return data; }
ComplexArray::operator Complex*() const { return data; }

ComplexArray& ComplexArray::operator=(const ComplexArray& a) {
  if (this != &a) std::copy(a.data, a.data + std::min(n, a.n), data);
  return *this;
}
ComplexArray& ComplexArray::operator=(const ComplexArray* a) { return *this = *a; }

ComplexArray& ComplexArray::operator+=(const ComplexArray& a) {
  for (unsigned int i = 0; i < n; i++) data[i] += a.data[i];
  return *this;
}
ComplexArray& ComplexArray::operator+=(const Complex a) {
  for (unsigned int i = 0; i < n; i++) data[i] += a;
  return *this;
}
ComplexArray& ComplexArray::operator-=(const ComplexArray& a) {
  for (unsigned int i = 0; i < n; i++) data[i] -= a.data[i];
  return *this;
}
ComplexArray& ComplexArray::operator*=(const ComplexArray& a) {
  for (unsigned int i = 0; i < n; i++) data[i] *= a.data[i];
  return *this;
}
ComplexArray& ComplexArray::operator*=(const Complex a) {
  for (unsigned int i = 0; i < n; i++) data[i] *= a;
  return *this;
}
ComplexArray& ComplexArray::operator*=(const double a) {
  for (unsigned int i = 0; i < n; i++) data[i] *= a;
  return *this;
}
ComplexArray& ComplexArray::operator/=(const ComplexArray& a) {
  for (unsigned int i = 0; i < n; i++) data[i] /= a.data[i];
  return *this;
}
ComplexArray& ComplexArray::operator/=(const Complex a) {
  for (unsigned int i = 0; i < n; i++) data[i] /= a;
  return *this;
}
ComplexArray& ComplexArray::operator/=(const double a) {
  for (unsigned int i = 0; i < n; i++) data[i] /= a;
  return *this;
}

ComplexArray operator+(ComplexArray const& a, ComplexArray const& b) { ComplexArray c(a); c += b; return c; }
ComplexArray operator-(ComplexArray const& a, ComplexArray const& b) { ComplexArray c(a); c -= b; return c; }
ComplexArray operator*(ComplexArray const& a, ComplexArray const& b) { ComplexArray c(a); c *= b; return c; }
ComplexArray operator*(ComplexArray const& a, Complex const& b) { ComplexArray c(a); c *= b; return c; }
ComplexArray operator*(ComplexArray const& a, double const& b) { ComplexArray c(a); c *= b; return c; }
ComplexArray operator/(ComplexArray const& a, ComplexArray const& b) { ComplexArray c(a); c /= b; return c; }
ComplexArray operator/(ComplexArray const& a, Complex const& b) { ComplexArray c(a); c /= b; return c; }
ComplexArray operator/(ComplexArray const& a, double const& b) { ComplexArray c(a); c /= b; return c; }

double sum(const std::vector<double> A) {
  double r = 0.;
  for (double x : A) r += x;
  return r;
}

double norm(const ComplexArray A) {
  double r = 0.;
  for (unsigned int i = 0; i < A.getSize(); i++) r += std::norm(A[i]);
  return std::sqrt(r);
}

std::vector<double> abs(const ComplexArray A) {
  std::vector<double> r(A.getSize());
  for (unsigned int i = 0; i < A.getSize(); i++) r[i] = std::abs(A[i]);
  return r;
}

ComplexArray conj(const ComplexArray A) {
  ComplexArray r(A.getSize());
  for (unsigned int i = 0; i < A.getSize(); i++) r[i] = std::conj(A[i]);
  return r;
}

ComplexArray exp(const ComplexArray A) {
  ComplexArray r(A.getSize());
  for (unsigned int i = 0; i < A.getSize(); i++) r[i] = std::exp(A[i]);
  return r;
}
