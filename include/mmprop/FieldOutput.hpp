// FieldOutput.hpp -- the text output of MM_Propagation's example driver
// (simulations/main.cpp:101-103): one line per sample j,
//   |A_0(j)| ; |A_1(j)| ; ... (initial fields) ; |B_0(j)| ; |B_1(j)| ; ... (propagated fields)
// written with the stream's current formatting (default: 6 significant digits).
#ifndef MMPROP_FIELD_OUTPUT_HPP
#define MMPROP_FIELD_OUTPUT_HPP

#include <ostream>

#include "ComplexArraysContainer.hpp"

void writeFieldMagnitudes(std::ostream& os, const ComplexArraysContainer& initial,
                          const ComplexArraysContainer& result);

#endif
