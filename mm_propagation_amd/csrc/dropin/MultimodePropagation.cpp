// MultimodePropagation.cpp -- the abstract solver interface (include/MultimodePropagation.hpp:35-45
// of MM_Propagation). Its destructor is the class's key function: defined here, out of line, it
// places the vtable and type information in the drop-in library once, instead of in every
// caller's translation unit. The reference's destructor is not virtual, so deleting a solver
// through a base pointer skipped the derived destructor there; here it runs.
#include "MultimodePropagation.hpp"

MultimodePropagation::~MultimodePropagation() {}
