// MultimodePropagation.cpp -- the abstract solver interface has no state of its own; this
// unit anchors its vtable in the drop-in library.
#include "MultimodePropagation.hpp"
