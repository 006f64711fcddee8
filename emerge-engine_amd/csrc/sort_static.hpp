// sort_static.hpp -- scratch of the allocation-free cell-key sort.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace emh {

// Points s.kbuf, s.vbuf and s.hist into the current device's static scratch
// array (EMH_SORT_NOALLOC_MAX_KEYS keys), resolving its address on the first
// call on that device.  Allocates nothing.
hipError_t noalloc_sort_scratch(SortBuf& s);

}  // namespace emh
