// sort_static.hip -- the scratch memory of the allocation-free cell-key sort
// (emh_sort_cell_keys_noalloc): one file-scope device array per GPU, sized
// for EMH_SORT_NOALLOC_MAX_KEYS keys.
//
// Host code never passes the array's name to a kernel or a copy: in host
// code that name is a stand-in, not a device address.  The first use on a
// device resolves the array's address there with hipGetSymbolAddress (that
// device current) and keeps it; every later use on that device takes the
// kept address.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/emerge_hip.h"
#include "kernels.hpp"
#include "sort_static.hpp"

namespace emh {

namespace {

constexpr uint32_t kMaxKeys = EMH_SORT_NOALLOC_MAX_KEYS;
// the radix sort's scratch: two key and two slot buffers of kMaxKeys, then
// the counts of one pass (at most 2^kSortMaxDigit buckets per workgroup) and
// its digit totals
constexpr size_t kHistWords = sort_hist_len_max(kMaxKeys);
constexpr size_t kScratchWords = 4 * (size_t)kMaxKeys + kHistWords;
static_assert(kScratchWords * 4 == EMH_SORT_NOALLOC_SCRATCH_BYTES, "header and source disagree on the scratch size");

__device__ __align__(256) uint32_t scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint32_t* g_ptr[kMaxDevices] = {};  // the array's address on each device, once resolved

}  // namespace

hipError_t noalloc_sort_scratch(SortBuf& s) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    uint32_t* p = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        if (!g_ptr[dev]) {
            void* addr = nullptr;
            e = hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch));
            if (e != hipSuccess) return e;
            if (!addr) return hipErrorInvalidSymbol;
            g_ptr[dev] = static_cast<uint32_t*>(addr);
        }
        p = g_ptr[dev];
    }
    s = SortBuf{};
    for (int k = 0; k < 2; ++k) {
        s.kbuf[k] = p + (size_t)(2 * k) * kMaxKeys;
        s.vbuf[k] = p + (size_t)(2 * k + 1) * kMaxKeys;
    }
    s.hist = p + 4 * (size_t)kMaxKeys;
    return hipSuccess;
}

}  // namespace emh
