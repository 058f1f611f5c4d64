// Stable LSD radix sort of the deterministic-scatter records (key = node id, value =
// barycentric weight, input in pid order) on the node bits only, via rocPRIM's device radix
// sort.  Stability keeps each node's weights in ascending pid.
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "mp_launch.h"

namespace mpd {

size_t radix_sort_temp_bytes(int64_t n, int end_bit) {
    size_t bytes = 0;
    (void)rocprim::radix_sort_pairs(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                              (const double*)nullptr, (double*)nullptr, (size_t)n, 0, end_bit);
    return bytes;
}

hipError_t radix_sort_pairs(void* temp, size_t temp_bytes, const uint32_t* keys_in, uint32_t* keys_out,
                            const double* vals_in, double* vals_out, int64_t n, int end_bit,
                            hipStream_t st) {
    if (n <= 0) return hipSuccess;
    size_t bytes = temp_bytes;
    return rocprim::radix_sort_pairs(temp, bytes, keys_in, keys_out, vals_in, vals_out, (size_t)n, 0,
                                     end_bit, st);
}

}  // namespace mpd
