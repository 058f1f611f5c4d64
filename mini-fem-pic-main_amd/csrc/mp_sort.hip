// Stable LSD radix sort of the deterministic-scatter records (key = node id, value =
// barycentric weight, input in pid order) on the node bits only, via rocPRIM's onesweep
// device radix sort.  Stability keeps each node's weights in ascending pid.
//
// The digit width is chosen from the key width so that node ids up to 22 bits (4M nodes) sort
// in two passes over the records: 18-bit keys (a 60^3 box) take 2 passes of 9 bits instead
// of the 3 passes of 8 bits rocPRIM's default width needs.  Each pass reads and writes 12 B per
// record, so a pass saved is a third of the sort's HBM traffic.
#include <algorithm>
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "mp_launch.h"

namespace mpd {

namespace {

// Tiles of the histogram and sorting kernels (threads, items per thread).  Measured on 80M
// records with 18-bit keys (deterministic steps of 20M particles): 1024 x 32 histogram and
// 512 x 20 sort tiles (up to 96 KB of LDS per workgroup, within gfx950's 160 KB) make the
// deterministic step 15% faster than 512 x 8 for both.
#ifndef MP_SORT_HIST_THREADS
#define MP_SORT_HIST_THREADS 1024
#endif
#ifndef MP_SORT_HIST_ITEMS
#define MP_SORT_HIST_ITEMS 32
#endif
#ifndef MP_SORT_THREADS
#define MP_SORT_THREADS 512
#endif
#ifndef MP_SORT_ITEMS
#define MP_SORT_ITEMS 20
#endif

// Largest record count rocPRIM sorts by block sort + merge passes instead of onesweep.  Its
// default (1M) puts the coupled loop's 10k-200k records on a dozen merge-pass launches per
// step; onesweep is a histogram and one launch per digit pass at any size.
#ifndef MP_SORT_MERGE_LIMIT
#define MP_SORT_MERGE_LIMIT 0
#endif

// Tiles for record counts up to MP_SORT_SMALL_MAX (the coupled loop's 10k-200k records): the
// large tiles above leave a 200k-record pass on ~20 workgroups, each working through 10k
// records, where 2k-record tiles spread it over ~100.
#ifndef MP_SORT_SMALL_MAX
#define MP_SORT_SMALL_MAX (1 << 22)
#endif
#ifndef MP_SORT_SMALL_THREADS
#define MP_SORT_SMALL_THREADS 256
#endif
#ifndef MP_SORT_SMALL_ITEMS
#define MP_SORT_SMALL_ITEMS 8
#endif

template <unsigned Bits, int HT, int HI, int ST, int SI>
using OnesweepCfg = rocprim::radix_sort_config<
    rocprim::default_config, rocprim::default_config,
    rocprim::radix_sort_onesweep_config<rocprim::kernel_config<HT, HI>, rocprim::kernel_config<ST, SI>, Bits,
                                        rocprim::block_radix_rank_algorithm::match>,
    MP_SORT_MERGE_LIMIT>;

template <unsigned Bits>
using SortCfg = OnesweepCfg<Bits, MP_SORT_HIST_THREADS, MP_SORT_HIST_ITEMS, MP_SORT_THREADS, MP_SORT_ITEMS>;
template <unsigned Bits>
using SortCfgSmall =
    OnesweepCfg<Bits, MP_SORT_SMALL_THREADS, MP_SORT_SMALL_ITEMS, MP_SORT_SMALL_THREADS, MP_SORT_SMALL_ITEMS>;

// Digit width for an end_bit-wide key: the fewest passes of at most 11 bits, then the
// narrowest width (>= 8) that keeps that pass count.
unsigned digit_bits(int end_bit) {
    if (end_bit <= 8) return 8;
    const int passes = (end_bit + 10) / 11;
    const int bits = (end_bit + passes - 1) / passes;
    return bits < 8 ? 8u : (unsigned)bits;
}

template <unsigned Bits>
hipError_t sort_bits(void* temp, size_t& bytes, const uint32_t* ki, uint32_t* ko, const double* vi,
                     double* vo, size_t n, int end_bit, hipStream_t st) {
    if (n <= (size_t)MP_SORT_SMALL_MAX)
        return rocprim::radix_sort_pairs<SortCfgSmall<Bits>>(temp, bytes, ki, ko, vi, vo, n, 0, (unsigned)end_bit,
                                                             st);
    return rocprim::radix_sort_pairs<SortCfg<Bits>>(temp, bytes, ki, ko, vi, vo, n, 0, (unsigned)end_bit, st);
}

hipError_t sort_any(void* temp, size_t& bytes, const uint32_t* ki, uint32_t* ko, const double* vi,
                    double* vo, size_t n, int end_bit, hipStream_t st) {
    switch (digit_bits(end_bit)) {
        case 9: return sort_bits<9>(temp, bytes, ki, ko, vi, vo, n, end_bit, st);
        case 10: return sort_bits<10>(temp, bytes, ki, ko, vi, vo, n, end_bit, st);
        case 11: return sort_bits<11>(temp, bytes, ki, ko, vi, vo, n, end_bit, st);
        default: return sort_bits<8>(temp, bytes, ki, ko, vi, vo, n, end_bit, st);
    }
}

}  // namespace

// Scratch for any count up to n: the small-tile configuration (more tiles, so more look-back
// state per record) serves counts up to MP_SORT_SMALL_MAX, the large one the rest.
size_t radix_sort_temp_bytes(int64_t n, int end_bit) {
    size_t bytes = 0, small = 0;
    (void)sort_any(nullptr, bytes, nullptr, nullptr, nullptr, nullptr, (size_t)n, end_bit, nullptr);
    if (n > (int64_t)MP_SORT_SMALL_MAX)
        (void)sort_any(nullptr, small, nullptr, nullptr, nullptr, nullptr, (size_t)MP_SORT_SMALL_MAX, end_bit,
                       nullptr);
    return std::max(bytes, small);
}

hipError_t radix_sort_pairs(void* temp, size_t temp_bytes, const uint32_t* keys_in, uint32_t* keys_out,
                            const double* vals_in, double* vals_out, int64_t n, int end_bit,
                            hipStream_t st) {
    if (n <= 0) return hipSuccess;
    size_t bytes = temp_bytes;
    return sort_any(temp, bytes, keys_in, keys_out, vals_in, vals_out, (size_t)n, end_bit, st);
}

}  // namespace mpd
