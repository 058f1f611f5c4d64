// Deterministic scatter by node segments.  The serial MoveParticles adds each node's weights
// in ascending pid (src/mini-pic.cpp:297-304 run one particle after another), so the fold of
// node n is ((0 + w_p1) + w_p2) + ... over the pids p1 < p2 < ... that end the step in a tet
// of n.  Here the push writes every slot's four records (node, weight) at 4*slot+v, in
// whatever order the slots are (cell-sorted for the gathers' locality), and counts the records
// per node through its atomic scatter with unit weights.  Then:
//   k_seg_counts   the counts as ints (the raw density buffer cleared for the next step);
//   exclusive scan the segment of each node in the grouped record array;
//   k_seg_bucket   each 2048-record tile groups its records by node in an LDS hash and claims
//                  one range per node with one atomic, writing (pid, weight) into it;
//   k_seg_sort     one wave per node: the segment's (pid, index) pairs sorted by pid in
//                  registers (bitonic network, up to 512 records) and the weights written back
//                  in that order;
//   k_seg_hot_sort segments over 512 records (dense clusters): a workgroup sorts each by pid
//                  (stable LSD radix on 8-bit digits through the record buffers);
//   k_fold         the serial fold of each segment, one lane per node (k_fold_long: one wave
//                  per node for the longest).
// No pass waits on another workgroup's progress, so small record counts are not serialised
// along a look-back chain as in a device radix sort.
#include <algorithm>

#include "mp_common.h"
#include "mp_launch.h"

namespace mpd {

namespace {

constexpr uint32_t kNoNode = 0xFFFFFFFFu;
constexpr int kBucketItems = 8;
constexpr int kBucketTile = 256 * kBucketItems;
constexpr int kBucketBits = 12;  // LDS hash slots: twice the tile, so every key finds a slot
constexpr int kSegWaveMax = 8 * kWave;

// cnt[n] = records of node n (the push's unit-weight sums, exact integers in FP64), raw[n]
// cleared for the next step; cursors and the two list counts cleared; cnt[nn] = 0 makes the
// scan's last entry the total.
__global__ void k_seg_counts(double* __restrict__ raw, int32_t* __restrict__ cnt, int32_t* __restrict__ cur,
                             int32_t* __restrict__ hot_count, int32_t* __restrict__ long_count, int32_t nn) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nn) {
        cnt[i] = (int32_t)raw[i];
        raw[i] = 0.0;
        cur[i] = 0;
    } else if (i == nn) {
        cnt[nn] = 0;
        *hot_count = 0;
        *long_count = 0;
    }
}

// Records [t0, t0 + 2048) of the slot-ordered array into their node segments.  A tile of
// sorted slots covers a few hundred nodes, so the global cursors see one returning atomic per
// node and tile; within a node's range the order is the LDS atomics' (k_seg_sort sorts it).
__global__ void __launch_bounds__(256)
k_seg_bucket(const uint32_t* __restrict__ keys, const double* __restrict__ vals, const int32_t* __restrict__ pid,
             int64_t m, const int32_t* __restrict__ start, int32_t* __restrict__ cur,
             uint32_t* __restrict__ out_pid, double* __restrict__ out_val) {
    constexpr int kSlots = 1 << kBucketBits;
    __shared__ uint32_t hk[kSlots];
    __shared__ int32_t hc[kSlots];
    for (int q = threadIdx.x; q < kSlots; q += blockDim.x) {
        hk[q] = kNoNode;
        hc[q] = 0;
    }
    const int64_t t0 = (int64_t)blockIdx.x * kBucketTile;
    uint32_t k[kBucketItems];
    double v[kBucketItems];
    int32_t p[kBucketItems];
#pragma unroll
    for (int it = 0; it < kBucketItems; it++) {
        const int64_t r = t0 + it * 256 + threadIdx.x;
        const bool in = r < m;
        k[it] = in ? keys[r] : kNoNode;
        v[it] = in ? vals[r] : 0.0;
        p[it] = in ? pid[r >> 2] : 0;
    }
    __syncthreads();
    int32_t slot[kBucketItems], rank[kBucketItems];
#pragma unroll
    for (int it = 0; it < kBucketItems; it++) {
        slot[it] = -1;
        rank[it] = 0;
        if (k[it] == kNoNode) continue;
        uint32_t h = (k[it] * 0x9E3779B1u) >> (32 - kBucketBits);
        // at most 2048 distinct keys in 4096 slots: the probe ends at the key or an empty slot
        for (int q = 0; q < kSlots; q++) {
            const uint32_t prev = atomicCAS(&hk[h], kNoNode, k[it]);
            if (prev == kNoNode || prev == k[it]) break;
            h = (h + 1) & (kSlots - 1);
        }
        slot[it] = (int32_t)h;
        rank[it] = atomicAdd(&hc[h], 1);
    }
    __syncthreads();
    for (int q = threadIdx.x; q < kSlots; q += blockDim.x) {
        const uint32_t node = hk[q];
        if (node != kNoNode) hc[q] = start[node] + atomicAdd(&cur[node], hc[q]);
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kBucketItems; it++) {
        if (slot[it] < 0) continue;
        const int64_t pos = (int64_t)hc[slot[it]] + rank[it];
        out_pid[pos] = (uint32_t)p[it];
        out_val[pos] = v[it];
    }
}

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

#ifndef MP_SEG_DPP
#define MP_SEG_DPP 1
#endif

// The value of lane ^ j.  With MP_SEG_DPP, lane distances 1, 2, 4 and 8 move through DPP
// (VALU, no LDS traffic: quad permutes, and row rotates by 8, and by 4 or 12 as `up4` picks
// the one that reads lane ^ 4), 16 through ds_swizzle; 32 (and all of them without
// MP_SEG_DPP) through ds_bpermute.
__device__ __forceinline__ uint32_t lane_xor(uint32_t v, int j, bool up4) {
#if MP_SEG_DPP
    if (j == 1) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, true);
    if (j == 2) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, true);
    if (j == 4) {
        const int a = __builtin_amdgcn_mov_dpp((int)v, 0x124, 0xF, 0xF, true);
        const int b = __builtin_amdgcn_mov_dpp((int)v, 0x12C, 0xF, 0xF, true);
        return (uint32_t)(up4 ? a : b);
    }
    if (j == 8) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x128, 0xF, 0xF, true);
    if (j == 16) return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x401F);
#endif
    (void)up4;
    return (uint32_t)__shfl_xor((int)v, j, kWave);
}

// Bitonic sort of 64*E (key, idx) pairs held by a wave, element g in lane g % 64, register
// g / 64; ascending in g.  Keys are unique except the padding (all 0xFFFFFFFF), whose order
// does not matter.  A compare-exchange across lanes is one unsigned compare against the
// partner's key, flipped on the lanes that keep the larger key of their pair (the upper lane of
// an ascending pair, the lower of a descending one: a function of the lane bits j and k, the
// same for every register while k < 64), then a select of key and index.
template <int E>
__device__ __forceinline__ void wave_bitonic(uint32_t (&key)[E], uint32_t (&idx)[E], int lane) {
    constexpr int N = kWave * E;
#if MP_SEG_DPP
    // which row rotate brings lane ^ 4 here: rotate by 4 for one half of each row, by 12 for
    // the other
    const bool up4 = __builtin_amdgcn_mov_dpp(lane, 0x124, 0xF, 0xF, true) == (lane ^ 4);
#else
    const bool up4 = false;
#endif
#pragma unroll
    for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j >= kWave) {
                // both elements in this lane: registers e and e | j/64; the direction depends on
                // the register only (k >= 128 > lane)
                const int je = j / kWave;
#pragma unroll
                for (int e = 0; e < E; e++) {
                    if (e & je) continue;
                    const int f = e | je;
                    const bool asc = ((e * kWave) & k) == 0;
                    const bool sw = asc ? key[f] < key[e] : key[e] < key[f];
                    const uint32_t ka = key[e], kb = key[f], ia = idx[e], ib = idx[f];
                    key[e] = sw ? kb : ka;
                    key[f] = sw ? ka : kb;
                    idx[e] = sw ? ib : ia;
                    idx[f] = sw ? ia : ib;
                }
            } else {
                const bool upper = (lane & j) != 0;
#pragma unroll
                for (int e = 0; e < E; e++) {
                    const bool desc = ((e * kWave + lane) & k) != 0;
                    const bool keep_max = upper != desc;
                    const uint32_t pk = lane_xor(key[e], j, up4);
                    const uint32_t pi = lane_xor(idx[e], j, up4);
                    // unique keys: the two lanes of a pair agree on whether to swap
                    const bool take = (pk < key[e]) != keep_max;
                    key[e] = take ? pk : key[e];
                    idx[e] = take ? pi : idx[e];
                }
            }
        }
    }
}

// Sorts the segment [s, s + m) of (pid, weight) records by pid and writes the weights back in
// that order (the pids stay: only the order of the weights is read afterwards).
template <int E>
__device__ __forceinline__ void seg_sort_wave(const uint32_t* __restrict__ spid, double* __restrict__ sval,
                                              int64_t s, int m, int lane, double* lv) {
    uint32_t key[E], idx[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int g = e * kWave + lane;
        idx[e] = (uint32_t)g;
        if (g < m) {
            key[e] = spid[s + g];
            lv[g] = sval[s + g];
        } else {
            key[e] = kNoNode;
        }
    }
    wave_bitonic<E>(key, idx, lane);
    wave_sync_lds();
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int g = e * kWave + lane;
        if (g < m) sval[s + g] = lv[idx[e]];
    }
}

// One wave per node (a scalar node index, so the branches on the segment length are
// wave-uniform).  Segments of more than 512 records are listed for the hot path.
__global__ void __launch_bounds__(256)
k_seg_sort(const uint32_t* __restrict__ spid, double* __restrict__ sval, const int32_t* __restrict__ start,
           int32_t nn, int32_t* __restrict__ hot_list, int32_t* __restrict__ hot_count) {
    __shared__ double lval[256 / kWave][kSegWaveMax];
    const int lane = lane_id();
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
    const int64_t n = (int64_t)blockIdx.x * (256 / kWave) + w;
    if (n >= nn) return;
    const int64_t s = start[n];
    const int m = start[n + 1] - (int32_t)s;
    if (m > kSegWaveMax) {
        if (lane == 0) hot_list[atomicAdd(hot_count, 1)] = (int32_t)n;
    } else if (m > 4 * kWave) {
        seg_sort_wave<8>(spid, sval, s, m, lane, lval[w]);
    } else if (m > 2 * kWave) {
        seg_sort_wave<4>(spid, sval, s, m, lane, lval[w]);
    } else if (m > kWave) {
        seg_sort_wave<2>(spid, sval, s, m, lane, lval[w]);
    } else if (m > 1) {
        seg_sort_wave<1>(spid, sval, s, m, lane, lval[w]);
    }
}

// Each workgroup takes listed hot segments in turn and sorts each by pid: `passes` (even) LSD
// passes of 8 bits, 256 records per round, stable within a round by wave ballots (rank among
// the lanes with the same digit) and per-wave digit counts in LDS.  Even passes end in
// (pa, va); (pb, vb) is scratch at the same offsets.
__global__ void __launch_bounds__(256)
k_seg_hot_sort(uint32_t* __restrict__ pa, double* __restrict__ va, uint32_t* __restrict__ pb,
               double* __restrict__ vb, const int32_t* __restrict__ start, const int32_t* __restrict__ hot_list,
               const int32_t* __restrict__ hot_count, int passes) {
    __shared__ int32_t base[256];
    __shared__ int32_t wcnt[256 / kWave][256];
    const int tid = threadIdx.x, lane = lane_id();
    const int w = tid / kWave;
    const int32_t count = *hot_count;
    for (int32_t h = blockIdx.x; h < count; h += gridDim.x) {
        const int32_t n = hot_list[h];
        const int64_t s = start[n];
        const int64_t m = start[n + 1] - s;
        uint32_t *sp = pa + s, *dp = pb + s;
        double *sv = va + s, *dv = vb + s;
        for (int pass = 0; pass < passes; pass++) {
            const int sh = 8 * pass;
            base[tid] = 0;
#pragma unroll
            for (int q = 0; q < 256 / kWave; q++) wcnt[q][tid] = 0;
            __syncthreads();
            for (int64_t i = tid; i < m; i += 256) atomicAdd(&base[(sp[i] >> sh) & 255], 1);
            __syncthreads();
            if (w == 0) {  // exclusive scan of the 256 digit counts, 4 per lane
                const int c0 = base[4 * lane], c1 = base[4 * lane + 1], c2 = base[4 * lane + 2],
                          c3 = base[4 * lane + 3];
                const int t = c0 + c1 + c2 + c3;
                int inc = t;
                for (int o = 1; o < kWave; o <<= 1) {
                    const int y = __shfl_up(inc, o, kWave);
                    if (lane >= o) inc += y;
                }
                const int ex = inc - t;
                base[4 * lane] = ex;
                base[4 * lane + 1] = ex + c0;
                base[4 * lane + 2] = ex + c0 + c1;
                base[4 * lane + 3] = ex + c0 + c1 + c2;
            }
            __syncthreads();
            for (int64_t r0 = 0; r0 < m; r0 += 256) {
                const int64_t i = r0 + tid;
                const bool valid = i < m;
                const uint32_t key = valid ? sp[i] : 0u;
                const double val = valid ? sv[i] : 0.0;
                const uint32_t d = (key >> sh) & 255;
                unsigned long long peers = __ballot(valid);
#pragma unroll
                for (int b = 0; b < 8; b++) {
                    const bool bit = (d >> b) & 1;
                    const unsigned long long bb = __ballot(bit);
                    peers &= bit ? bb : ~bb;
                }
                const int rank = __popcll(peers & ((1ull << lane) - 1));
                if (valid && rank == 0) wcnt[w][d] = __popcll(peers);
                __syncthreads();
                if (valid) {
                    int before = 0;
                    for (int q = 0; q < w; q++) before += wcnt[q][d];
                    const int64_t pos = base[d] + before + rank;
                    dp[pos] = key;
                    dv[pos] = val;
                }
                __syncthreads();
                int add = 0;
#pragma unroll
                for (int q = 0; q < 256 / kWave; q++) {
                    add += wcnt[q][tid];
                    wcnt[q][tid] = 0;
                }
                base[tid] += add;
                __syncthreads();
            }
            uint32_t* tp = sp;
            sp = dp;
            dp = tp;
            double* tv = sv;
            sv = dv;
            dv = tv;
        }
    }
}

unsigned grid_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

hipError_t launch_seg_fold(const SegFoldArgs& a, hipStream_t st) {
    const int32_t nn = a.nn;
    int32_t* cnt = a.tab;
    int32_t* start = cnt + (int64_t)nn + 1;
    int32_t* cur = start + (int64_t)nn + 1;
    int32_t* hot_list = cur + nn;
    int32_t* hot_count = hot_list + nn;
    int32_t* long_list = hot_count + 1;
    int32_t* long_count = long_list + nn;
    hipLaunchKernelGGL(k_seg_counts, dim3(grid_for((int64_t)nn + 1, 256)), dim3(256), 0, st, a.raw, cnt, cur,
                       hot_count, long_count, nn);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = launch_excl_scan(cnt, (int64_t)nn + 1, start, a.scan_tmp, st);
    if (e != hipSuccess) return e;
    if (a.m > 0)
        hipLaunchKernelGGL(k_seg_bucket, dim3(grid_for(a.m, kBucketTile)), dim3(256), 0, st, a.keys, a.vals, a.pid,
                           a.m, start, cur, a.seg_pid, a.seg_val);
    hipLaunchKernelGGL(k_seg_sort, dim3(grid_for(nn, 256 / kWave)), dim3(256), 0, st, a.seg_pid, a.seg_val, start,
                       nn, hot_list, hot_count);
    // hot segments: as many workgroups as a few per CU, each taking listed nodes in turn (none
    // listed: they return at once)
    const int passes = std::max(2, ((a.pid_bits + 7) / 8 + 1) / 2 * 2);
    const unsigned gh = (unsigned)std::min<int64_t>(1024, std::max<int64_t>(1, nn));
    hipLaunchKernelGGL(k_seg_hot_sort, dim3(gh), dim3(256), 0, st, a.seg_pid, a.seg_val, a.keys, a.vals, start,
                       hot_list, hot_count, passes);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    // every segment now holds its weights in ascending pid: the serial folds
    return launch_fold_nodes(a.seg_val, start, start + 1, a.nvol, nn, a.spwt, nullptr, nn, a.out, long_list,
                             long_count, st);
}

}  // namespace mpd
