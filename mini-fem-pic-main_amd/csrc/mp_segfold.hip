// Deterministic scatter by node segments.  The serial MoveParticles adds each node's weights
// in ascending pid (src/mini-pic.cpp:297-304 run one particle after another), so the fold of
// node n is ((0 + w_p1) + w_p2) + ... over the pids p1 < p2 < ... that end the step in a tet
// of n.  Here the push counts each node's records through its atomic scatter with unit weights
// and writes no records: the tile passes rebuild every particle's four records (node, weight)
// from the pushed state, in slot order (cell-sorted for the gathers' locality).  Each node's
// records are cut into slabs by pid range so that every slab fits one wave:
//   k_seg_counts   the counts as ints (the raw density buffer cleared for the next step); a
//                  node over 512 records ("hot") gets 2^b slabs, b the smallest with at most
//                  128 records a slab on average; the nodes listed hot first, cold after;
//   exclusive scan the first slab of each node (a cold node has one);
//   k_slab_prep    per node (first slab, b); a cold node's slab count is its record count;
//   k_seg_tile<1>  hot nodes only (it returns at once without): the records per slab, counted
//                  in an LDS hash per 512-particle tile and added with one atomic per slab;
//   exclusive scan the start of every slab in the grouped record array;
//   k_seg_tile<0>  each 512-particle tile rebuilds its records, groups them by slab in an LDS
//                  hash and claims one range per slab with one atomic, writing (pid, weight);
//   k_seg_sortfold one wave per node, hot nodes first: slab after slab, each record ranked
//                  by pid within a sub-bucket of the slab's pid range (a register bitonic
//                  network where the pids bunch), the weights put in that order in LDS and
//                  added by the first lane, so the fold runs through the node's records in
//                  ascending pid; the normalised density is written at the end.
// A slab's pids are those whose 16-bit digit d(pid) = (pid * floor(2^48 / P)) >> 32 (P = the
// next pid to issue) has its top b bits equal to the slab's index: d is non-decreasing in the
// pid, so slab order is pid order.  With the pids of a hot node spread over [0, P) a slab holds
// 64..128 records on average; where they are bunched and a slab exceeds 256, the node goes to
// a slower path (a workgroup LSD radix sort of its whole range, then a wave's serial fold).
// No pass waits on another workgroup's progress.
// det_algo 2 keeps the slab cut of an exact step and has the push place the records of the next
// steps (mp_common.h, place_records): k_place_lim / k_place_cap lay out each slab's segment
// from the exact step's slab counts, k_place_check turns the placed counts into the next
// layout, k_seg_sortfold<true, ...> folds the placed slabs (k_seg_slab_cont finishing a node
// whose slab outgrew 256 records, k_place_unpack feeding a cold node over 512 to the workgroup
// sort), and k_place_to_raw hands the counts to the exact path when a step is rebuilt.
#include <algorithm>

#include "mp_common.h"
#include "mp_launch.h"

namespace mpd {

namespace {

constexpr uint32_t kNoKey = 0xFFFFFFFFu;
constexpr int kBucketBits = 12;  // LDS hash slots: twice a tile's records, so every key finds a slot
constexpr int kSegWaveMax = 8 * kWave;
constexpr int kSlabShift = 7;  // hot slabs: (64, 128] records on average
constexpr int kBigBits = 8;    // nodes of 2^8 slabs and more (over 16k records) are listed first

// The table of launch_seg_fold, carved from SegFoldArgs::tab (ints).
struct SegTab {
    int32_t* cnt;         // nn: records per node
    int32_t* bits;        // nn: slab bits b (0: cold, one slab)
    int32_t* nsl;         // nn + 1: 2^b, then 0
    int32_t* sbase;       // nn + 1: first slab of each node (exclusive scan of nsl)
    int32_t* order;       // nn: hot nodes under 2^kBigBits slabs from the front, cold nodes from the back
    int32_t* big;         // nn: hot nodes of 2^kBigBits slabs and more
    int32_t* ovf_list;    // nn: hot nodes with a slab over 256 records
    int32_t* nstart;      // nn: record range of a listed node
    int32_t* nend;        // nn
    int2* sinfo;          // nn: (first slab, b)
    int32_t* ctrs;        // 8: hot nodes (all), cold, overflowing, big hot, other hot, unused
    int32_t* slab_cnt;    // smax + 1: records per slab
    int32_t* cur;         // smax: claim cursors
    int32_t* slab_start;  // smax + 1: exclusive scan of slab_cnt
};

// Slabs: one per cold node, and at most c / 64 for a hot node of c records, so at most
// nn + m / 64 in all.
int64_t slab_max(int32_t nn, int64_t m) { return (int64_t)nn + m / 64 + 1; }

SegTab carve(int32_t* t, int32_t nn, int64_t m) {
    SegTab s;
    const int64_t n = nn, smax = slab_max(nn, m);
    s.cnt = t;
    s.bits = s.cnt + n;
    s.nsl = s.bits + n;
    s.sbase = s.nsl + n + 1;
    s.order = s.sbase + n + 1;
    s.big = s.order + n;
    s.ovf_list = s.big + n;
    s.nstart = s.ovf_list + n;
    s.nend = s.nstart + n;
    int32_t* p = s.nend + n;
    p += (p - t) & 1;  // int2 alignment
    s.sinfo = reinterpret_cast<int2*>(p);
    s.ctrs = p + 2 * n;
    s.slab_cnt = s.ctrs + 8;
    s.cur = s.slab_cnt + smax + 1;
    s.slab_start = s.cur + smax;
    return s;
}

__device__ __forceinline__ int bit_width_u32(uint32_t v) { return v ? 32 - __clz((int)v) : 0; }

__device__ __forceinline__ unsigned long long lanes_below() {
    return (1ull << lane_id()) - 1;
}

// cnt = records of node n (the push's unit-weight sums, exact integers in FP64), raw cleared
// for the next step; the node's slab bits; the node listed: big hot nodes in `big`, other hot
// nodes at the front of `order`, cold nodes at its back (wave-aggregated list atomics).
__global__ void __launch_bounds__(256)
k_seg_counts(double* __restrict__ raw, int32_t nn, int32_t* __restrict__ cnt, int32_t* __restrict__ bits,
             int32_t* __restrict__ nsl, int32_t* __restrict__ order, int32_t* __restrict__ big,
             int32_t* __restrict__ ctrs) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < nn;
    int b = 0;
    if (valid) {
        const int32_t c = (int32_t)raw[i];
        raw[i] = 0.0;
        cnt[i] = c;
        if (c > kSegWaveMax) b = std::min(kSlabFineBits, bit_width_u32((uint32_t)(c - 1) >> kSlabShift));
        bits[i] = b;
        nsl[i] = 1 << b;
    } else if (i == nn) {
        nsl[nn] = 0;
    }
    const bool isbig = valid && b >= kBigBits, hot = valid && b > 0 && b < kBigBits, cold = valid && b == 0;
    const unsigned long long gm = __ballot(isbig), hm = __ballot(hot), cm = __ballot(cold);
    int32_t gb = 0, hb = 0, cb = 0;
    if (lane_id() == 0) {
        if (gm | hm) atomicAdd(ctrs + 0, __popcll(gm | hm));
        if (gm) gb = atomicAdd(ctrs + 3, __popcll(gm));
        if (hm) hb = atomicAdd(ctrs + 4, __popcll(hm));
        if (cm) cb = atomicAdd(ctrs + 1, __popcll(cm));
    }
    gb = __shfl(gb, 0, kWave);
    hb = __shfl(hb, 0, kWave);
    cb = __shfl(cb, 0, kWave);
    if (isbig) big[gb + __popcll(gm & lanes_below())] = (int32_t)i;
    if (hot) order[hb + __popcll(hm & lanes_below())] = (int32_t)i;
    if (cold) order[nn - 1 - (cb + __popcll(cm & lanes_below()))] = (int32_t)i;
}

__global__ void __launch_bounds__(256)
k_slab_prep(const int32_t* __restrict__ cnt, const int32_t* __restrict__ bits, const int32_t* __restrict__ sbase,
            int32_t nn, int2* __restrict__ sinfo, int32_t* __restrict__ slab_cnt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    const int32_t b = bits[i], s = sbase[i];
    sinfo[i] = make_int2(s, b);
    if (b == 0) slab_cnt[s] = cnt[i];
}

constexpr int kTileParts = 512;  // particles of a tile (2048 records)
constexpr int kTileItems = kTileParts / 256;

// The LDS hash slot of key (at most 2048 distinct keys in 4096 slots: the probe ends at the key
// or an empty slot).
__device__ __forceinline__ int tile_slot(uint32_t* hk, uint32_t key) {
    constexpr int kSlots = 1 << kBucketBits;
    uint32_t h = (key * 0x9E3779B1u) >> (32 - kBucketBits);
#if MP_PROBE_ROLLED
#pragma unroll 1
#endif
    for (int q = 0; q < kSlots; q++) {
        const uint32_t prev = atomicCAS(&hk[h], kNoKey, key);
        if (prev == kNoKey || prev == key) break;
        h = (h + 1) & (kSlots - 1);
    }
    return (int)h;
}

// Slot and rank of this lane's key (kNoKey: none) among the tile's keys.  Called by every lane
// of the wave: the lanes holding the wave's first valid key (on cell-sorted particles, most
// lanes: neighbouring particles share the tet and so the node of vertex v) share the leader's
// slot lookup and one LDS atomic, ranked by lane; the others take their own.
__device__ __forceinline__ void tile_take(uint32_t* hk, int32_t* hc, uint32_t key, int& slot, int& rank) {
    slot = -1;
    rank = 0;
    const unsigned long long valid = __ballot(key != kNoKey);
    if (valid == 0) return;
    const int lane = lane_id();
    const int lead = __ffsll((long long)valid) - 1;
    const uint32_t k0 = (uint32_t)__builtin_amdgcn_readlane((int)key, lead);
    const unsigned long long same = __ballot(key == k0);
    int s0 = 0, r0 = 0;
    if (lane == lead) {
        s0 = tile_slot(hk, k0);
        r0 = atomicAdd(&hc[s0], __popcll(same));
    }
    s0 = __builtin_amdgcn_readlane(s0, lead);
    r0 = __builtin_amdgcn_readlane(r0, lead);
    if (key == k0) {
        slot = s0;
        rank = r0 + __popcll(same & lanes_below());
    } else if (key != kNoKey) {
        slot = tile_slot(hk, key);
        rank = atomicAdd(&hc[slot], 1);
    }
}

// The records of particles [t0, t0 + 512) of the pushed state, by slab.  A particle still in the
// domain (cell >= 0) ends its step in tet c at (x, y, z); its records are (con_v(c),
// bary_v(c; x, y, z)), v = 0..3, the weights the walk computed there (same arithmetic, same
// operands).  COUNT: hot nodes' records only, their counts added to slab_cnt (nothing to do
// without a hot node).  Else: every record, written as (pid, weight) into its slab's range; a
// tile of cell-sorted particles covers a few dozen tets, so the global cursors see one
// returning atomic per slab and tile; within a slab the order is the LDS atomics' (the sort
// puts it right).  Tiles are mapped to XCDs in contiguous eighths (xcd_block): the tiles that
// fill one slab's range, neighbours in cell-sorted order, run on one XCD and its L2 merges their
// partial lines.
template <bool COUNT>
__global__ void __launch_bounds__(256)
k_seg_tile(const TetBlock* __restrict__ tets, Soa st, int64_t n, const int2* __restrict__ sinfo,
           unsigned long long mul, const int32_t* __restrict__ ctrs, int32_t* __restrict__ slab_cnt,
           const int32_t* __restrict__ slab_start, int32_t* __restrict__ cur, uint32_t* __restrict__ out_pid,
           double* __restrict__ out_val) {
    const bool hot = ctrs[0] != 0;
    if (COUNT && !hot) return;
    constexpr int kSlots = 1 << kBucketBits;
    __shared__ uint32_t hk[kSlots];
    __shared__ int32_t hc[kSlots];
    const int64_t i0 = xcd_block() * kTileParts + threadIdx.x;
    int32_t c[kTileItems];
    uint32_t p[kTileItems];
    double x[kTileItems], y[kTileItems], z[kTileItems];
#pragma unroll
    for (int j = 0; j < kTileItems; j++) {
        const int64_t i = i0 + j * 256;
        c[j] = i < n ? st.cell[i] : -1;
        p[j] = i < n ? (uint32_t)st.pid[i] : 0u;
        x[j] = y[j] = z[j] = 0.0;
        if (!COUNT && i < n) {
            x[j] = ld(st.x + i);
            y[j] = ld(st.y + i);
            z[j] = ld(st.z + i);
        }
    }
    for (int q = threadIdx.x; q < kSlots; q += blockDim.x) {
        hk[q] = kNoKey;
        hc[q] = 0;
    }
    uint32_t k[4 * kTileItems];
    double w[4 * kTileItems];
#pragma unroll
    for (int j = 0; j < kTileItems; j++) {
        const bool live = c[j] >= 0;
        const TetView t = tet_at(tets, live ? c[j] : 0);
#pragma unroll
        for (int v = 0; v < 4; v++) {
            k[4 * j + v] = live ? (uint32_t)t.con(v) : kNoKey;
            w[4 * j + v] = 0.0;
        }
        if (!COUNT && live) {
            const double vol = t.vol();
#pragma unroll
            for (int v = 0; v < 4; v++)
                w[4 * j + v] = bary(t.c(v, 0), t.c(v, 1), t.c(v, 2), t.c(v, 3), x[j], y[j], z[j], vol);
        }
    }
    // without a hot node every node has one slab and slab n is node n's (the scan of ones)
    if (hot) {
#pragma unroll
        for (int r = 0; r < 4 * kTileItems; r++) {
            if (k[r] == kNoKey) continue;
            const int2 si = sinfo[k[r]];
            k[r] = (COUNT && si.y == 0) ? kNoKey : slab_of(si, p[r / 4], mul);
        }
    }
    __syncthreads();
    int slot[4 * kTileItems], rank[4 * kTileItems];
#pragma unroll
    for (int r = 0; r < 4 * kTileItems; r++) tile_take(hk, hc, k[r], slot[r], rank[r]);
    __syncthreads();
    // one atomic per occupied slot; a thread's returning claims go four at a time, all four
    // in flight before any is waited for
    constexpr int kClaims = 4;
    static_assert(kSlots % (kClaims * 256) == 0, "whole claim groups per thread");
    for (int q0 = threadIdx.x; q0 < kSlots; q0 += kClaims * 256) {
        uint32_t key[kClaims];
        int32_t got[kClaims];
#pragma unroll
        for (int u = 0; u < kClaims; u++) {
            const int q = q0 + u * 256;
            key[u] = hk[q];
            got[u] = key[u] != kNoKey ? hc[q] : 0;
        }
#pragma unroll
        for (int u = 0; u < kClaims; u++) {
            if (key[u] == kNoKey) continue;
            if (COUNT)
                atomicAdd(slab_cnt + key[u], got[u]);
            else
                got[u] = slab_start[key[u]] + atomicAdd(&cur[key[u]], got[u]);
        }
        if (!COUNT) {
#pragma unroll
            for (int u = 0; u < kClaims; u++)
                if (key[u] != kNoKey) hc[q0 + u * 256] = got[u];
        }
    }
    if (COUNT) return;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4 * kTileItems; r++) {
        if (slot[r] < 0) continue;
        const int64_t pos = (int64_t)hc[slot[r]] + rank[r];
        out_pid[pos] = p[r / 4];
        out_val[pos] = w[r];
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

// The first lane adds lv[0..m) to acc in order, the next eight reads in flight while eight are
// added.
__device__ __forceinline__ double lds_fold(const double* lv, int m, double acc) {
    int r = 0;
    if (m >= 8) {
        double w[8];
#pragma unroll
        for (int q = 0; q < 8; q++) w[q] = lv[q];
        for (r = 8; r + 8 <= m; r += 8) {
            double x[8];
#pragma unroll
            for (int q = 0; q < 8; q++) x[q] = lv[r + q];
#pragma unroll
            for (int q = 0; q < 8; q++) acc = acc + w[q];
#pragma unroll
            for (int q = 0; q < 8; q++) w[q] = x[q];
        }
#pragma unroll
        for (int q = 0; q < 8; q++) acc = acc + w[q];
    }
    for (; r < m; r++) acc = acc + lv[r];
    return acc;
}

// Adds a slab of 1 < m <= 64 * E records to acc in ascending pid.  Record g = e * 64 + lane
// is (key[e], val[e]) (padding: key 0xFFFFFFFF); the wave sorts the (pid, index) pairs in
// registers, puts the weights in sorted order in its LDS slice, and the first lane adds them
// one after the other (acc is meaningful in lane 0 only).
template <int E, int R>
__device__ __forceinline__ double slab_fold_regs(const uint32_t (&key_in)[R], const double (&val)[R], int m, int lane,
                                                 double* lv, double acc) {
    uint32_t key[E], idx[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int g = e * kWave + lane;
        key[e] = key_in[e];
        idx[e] = (uint32_t)g;
        if (g < m) lv[g] = val[e];
    }
    wave_bitonic<E>(key, idx, lane);
    wave_sync_lds();
    // sorted position g = e * 64 + lane holds record idx[e]; the pids are under 2^31, so the
    // padding sorts behind the m records
    double v[E];
#pragma unroll
    for (int e = 0; e < E; e++) v[e] = e * kWave + lane < m ? lv[idx[e]] : 0.0;
    wave_sync_lds();
#pragma unroll
    for (int e = 0; e < E; e++)
        if (e * kWave + lane < m) lv[e * kWave + lane] = v[e];
    wave_sync_lds();
    if (lane == 0) acc = lds_fold(lv, m, acc);
    // lane 0's reads complete before the next slab overwrites the slice
    wave_sync_lds();
    return acc;
}

// Per-wave LDS of the rank sort: the pids grouped by sub-bucket, and the sub-buckets' counts
// and starts (up to 128 each).
struct RankLds {
    uint32_t pid[kSegWaveMax];
    int32_t cnt[2 * kWave];
    int32_t base[2 * kWave];
};

// A sub-bucket longer than this sends the slab to the bitonic network instead.
constexpr int kRankMaxBucket = 24;

// Wave-wide scans through DPP (VALU lane moves, no LDS round trip): within each row of 16
// lanes by row_shr 1, 2, 4 and 8, then across rows by row_bcast:15 (lane 15 of each row into
// the next; rows 1 and 3) and row_bcast:31 (lane 31 into rows 2 and 3).  A lane whose source
// lies outside its row, or whose row is masked, takes the identity id.
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t id, uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)v, CTRL, ROWS, 0xF, false);
}

template <class Op>
__device__ __forceinline__ uint32_t wave_scan_u32(uint32_t v, uint32_t id, Op op) {
    v = op(v, dpp_u32<0x111>(id, v));
    v = op(v, dpp_u32<0x112>(id, v));
    v = op(v, dpp_u32<0x114>(id, v));
    v = op(v, dpp_u32<0x118>(id, v));
    v = op(v, dpp_u32<0x142, 0xA>(id, v));
    v = op(v, dpp_u32<0x143, 0xC>(id, v));
    return v;
}

struct OpMin {
    __device__ __forceinline__ uint32_t operator()(uint32_t a, uint32_t b) const { return min(a, b); }
};
struct OpMax {
    __device__ __forceinline__ uint32_t operator()(uint32_t a, uint32_t b) const { return max(a, b); }
};
struct OpAdd {
    __device__ __forceinline__ uint32_t operator()(uint32_t a, uint32_t b) const { return a + b; }
};

// The wave's minimum / maximum (uniform: the last lane's inclusive scan).
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_readlane((int)wave_scan_u32(v, 0xFFFFFFFFu, OpMin{}), kWave - 1);
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_readlane((int)wave_scan_u32(v, 0u, OpMax{}), kWave - 1);
}

// Adds a slab of 1 < m <= 64 * E records to acc in ascending pid, by rank: the pids' range
// [lo, hi] is cut into 16 * E sub-buckets by b = floor((pid - lo) * 16E / (hi - lo + 1)) (an
// FP64 product of exact integers by one positive scale, so b is non-decreasing in the pid:
// sub-bucket order is pid order); the records are grouped by sub-bucket through LDS counters,
// and each record's rank is the number of smaller pids in its own sub-bucket (2 to 4 records on
// average).  Its weight goes to LDS at (sub-bucket start + rank), and the first lane adds them
// in order.  Returns false, having changed nothing, when a sub-bucket holds more than 24
// records (pids bunched within the range): the caller then sorts the slab with the bitonic
// network.  Record g = e * 64 + lane is (key[e], val[e]).
template <int E, int R>
__device__ __forceinline__ bool slab_fold_rank(const uint32_t (&key)[R], const double (&val)[R], int m, int lane,
                                               RankLds& rl, double* lv, double& acc) {
    constexpr int NB = 16 * E;            // sub-buckets
    constexpr int NBL = NB > kWave ? 2 : 1;  // counters per lane
    uint32_t lo = 0xFFFFFFFFu, hi = 0;
#pragma unroll
    for (int e = 0; e < E; e++)
        if (e * kWave + lane < m) {
            lo = min(lo, key[e]);
            hi = max(hi, key[e]);
        }
    lo = wave_min_u32(lo);
    hi = wave_max_u32(hi);
    const double scale = (double)NB / ((double)(hi - lo) + 1.0);
#pragma unroll
    for (int q = 0; q < NBL; q++)
        if (q * kWave + lane < NB) rl.cnt[q * kWave + lane] = 0;
    wave_sync_lds();
    int b[E], slot[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
        const bool valid = e * kWave + lane < m;
        b[e] = valid ? min(NB - 1, (int)((double)(key[e] - lo) * scale)) : 0;
        slot[e] = valid ? atomicAdd(&rl.cnt[b[e]], 1) : 0;
    }
    wave_sync_lds();
    // lane l owns sub-buckets l * NBL .. l * NBL + NBL - 1
    int c[NBL], t = 0, cmax = 0;
#pragma unroll
    for (int q = 0; q < NBL; q++) {
        c[q] = lane * NBL + q < NB ? rl.cnt[lane * NBL + q] : 0;
        cmax = max(cmax, c[q]);
        t += c[q];
    }
    // exclusive scan of the counts, NBL consecutive buckets per lane
    const int inc = (int)wave_scan_u32((uint32_t)t, 0u, OpAdd{});
    cmax = (int)wave_max_u32((uint32_t)cmax);
    if (cmax > kRankMaxBucket) return false;
    int run = inc - t;
#pragma unroll
    for (int q = 0; q < NBL; q++) {
        if (lane * NBL + q < NB) rl.base[lane * NBL + q] = run;
        run += c[q];
    }
    wave_sync_lds();
    int s0[E], n0[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
        const bool valid = e * kWave + lane < m;
        s0[e] = rl.base[b[e]];
        n0[e] = valid ? rl.cnt[b[e]] : 0;
        if (valid) rl.pid[s0[e] + slot[e]] = key[e];
    }
    wave_sync_lds();
    int r[E];
#pragma unroll
    for (int e = 0; e < E; e++) r[e] = 0;
    // branch-free: every lane reads entry j of its sub-bucket (or, past its end, an in-bounds
    // entry the mask drops), so the unrolled iterations keep several LDS reads in flight; the
    // tests are sign bits (pids are under 2^31: o < key is the sign of o - key; j < n0 that of
    // j - n0), plain VALU with no lane masks
#pragma unroll 4
    for (int j = 0; j < cmax; j++) {
#pragma unroll
        for (int e = 0; e < E; e++) {
            const uint32_t o = rl.pid[min(s0[e] + j, kSegWaveMax - 1)];
            r[e] += (int)(((o - key[e]) & (uint32_t)(j - n0[e])) >> 31);
        }
    }
#pragma unroll
    for (int e = 0; e < E; e++)
        if (e * kWave + lane < m) lv[s0[e] + r[e]] = val[e];
    wave_sync_lds();
    if (lane == 0) acc = lds_fold(lv, m, acc);
    // lane 0's reads complete before the next slab overwrites the slices
    wave_sync_lds();
    return true;
}

// Loads the slab [s, s + m) into registers: record g = e * 64 + lane in (key[e], val[e]).
template <int R>
__device__ __forceinline__ void slab_load(const uint32_t* __restrict__ spid, const double* __restrict__ sval,
                                          int64_t s, int m, int lane, uint32_t (&key)[R], double (&val)[R]) {
#pragma unroll
    for (int e = 0; e < R; e++) {
        const int g = e * kWave + lane;
        key[e] = g < m ? spid[s + g] : kNoKey;
        val[e] = g < m ? sval[s + g] : 0.0;
    }
}

// The same from placed 12-byte records (PlacedRec).
template <int R>
__device__ __forceinline__ void slab_load_rec(const uint32_t* __restrict__ rec, int64_t s, int m, int lane,
                                              uint32_t (&key)[R], double (&val)[R]) {
#pragma unroll
    for (int e = 0; e < R; e++) {
        const int g = e * kWave + lane;
        key[e] = kNoKey;
        val[e] = 0.0;
        if (g < m) get_placed(rec, s + g, key[e], val[e]);
    }
}

// Adds the slab (already in registers) to acc: the smallest register count that holds it, by
// rank, or by the bitonic network where the rank sort declines.
template <int E, int R>
__device__ __forceinline__ double slab_fold_e(const uint32_t (&key)[R], const double (&val)[R], int m, int lane,
                                             RankLds& rl, double* lv, double acc) {
    if (slab_fold_rank<E>(key, val, m, lane, rl, lv, acc)) return acc;
    return slab_fold_regs<E>(key, val, m, lane, lv, acc);
}

template <int R>
__device__ __forceinline__ double slab_fold(const uint32_t (&key)[R], const double (&val)[R], int m, int lane,
                                           RankLds& rl, double* lv, double acc) {
    if (R > 4 && m > 4 * kWave) return slab_fold_e<(R > 4 ? 8 : 4)>(key, val, m, lane, rl, lv, acc);
    if (m > 2 * kWave) return slab_fold_e<4>(key, val, m, lane, rl, lv, acc);
    if (m > kWave) return slab_fold_e<2>(key, val, m, lane, rl, lv, acc);
    if (m > 1) return slab_fold_e<1>(key, val, m, lane, rl, lv, acc);
    if (m == 1 && lane == 0) acc = acc + val[0];  // record 0 is lane 0's first register
    return acc;
}

// One wave per node: the big hot nodes first, then the other hot nodes, then the cold ones (or,
// without a hot node, wave q takes node q).  The node index is a scalar, so every branch on a
// slab's length is wave-uniform.  Slab j + 1 is loaded into registers while slab j is sorted
// and folded.  Slab j's records: [sstart[j], sstart[j + 1]) of (spid, sval) grouped by the tile
// pass, or, PLACED, [sstart[j], sstart[j] + scnt[j]) placed by the push, as (spid, sval) or, AOS,
// PlacedRec records in spid (nothing is done when flags report a record
// that did not fit or a hot slab over its limit: the step is rebuilt).  A hot node with a slab
// over 256 records (never PLACED: the check before it flags such a slab), or a placed cold node
// over 512, is listed for the fallback with its record range instead; a placed hot node's
// chain stops at a slab over 256 and k_seg_slab_cont finishes it.
template <bool PLACED, bool AOS>
__global__ void __launch_bounds__(256)
k_seg_sortfold(const uint32_t* __restrict__ spid, const double* __restrict__ sval, const int32_t* __restrict__ sbase,
               const int32_t* __restrict__ sstart, const int32_t* __restrict__ scnt,
               const int32_t* __restrict__ order, const int32_t* __restrict__ big, int32_t* __restrict__ ctrs,
               int32_t nn, const double* __restrict__ nvol, double spwt, double* __restrict__ out,
               int32_t* __restrict__ ovf_list, int32_t* __restrict__ nstart, int32_t* __restrict__ nend,
               const int32_t* __restrict__ flags) {
    __shared__ double lval[256 / kWave][kSegWaveMax];
    __shared__ RankLds lrank[256 / kWave];
    if (PLACED && (flags[0] | flags[1])) return;
    const int lane = lane_id();
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
    const int32_t q = (int32_t)((int64_t)blockIdx.x * (256 / kWave) + w);
    if (q >= nn) return;
    // a slab's records into registers (AOS: spid holds PlacedRec records, sval is unused)
    auto load = [&](int64_t st, int m, auto& k, auto& v) {
        if constexpr (AOS)
            slab_load_rec(spid, st, m, lane, k, v);
        else
            slab_load(spid, sval, st, m, lane, k, v);
    };
    const int32_t n_hot = __builtin_amdgcn_readfirstlane(ctrs[0]);
    int32_t n = q;
    if (n_hot) {
        const int32_t n_big = __builtin_amdgcn_readfirstlane(ctrs[3]);
        n = q < n_big ? big[q] : q < n_hot ? order[q - n_big] : order[q];
        n = __builtin_amdgcn_readfirstlane(n);
    }
    // the record range of slab j
    auto range = [&](int32_t j, int64_t& s, int& m) {
        s = __builtin_amdgcn_readfirstlane(sstart[j]);
        m = PLACED ? __builtin_amdgcn_readfirstlane(scnt[j])
                   : __builtin_amdgcn_readfirstlane(sstart[j + 1]) - (int32_t)s;
    };
    const int32_t b0 = __builtin_amdgcn_readfirstlane(sbase[n]);
    const int32_t ns = __builtin_amdgcn_readfirstlane(sbase[n + 1]) - b0;
    double* lv = lval[w];
    RankLds& rl = lrank[w];
    double acc = 0.0;
    if (ns == 1) {
        int64_t s;
        int m;
        range(b0, s, m);
        if (PLACED && m > kSegWaveMax) {
            // a cold node's placed slab over one wave's 512: its range is contiguous, so the
            // workgroup fallback sorts it
            if (lane == 0) {
                ovf_list[atomicAdd(ctrs + 2, 1)] = n;
                nstart[n] = (int32_t)s;
                nend[n] = (int32_t)s + m;
            }
            return;
        }
        uint32_t key[8];
        double val[8];
        load(s, m, key, val);
        acc = slab_fold(key, val, m, lane, rl, lv, acc);
    } else if (ns > 1) {
        // a hot node's slabs hold 64..128 records on average; one over 256 (pids bunched within
        // the node's range) sends the node to the fallback
        if (!PLACED) {
            bool over = false;
            for (int32_t j = lane; j < ns; j += kWave)
                over |= sstart[b0 + j + 1] - sstart[b0 + j] > 4 * kWave;
            if (__ballot(over)) {
                if (lane == 0) {
                    ovf_list[atomicAdd(ctrs + 2, 1)] = n;
                    nstart[n] = sstart[b0];
                    nend[n] = sstart[b0 + ns];
                }
                return;
            }
        }
        // the slab ranges 64 at a time, slab q + lane's in this lane (one load for 64 slabs,
        // not two dependent scalar loads before each slab's data), read out by lane index
        int32_t q = 0, rs = 0, rm = 0;
        auto chunk = [&](int32_t q0) {
            q = q0;
            const int32_t j = q0 + lane;
            rs = j < ns ? sstart[b0 + j] : 0;
            rm = j >= ns ? 0 : PLACED ? scnt[b0 + j] : sstart[b0 + j + 1] - rs;
        };
        auto range_at = [&](int32_t j, int64_t& st, int& m) {
            if (j - q >= kWave) chunk(j);
            st = __builtin_amdgcn_readlane(rs, j - q);
            m = __builtin_amdgcn_readlane(rm, j - q);
        };
        chunk(0);
        // slab j + 1 is loaded into the prefetch registers while slab j is folded
        uint32_t pk[4];
        double pv[4];
        int64_t ps;
        int pm;
        range_at(0, ps, pm);
        load(ps, pm, pk, pv);
        for (int32_t j = 0; j < ns; j++) {
            uint32_t ck[4];
            double cv[4];
#pragma unroll
            for (int e = 0; e < 4; e++) {
                ck[e] = pk[e];
                cv[e] = pv[e];
            }
            const int64_t s = ps;
            const int m = pm;
            (void)s;
            if (PLACED && m > 4 * kWave) {
                // a placed slab grown past 256 records since the layout's exact step: the node's
                // chain goes on from slab j in k_seg_slab_cont (listed from the back of
                // ovf_list, its slab index in nstart[n], the sum so far in out[n])
                if (lane == 0) {
                    ovf_list[nn - 1 - atomicAdd(ctrs + 6, 1)] = n;
                    nstart[n] = j;
                    out[n] = acc;
                }
                return;
            }
            if (j + 1 < ns) {
                range_at(j + 1, ps, pm);
                load(ps, pm, pk, pv);
            }
            acc = slab_fold(ck, cv, m, lane, rl, lv, acc);
        }
    }
    if (lane == 0) out[n] = acc * (spwt / nvol[n]);
}

// The placed fold's continuation for the nodes k_seg_sortfold listed with a slab over 256
// records (at most 512: the check flags more): one wave per node, from its recorded slab on,
// each slab loaded whole into 8 registers per lane (a kernel of its own, so the common slab
// loop keeps its smaller register budget); the waves stride over the list (count ctrs[6]).
template <bool AOS>
__global__ void __launch_bounds__(256)
k_seg_slab_cont(const uint32_t* __restrict__ spid, const double* __restrict__ sval, const int32_t* __restrict__ sbase,
                const int32_t* __restrict__ sstart, const int32_t* __restrict__ scnt, const int32_t* __restrict__ ctrs,
                int32_t nn, const double* __restrict__ nvol, double spwt, double* __restrict__ out,
                const int32_t* __restrict__ list, const int32_t* __restrict__ nstart,
                const int32_t* __restrict__ flags) {
    __shared__ double lval[256 / kWave][kSegWaveMax];
    __shared__ RankLds lrank[256 / kWave];
    if (flags[0] | flags[1]) return;
    const int lane = lane_id();
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
    const int32_t count = __builtin_amdgcn_readfirstlane(ctrs[6]);
    const int32_t waves = (int32_t)(gridDim.x * (256 / kWave));
    for (int32_t i = (int32_t)blockIdx.x * (256 / kWave) + w; i < count; i += waves) {
        const int32_t n = __builtin_amdgcn_readfirstlane(list[nn - 1 - i]);
        const int32_t b0 = __builtin_amdgcn_readfirstlane(sbase[n]);
        const int32_t ns = __builtin_amdgcn_readfirstlane(sbase[n + 1]) - b0;
        double acc = out[n];
        for (int32_t j = __builtin_amdgcn_readfirstlane(nstart[n]); j < ns; j++) {
            const int64_t st = __builtin_amdgcn_readfirstlane(sstart[b0 + j]);
            const int m = __builtin_amdgcn_readfirstlane(scnt[b0 + j]);
            uint32_t key[8];
            double val[8];
            if constexpr (AOS)
                slab_load_rec(spid, st, m, lane, key, val);
            else
                slab_load(spid, sval, st, m, lane, key, val);
            acc = slab_fold(key, val, m, lane, lrank[w], lval[w], acc);
        }
        if (lane == 0) out[n] = acc * (spwt / nvol[n]);
    }
}

// The fallback for a hot node with a slab over 256 records (pids bunched within its range):
// each workgroup takes listed nodes in turn and sorts the node's whole range [start, end) by
// pid with `passes` (even) LSD passes of 8 bits, 256 records per round, stable within a round
// by wave ballots (rank among the lanes with the same digit) and per-wave digit counts in LDS.
// Even passes end in (pa, va); (pb, vb) is scratch at the same offsets.
__global__ void __launch_bounds__(256)
k_seg_hot_sort(uint32_t* __restrict__ pa, double* __restrict__ va, uint32_t* __restrict__ pb,
               double* __restrict__ vb, const int32_t* __restrict__ start, const int32_t* __restrict__ end,
               const int32_t* __restrict__ hot_list, const int32_t* __restrict__ hot_count, int passes) {
    __shared__ int32_t base[256];
    __shared__ int32_t wcnt[256 / kWave][256];
    const int tid = threadIdx.x, lane = lane_id();
    const int w = tid / kWave;
    const int32_t count = *hot_count;
    for (int32_t h = blockIdx.x; h < count; h += gridDim.x) {
        const int32_t n = hot_list[h];
        const int64_t s = start[n];
        const int64_t m = end[n] - s;
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

// ---- records placed by the push (det_algo 2) -------------------------------------------------

// The listed nodes' placed ranges [start[n], end[n]) copied from PlacedRec records into (pid,
// weight) arrays at the same offsets, for the workgroup sort; each workgroup takes listed nodes
// in turn.
__global__ void __launch_bounds__(256)
k_place_unpack(const uint32_t* __restrict__ rec, const int32_t* __restrict__ start, const int32_t* __restrict__ end,
               const int32_t* __restrict__ list, const int32_t* __restrict__ count, uint32_t* __restrict__ pid,
               double* __restrict__ val) {
    const int32_t nl = *count;
    for (int32_t h = blockIdx.x; h < nl; h += gridDim.x) {
        const int32_t n = list[h];
        for (int64_t i = start[n] + threadIdx.x; i < end[n]; i += blockDim.x) get_placed(rec, i, pid[i], val[i]);
    }
}

// A slab's segment for the next step: its count now, a quarter more and 32 (the counts change
// by a few per cent from one step to the next; a tighter 1/8 + 16 for hot slabs rebuilt one
// step in 50 of the 200M clustered run).  Capacity costs memory only: the fold reads counts.
__device__ __forceinline__ int32_t place_capacity(int32_t c, int32_t lim) {
    (void)lim;
    return c + (c >> 2) + 32;
}

// The most records a slab may hold for the placed fold: a hot node's slabs 512 (the prefetching
// slab loop takes up to 256, k_seg_slab_cont the node's rest from a larger one on); a cold
// node's one slab has no limit (lim 0: over 512 records its node is sorted by the workgroup
// fallback, its range being contiguous).  One thread per node.
__global__ void __launch_bounds__(256)
k_place_lim(const int32_t* __restrict__ sbase, const int32_t* __restrict__ bits, int32_t nn,
            int32_t* __restrict__ lim) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    const int32_t s0 = sbase[i], b = bits[i];
    if (b == 0) {
        lim[s0] = 0;
        return;
    }
    for (int32_t j = 0; j < (1 << b); j++) lim[s0 + j] = kSegWaveMax;
}

// Capacities (smax + 1 entries, 0 past the *ns slabs) from the exact path's slab counts.
__global__ void __launch_bounds__(256)
k_place_cap(const int32_t* __restrict__ cnt, const int32_t* __restrict__ lim, const int32_t* __restrict__ ns,
            int64_t smax, int32_t* __restrict__ cap) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > smax) return;
    cap[i] = i < *ns ? place_capacity(cnt[i], lim[i]) : 0;
}

// After a placing push: the cursors are the slab counts; a hot node's slab over its limit raises
// flags[1], and the next step's capacities are written (0 past the *ns slabs, for the scan's
// total).  The fallback lists' counts (ctrs[2], ctrs[6]) are cleared for the placed fold.
__global__ void __launch_bounds__(256)
k_place_check(const int32_t* __restrict__ cur, const int32_t* __restrict__ lim, const int32_t* __restrict__ ns,
              int64_t smax, int32_t* __restrict__ flags, int32_t* __restrict__ cap, int32_t* __restrict__ ctrs) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > smax) return;
    if (i == 0) ctrs[2] = ctrs[6] = 0;
    if (i < *ns) {
        const int32_t c = cur[i], l = lim[i];
        if (l && c > l) flags[1] = 1;
        cap[i] = place_capacity(c, l);
    } else {
        cap[i] = 0;
    }
}

// The node counts as the exact path's input (FP64 unit sums), when a placed step is rebuilt:
// node n's slabs [sbase[n], sbase[n + 1]) summed through the exclusive scan of the slab counts.
__global__ void __launch_bounds__(256)
k_place_to_raw(const int32_t* __restrict__ scur, const int32_t* __restrict__ sbase, int32_t nn,
               double* __restrict__ raw) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nn) raw[i] = (double)(scur[sbase[i + 1]] - scur[sbase[i]]);
}

unsigned grid_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

int64_t seg_table_ints(int32_t nn, int64_t m) {
    return 12 * (int64_t)nn + 16 + 3 * (slab_max(nn, m) + 1);
}

int64_t seg_scan_len(int32_t nn, int64_t m) { return slab_max(nn, m) + 1; }

hipError_t launch_seg_fold(const SegFoldArgs& a, hipStream_t st) {
    const int32_t nn = a.nn;
    const int64_t m = 4 * a.n;
    const int64_t smax = slab_max(nn, m);
    const SegTab t = carve(a.tab, nn, m);
    // counters, slab counts and cursors start at zero (one contiguous range)
    hipError_t e = launch_fill_i32(t.ctrs, 8 + (smax + 1) + smax, 0, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_seg_counts, dim3(grid_for((int64_t)nn + 1, 256)), dim3(256), 0, st, a.raw, nn, t.cnt, t.bits,
                       t.nsl, t.order, t.big, t.ctrs);
    e = hipGetLastError();
    if (e == hipSuccess) e = launch_excl_scan(t.nsl, (int64_t)nn + 1, t.sbase, a.scan_tmp, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_slab_prep, dim3(grid_for(nn, 256)), dim3(256), 0, st, t.cnt, t.bits, t.sbase, nn, t.sinfo,
                       t.slab_cnt);
    // tiles of 512 particles, the grid a multiple of the XCD count (xcd_block)
    const unsigned gt = (unsigned)((grid_for(a.n, kTileParts) + kXcds - 1) / kXcds * kXcds);
    if (a.n > 0)
        hipLaunchKernelGGL(k_seg_tile<true>, dim3(gt), dim3(256), 0, st, a.tets, a.soa, a.n, t.sinfo, a.mul, t.ctrs,
                           t.slab_cnt, t.slab_start, t.cur, a.seg_pid, a.seg_val);
    e = hipGetLastError();
    if (e == hipSuccess) e = launch_excl_scan(t.slab_cnt, smax + 1, t.slab_start, a.scan_tmp, st);
    if (e != hipSuccess) return e;
    if (a.n > 0)
        hipLaunchKernelGGL(k_seg_tile<false>, dim3(gt), dim3(256), 0, st, a.tets, a.soa, a.n, t.sinfo, a.mul, t.ctrs,
                           t.slab_cnt, t.slab_start, t.cur, a.seg_pid, a.seg_val);
    hipLaunchKernelGGL((k_seg_sortfold<false, false>), dim3(grid_for(nn, 256 / kWave)), dim3(256), 0, st, a.seg_pid,
                       a.seg_val, t.sbase, t.slab_start, (const int32_t*)nullptr, t.order, t.big, t.ctrs, nn, a.nvol,
                       a.spwt, a.out, t.ovf_list, t.nstart, t.nend, (const int32_t*)nullptr);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    // the fallback (none listed: its workgroups return at once): sort each listed node's range
    // by pid, then fold it with one wave
    const int passes = std::max(2, ((a.pid_bits + 7) / 8 + 1) / 2 * 2);
    const unsigned gh = (unsigned)std::min<int64_t>(1024, std::max<int64_t>(1, nn));
    hipLaunchKernelGGL(k_seg_hot_sort, dim3(gh), dim3(256), 0, st, a.seg_pid, a.seg_val, a.tmp_pid, a.tmp_val, t.nstart,
                       t.nend, t.ovf_list, t.ctrs + 2, passes);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_fold_list(a.seg_val, t.nstart, t.nend, a.nvol, a.spwt, a.out, t.ovf_list, t.ctrs + 2, nn, st);
}

int64_t seg_slab_max(int32_t nn, int64_t m) { return slab_max(nn, m); }

const int2* seg_slab_info(const int32_t* tab, int32_t nn) {
    return carve(const_cast<int32_t*>(tab), nn, 0).sinfo;
}

const int32_t* seg_hot_count(const int32_t* tab, int32_t nn) { return carve(const_cast<int32_t*>(tab), nn, 0).ctrs; }

const int32_t* seg_slab_count(const int32_t* tab, int32_t nn) {
    return carve(const_cast<int32_t*>(tab), nn, 0).sbase + nn;
}

hipError_t launch_place_layout(const int32_t* tab, int32_t nn, int64_t m, int32_t* lim, int32_t* cap, int32_t* base,
                               int32_t* scan_tmp, hipStream_t st) {
    const SegTab t = carve(const_cast<int32_t*>(tab), nn, m);
    const int64_t smax = slab_max(nn, m);
    hipLaunchKernelGGL(k_place_lim, dim3(grid_for(nn, 256)), dim3(256), 0, st, t.sbase, t.bits, nn, lim);
    hipLaunchKernelGGL(k_place_cap, dim3(grid_for(smax + 1, 256)), dim3(256), 0, st, t.slab_cnt, lim,
                       t.sbase + nn, smax, cap);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = launch_excl_scan(cap, smax + 1, base, scan_tmp, st);
    return e;
}

hipError_t launch_place_check(const int32_t* tab, int32_t nn, int64_t smax, const int32_t* cur, const int32_t* lim,
                              int32_t* flags, int32_t* cap, hipStream_t st) {
    hipLaunchKernelGGL(k_place_check, dim3(grid_for(smax + 1, 256)), dim3(256), 0, st, cur, lim,
                       seg_slab_count(tab, nn), smax, flags, cap,
                       carve(const_cast<int32_t*>(tab), nn, 0).ctrs);
    return hipGetLastError();
}

hipError_t launch_place_to_raw(const int32_t* tab, int32_t nn, int64_t smax, const int32_t* cur, int32_t* scur,
                               int32_t* scan_tmp, double* raw, hipStream_t st) {
    hipError_t e = launch_excl_scan(cur, smax + 1, scur, scan_tmp, st);
    if (e != hipSuccess) return e;
    const SegTab t = carve(const_cast<int32_t*>(tab), nn, 0);
    hipLaunchKernelGGL(k_place_to_raw, dim3(grid_for(nn, 256)), dim3(256), 0, st, scur, t.sbase, nn, raw);
    return hipGetLastError();
}

hipError_t launch_place_fold(const PlaceFoldArgs& a, hipStream_t st) {
    const int32_t nn = a.nn;
    const SegTab t = carve(const_cast<int32_t*>(a.tab), nn, 0);
    if (a.rec)
        hipLaunchKernelGGL((k_seg_sortfold<true, true>), dim3(grid_for(nn, 256 / kWave)), dim3(256), 0, st, a.rec,
                           (const double*)nullptr, t.sbase, a.base, a.cur, t.order, t.big, t.ctrs, nn, a.nvol, a.spwt,
                           a.out, t.ovf_list, t.nstart, t.nend, a.flags);
    else
        hipLaunchKernelGGL((k_seg_sortfold<true, false>), dim3(grid_for(nn, 256 / kWave)), dim3(256), 0, st, a.spid,
                           a.sval, t.sbase, a.base, a.cur, t.order, t.big, t.ctrs, nn, a.nvol, a.spwt, a.out,
                           t.ovf_list, t.nstart, t.nend, a.flags);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // the cold nodes listed over 512 records (usually none: the workgroups return at once),
    // sorted in (spid, sval) (PlacedRec records unpacked there first)
    const int passes = std::max(2, ((a.pid_bits + 7) / 8 + 1) / 2 * 2);
    const unsigned gh = (unsigned)std::min<int64_t>(1024, std::max<int64_t>(1, nn));
    if (a.rec) {
        // the nodes whose chain stopped at a slab over 256 records (usually none)
        hipLaunchKernelGGL(k_seg_slab_cont<true>, dim3(64), dim3(256), 0, st, a.rec, (const double*)nullptr, t.sbase,
                           a.base, a.cur, t.ctrs, nn, a.nvol, a.spwt, a.out, t.ovf_list, t.nstart, a.flags);
        hipLaunchKernelGGL(k_place_unpack, dim3(gh), dim3(256), 0, st, a.rec, t.nstart, t.nend, t.ovf_list,
                           t.ctrs + 2, a.spid, a.sval);
    } else {
        hipLaunchKernelGGL(k_seg_slab_cont<false>, dim3(64), dim3(256), 0, st, a.spid, a.sval, t.sbase, a.base, a.cur,
                           t.ctrs, nn, a.nvol, a.spwt, a.out, t.ovf_list, t.nstart, a.flags);
    }
    hipLaunchKernelGGL(k_seg_hot_sort, dim3(gh), dim3(256), 0, st, a.spid, a.sval, a.tmp_pid, a.tmp_val, t.nstart,
                       t.nend, t.ovf_list, t.ctrs + 2, passes);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_fold_list(a.sval, t.nstart, t.nend, a.nvol, a.spwt, a.out, t.ovf_list, t.ctrs + 2, nn, st);
}

}  // namespace mpd
