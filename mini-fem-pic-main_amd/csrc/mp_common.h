// Device helpers shared by the mover kernels (mp_kernels.hip, mp_batch.hip): wave
// reductions, XCD-aware block mapping, streaming loads, per-block step tallies, the
// aggregated atomic scatters and the deterministic record writer.
#pragma once
#include "mp_device.h"

namespace mpd {

constexpr int kWave = 64;
constexpr int kXcds = 8;

#ifndef MP_NT_STREAM
#define MP_NT_STREAM 1
#endif
// Streaming accesses of the resident particle arrays (read once, written once per step):
// non-temporal, so they do not evict the tet records the walk gathers from L2.
__device__ __forceinline__ double ld(const double* p) {
#if MP_NT_STREAM
    return __builtin_nontemporal_load(p);
#else
    return *p;
#endif
}
__device__ __forceinline__ void st_(double* p, double v) {
#if MP_NT_STREAM
    __builtin_nontemporal_store(v, p);
#else
    *p = v;
#endif
}

// Workgroups are dispatched to the 8 XCDs round-robin.  Map physical block b to a logical
// chunk so that each XCD walks one contiguous eighth of the (cell-sorted) particle array and
// its L2 holds the tets of one slab of space; launchers pad the grid to a multiple of 8.
__device__ __forceinline__ int64_t xcd_block() {
    const int64_t g = gridDim.x, b = blockIdx.x;
    return (b % kXcds) * (g / kXcds) + b / kXcds;
}

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (kWave - 1)); }

__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Block-wide step tallies from per-thread counts (persistent kernels accumulate over chunks).
__device__ __forceinline__ void tally_counts(int32_t* __restrict__ part, int a, int l, int x, int m, int k,
                                             int r) {
    __shared__ int32_t shc[1024 / kWave][6];
    int v[6] = {wave_sum(a), wave_sum(l), wave_sum(x), wave_sum(m), wave_sum(k), wave_sum(r)};
    const int w = threadIdx.x / kWave;
    if (lane_id() == 0)
#pragma unroll
        for (int q = 0; q < 6; q++) shc[w][q] = v[q];
    __syncthreads();
    if (threadIdx.x < 6) {
        int s = 0;
        for (int q = 0; q < (int)(blockDim.x / kWave); q++) s += shc[q][threadIdx.x];
        part[(int64_t)blockIdx.x * 6 + threadIdx.x] = s;
    }
}

// Per-block step tallies: waves reduce with ballots/shuffles, the block sums its waves in
// LDS and thread 0 stores kTally int32 partials to part[blockIdx.x * kTally + k] with plain
// stores.  k_tally_finish folds the partials into the 64-bit counters.  (One atomic per
// wave on a shared counter line serialises in a single L2 channel: ~10 ms at 20M particles.)
constexpr int kTally = 6;

__device__ __forceinline__ void tally(int32_t* __restrict__ part, bool act, bool lost, bool crossed,
                                      int moves, bool capped, int refl) {
    __shared__ int32_t sh[1024 / kWave][kTally];
    int v[kTally];
    v[0] = __popcll(__ballot(act));
    v[1] = __popcll(__ballot(lost));
    v[2] = __popcll(__ballot(crossed));
    v[3] = wave_sum(moves);
    v[4] = __popcll(__ballot(capped));
    v[5] = wave_sum(refl);
    const int w = threadIdx.x / kWave;
    if (lane_id() == 0)
#pragma unroll
        for (int k = 0; k < kTally; k++) sh[w][k] = v[k];
    __syncthreads();
    if (threadIdx.x < kTally) {
        int s = 0;
        for (int q = 0; q < (int)(blockDim.x / kWave); q++) s += sh[q][threadIdx.x];
        part[(int64_t)blockIdx.x * kTally + threadIdx.x] = s;
    }
}

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Atomic scatter aggregated per workgroup by node: the block's (node, weight) pairs are summed
// in an LDS hash table (open addressing, linear probing, ds_cmpst + ds_add_f64), then each
// occupied slot issues one global atomic.  A block of 256 cell-sorted particles touches a few
// dozen nodes, so ~1024 contributions become ~40 global atomics, however the particles
// reordered since the last sort.  A contribution that finds no slot within kNodeProbe probes
// (unsorted particles on a large mesh) is added to HBM directly.  Every thread of the block
// must call it.
constexpr int kHashSlots = 1024;  // tables keyed by tet (one insert per particle)
// 256 slots: ~40 nodes per workgroup after a cell sort; a smaller table is cheaper to clear
// and flush (measured 2% on the whole push against 1024 slots).
#ifndef MP_NODE_HASH_BITS
#define MP_NODE_HASH_BITS 8
#endif
constexpr int kNodeSlots = 1 << MP_NODE_HASH_BITS;
// 1 (default): the hash probe loops stay loops.  Fully unrolled, each of the 32 probes nests
// an exec-mask save: the push then needs 106 SGPRs and spills 264 of them to VGPR lanes (38
// SGPRs and no spill rolled, and 1% faster).
#ifndef MP_PROBE_ROLLED
#define MP_PROBE_ROLLED 1
#endif
constexpr int kNodeProbe = 32;

__device__ __forceinline__ void lds_insert(int32_t* hk, double* hv, double* __restrict__ den, int32_t node,
                                           double w) {
    uint32_t h = ((uint32_t)node * 0x9E3779B1u) >> (32 - MP_NODE_HASH_BITS);
#if MP_PROBE_ROLLED
#pragma unroll 1
#endif
    for (int p = 0; p < kNodeProbe; p++) {
        const int32_t prev = atomicCAS(&hk[h], -1, node);
        if (prev == -1 || prev == node) {
            atomicAdd(&hv[h], w);
            return;
        }
        h = (h + 1) & (kNodeSlots - 1);
    }
    unsafeAtomicAdd(den + node, w);
}

// The same over the final tet's node ids (already in registers).
__device__ __forceinline__ void scatter_lds_nodes(double* __restrict__ den, bool ok, int32_t n0, int32_t n1,
                                                  int32_t n2, int32_t n3, double l0, double l1, double l2,
                                                  double l3) {
    __shared__ int32_t hk[kNodeSlots];
    __shared__ double hv[kNodeSlots];
    for (int s = threadIdx.x; s < kNodeSlots; s += blockDim.x) {
        hk[s] = -1;
        hv[s] = 0.0;
    }
    __syncthreads();
    if (ok) {
        lds_insert(hk, hv, den, n0, l0);
        lds_insert(hk, hv, den, n1, l1);
        lds_insert(hk, hv, den, n2, l2);
        lds_insert(hk, hv, den, n3, l3);
    }
    __syncthreads();
    for (int s = threadIdx.x; s < kNodeSlots; s += blockDim.x) {
        const int32_t node = hk[s];
        if (node >= 0) unsafeAtomicAdd(den + node, hv[s]);
    }
}

__device__ __forceinline__ void scatter_lds(const TetBlock* __restrict__ tets, double* __restrict__ den,
                                            bool ok, int32_t fc, double l0, double l1, double l2,
                                            double l3, int32_t base = 0) {
    int32_t n0 = 0, n1 = 0, n2 = 0, n3 = 0;
    if (ok) {
        const TetView t = tet_at(tets, fc);
        n0 = base + t.con(0);
        n1 = base + t.con(1);
        n2 = base + t.con(2);
        n3 = base + t.con(3);
    }
    scatter_lds_nodes(den, ok, n0, n1, n2, n3, l0, l1, l2, l3);
}

// Deterministic records: particle pid writes its 4 (node, weight) pairs at 4*pid + v, so the
// record array is in pid order; slots of absent pids hold the sentinel node 0xFFFFFFFF
// (pre-filled).  A stable sort on the node bits then leaves each node's weights in
// ascending pid: the order of the serial MoveParticles fold.
__device__ __forceinline__ void put_records_nodes(uint32_t* __restrict__ keys, double* __restrict__ vals,
                                                  int64_t pid, uint32_t n0, uint32_t n1, uint32_t n2, uint32_t n3,
                                                  double l0, double l1, double l2, double l3) {
    const int64_t r = 4 * pid;
    keys[r + 0] = n0;
    keys[r + 1] = n1;
    keys[r + 2] = n2;
    keys[r + 3] = n3;
    vals[r + 0] = l0;
    vals[r + 1] = l1;
    vals[r + 2] = l2;
    vals[r + 3] = l3;
}

__device__ __forceinline__ void put_records(uint32_t* __restrict__ keys, double* __restrict__ vals,
                                            int64_t pid, const TetView tf, double l0,
                                            double l1, double l2, double l3, uint32_t base = 0) {
    put_records_nodes(keys, vals, pid, base + (uint32_t)tf.con(0), base + (uint32_t)tf.con(1),
                      base + (uint32_t)tf.con(2), base + (uint32_t)tf.con(3), l0, l1, l2, l3);
}

// Slabs of a node's records: a node with (first slab, b) = si owns slabs si.x .. si.x + 2^b - 1;
// a record of pid p falls in slab si.x + (d(p) >> (kSlabFineBits - b)), d(p) = (p * mul) >> 32
// the 16-bit pid digit (mul = floor(2^48 / P) for pids below P).  d is non-decreasing in p, so
// slab order is pid order within a node.
constexpr int kSlabFineBits = 16;

__device__ __forceinline__ uint32_t slab_of(int2 si, uint32_t pid, unsigned long long mul) {
    if (si.y == 0) return (uint32_t)si.x;
    const uint32_t d = (uint32_t)(((unsigned long long)pid * mul) >> 32);
    return (uint32_t)si.x + (d >> (kSlabFineBits - si.y));
}

// A placed record: pid and FP64 weight in one aligned 16-byte slot (one dwordx4 store or load;
// a run of records claimed together is one contiguous range instead of two).  Where a
// workgroup's records spread over many pid slabs (runs of a few records per claim), the fewer
// partial 32-byte sectors outweigh the third more bytes (200M clustered particles: placing
// push 13.6 against 15.2 ms); where every node is one slab the runs are long and the separate
// pid and weight arrays are faster (20M uniform particles: 1.18 against 1.29 ms).
struct alignas(16) PlacedRec {
    uint32_t w[kPlacedWords];
};

__device__ __forceinline__ void put_placed(uint32_t* __restrict__ rec, int64_t i, uint32_t pid, double v) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    *reinterpret_cast<PlacedRec*>(rec + kPlacedWords * i) = PlacedRec{{pid, 0u, (uint32_t)b, (uint32_t)(b >> 32)}};
}

__device__ __forceinline__ void get_placed(const uint32_t* __restrict__ rec, int64_t i, uint32_t& pid, double& v) {
    const PlacedRec r = *reinterpret_cast<const PlacedRec*>(rec + kPlacedWords * i);
    pid = r.w[0];
    v = __longlong_as_double((long long)((uint64_t)r.w[2] | ((uint64_t)r.w[3] << 32)));
}

// Deterministic records placed by the push (det_algo 2): a record goes straight into its slab's
// segment [base[s], base[s + 1]) of the placed record array, sized from the previous step's
// count with slack (a cold node's one slab is the node; a hot node's slab is picked by the pid
// digit, pl.sinfo).  Every thread of the block calls it.  The block groups its records by slab
// in an LDS hash (one insert per wave for the lanes that share the wave's first slab, as
// neighbouring cell-sorted particles do), claims one range per slab with one returning atomic on
// the slab's cursor, and writes its PlacedRec at base + claim + rank; a record that finds no
// slot within the probe limit claims its place alone.  A record past its segment's capacity is
// not written and raises flags[0]: the caller then rebuilds the step's records exactly.  The
// cursors end as the slabs' record counts.
// The placing push's workgroups: 256 threads (a 2^10-slot hash) or 512 (2^11): 4 slots per
// thread.  512 doubles the records per slab claim, which pays where every node is one slab (9%
// on the placing push at 20M uniform particles) and costs where the records of a workgroup
// spread over many pid slabs (3% at 200M clustered ones).
constexpr int place_bits(int threads) { return threads > 256 ? 11 : 10; }
constexpr int kPlaceProbe = 32;

template <int BITS>
__device__ __forceinline__ int place_slot(int32_t* hk, int32_t key) {
    uint32_t h = ((uint32_t)key * 0x9E3779B1u) >> (32 - BITS);
    constexpr uint32_t mask = (1u << BITS) - 1u;
#if MP_PROBE_ROLLED
#pragma unroll 1
#endif
    for (int p = 0; p < kPlaceProbe; p++) {
        const int32_t prev = atomicCAS(&hk[h], -1, key);
        if (prev == -1 || prev == key) return (int)h;
        h = (h + 1) & mask;
    }
    return -1;
}

// Slot (-1: none) and rank of this lane's slab key (-1: no record) in the block's table; every
// lane of the wave calls it.
template <int BITS>
__device__ __forceinline__ void place_take(int32_t* hk, int32_t* hc, int32_t key, int& slot, int& rank) {
    slot = -1;
    rank = 0;
    const unsigned long long valid = __ballot(key >= 0);
    if (valid == 0) return;
    const int lane = lane_id();
    const int lead = __ffsll((long long)valid) - 1;
    const int32_t k0 = __builtin_amdgcn_readlane(key, lead);
    const unsigned long long same = __ballot(key == k0);
    int s0 = -1, r0 = 0;
    if (lane == lead) {
        s0 = place_slot<BITS>(hk, k0);
        if (s0 >= 0) r0 = atomicAdd(&hc[s0], __popcll(same));
    }
    s0 = __builtin_amdgcn_readlane(s0, lead);
    r0 = __builtin_amdgcn_readlane(r0, lead);
    if (key == k0) {
        if (s0 >= 0) {
            slot = s0;
            rank = r0 + __popcll(same & ((1ull << lane) - 1ull));
        }
    } else if (key >= 0) {
        slot = place_slot<BITS>(hk, key);
        if (slot >= 0) rank = atomicAdd(&hc[slot], 1);
    }
}

// THREADS: the workgroup size.  `again`: not the block's first call (several rounds per
// thread): its records' reads of the claimed offsets (any slot) finish before the table is
// cleared.
// AOS: records as PlacedRec into out_pid (out_val unused); else pid and weight into out_pid and
// out_val at the same index.
template <int THREADS, bool AOS>
__device__ __forceinline__ void place_records(const TetBlock* __restrict__ tets, bool ok, int32_t fc, uint32_t pid,
                                              double l0, double l1, double l2, double l3, const PlaceArgs& pl,
                                              uint32_t* __restrict__ out_pid, double* __restrict__ out_val,
                                              bool again) {
    constexpr int kBits = place_bits(THREADS);
    constexpr int slots = 1 << kBits;
    __shared__ int32_t hk[slots];
    __shared__ int32_t hc[slots];
    if (again) __syncthreads();
    for (int s = threadIdx.x; s < slots; s += blockDim.x) {
        hk[s] = -1;
        hc[s] = 0;
    }
    int32_t nd[4] = {-1, -1, -1, -1};
    int2 seg[4];  // each record's segment [base[s], base[s + 1])
    if (ok) {
        const TetView t = tet_at(tets, fc);
        nd[0] = t.con(0);
        nd[1] = t.con(1);
        nd[2] = t.con(2);
        nd[3] = t.con(3);
        if (pl.sinfo) {
#pragma unroll
            for (int v = 0; v < 4; v++) nd[v] = (int32_t)slab_of(pl.sinfo[nd[v]], pid, pl.mul);
        }
    }
    __syncthreads();
    int slot[4], rank[4];
#pragma unroll
    for (int v = 0; v < 4; v++) place_take<kBits>(hk, hc, nd[v], slot[v], rank[v]);
    __syncthreads();
    // the segment bounds (they do not depend on the claims) and the block's claims are all in
    // flight together: one returning atomic per occupied slot, a thread's up to kPlaceClaim
    // claims issued before any is waited for (a barrier waits on the loads issued before it)
    if (ok) {
#pragma unroll
        for (int v = 0; v < 4; v++) seg[v] = make_int2(pl.base[nd[v]], pl.base[nd[v] + 1]);
    }
    constexpr int kPlaceClaim = 4;
    for (int s0 = threadIdx.x; s0 < slots; s0 += kPlaceClaim * blockDim.x) {
        int32_t key[kPlaceClaim], got[kPlaceClaim];
#pragma unroll
        for (int u = 0; u < kPlaceClaim; u++) {
            const int s = s0 + u * blockDim.x;
            key[u] = s < slots ? hk[s] : -1;
            got[u] = key[u] >= 0 ? hc[s] : 0;
        }
#pragma unroll
        for (int u = 0; u < kPlaceClaim; u++)
            if (key[u] >= 0) got[u] = atomicAdd(pl.cur + key[u], got[u]);
#pragma unroll
        for (int u = 0; u < kPlaceClaim; u++)
            if (key[u] >= 0) hc[s0 + u * blockDim.x] = got[u];
    }
    __syncthreads();
    if (!ok) return;
    // every offset first (the claimed ones from LDS, then the rare record that found no slot
    // claims its own), the stores after: no wait on a memory result follows a store, so none
    // waits for an earlier record's stores to complete
    int32_t g[4];
#pragma unroll
    for (int v = 0; v < 4; v++) g[v] = slot[v] >= 0 ? hc[slot[v]] + rank[v] : -1;
#pragma unroll
    for (int v = 0; v < 4; v++)
        if (slot[v] < 0) g[v] = atomicAdd(pl.cur + nd[v], 1);
    const double w[4] = {l0, l1, l2, l3};
    bool fit = true;
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int32_t b = seg[v].x;
        if (g[v] < seg[v].y - b) {
            if (AOS) {
                put_placed(out_pid, (int64_t)b + g[v], pid, w[v]);
            } else {
                out_pid[b + g[v]] = pid;
                out_val[b + g[v]] = w[v];
            }
        } else {
            fit = false;
        }
    }
    if (!fit) pl.flags[0] = 1;
}

// The atomic scatter a kernel's template arguments select (nothing when DET == 1: the records
// are written where each walk ends; unit weights when DET == 2, which counts each node's
// records): AGG 0 one atomic per weight, 2 LDS hash per workgroup (every thread of the block
// calls it).
template <int DET, int AGG>
__device__ __forceinline__ void scatter_any(const TetBlock* __restrict__ tets, double* __restrict__ den, bool ok,
                                            int32_t fc, double l0, double l1, double l2, double l3) {
    if (DET == 1) return;
    if (DET == 2) l0 = l1 = l2 = l3 = 1.0;  // record counts per node
    if (AGG == 2) {
        scatter_lds(tets, den, ok, fc, l0, l1, l2, l3);
    } else if (ok) {
        const TetView t = tet_at(tets, fc);
        unsafeAtomicAdd(den + t.con(0), l0);
        unsafeAtomicAdd(den + t.con(1), l1);
        unsafeAtomicAdd(den + t.con(2), l2);
        unsafeAtomicAdd(den + t.con(3), l3);
    }
}

// Multi-step kernels: one LDS hash per workgroup keyed by (step, node) = the offset of the
// contribution in the launch's densities (step k of a launch at k * nodes per step), so the
// steps need no workgroup barrier between them and lanes may be at different steps; the table
// is flushed (one global atomic per occupied slot) once, by the workgroup's last wave to
// finish (steps_finish), at the end of a launch of at most kStepMaxLaunch steps.  A
// contribution that finds no slot within kStepProbe probes (more distinct keys than the table
// holds, e.g. unsorted particles) is added to HBM directly.
// The table has 1 << BITS slots (12 B each).
constexpr int kStepBits = 11;
constexpr int kStepSlots = 1 << kStepBits;
constexpr int kStepProbe = 32;

template <int BITS = kStepBits>
__device__ __forceinline__ void step_insert(uint32_t* hk, double* hv, double* __restrict__ den, uint32_t key,
                                            double w) {
    uint32_t h = (key * 0x9E3779B1u) >> (32 - BITS);
#if MP_PROBE_ROLLED
#pragma unroll 1
#endif
    for (int p = 0; p < kStepProbe; p++) {
        const uint32_t prev = atomicCAS(&hk[h], 0xFFFFFFFFu, key);
        if (prev == 0xFFFFFFFFu || prev == key) {
            atomicAdd(&hv[h], w);
            return;
        }
        h = (h + 1) & ((1u << BITS) - 1);
    }
    unsafeAtomicAdd(den + key, w);
}

// The end of a multi-step workgroup without a barrier: each wave posts its counters in LDS
// (shc, 6 per wave) and leaves; the last wave to arrive (the done counter) flushes the
// (step, node) hash (HASH) and writes the workgroup's six tallies, as tally_counts does.  Waves
// that finish early free their slots instead of waiting for the workgroup's slowest lane.
// Every lane of every wave calls it once, after its last use of the hash.
template <int BITS, bool HASH>
__device__ __forceinline__ void steps_finish(uint32_t* hk, double* hv, double* __restrict__ den,
                                             int32_t* __restrict__ part, int32_t* shc, int32_t* done, int a,
                                             int l, int x, int m, int k, int r) {
    const int v[6] = {wave_sum(a), wave_sum(l), wave_sum(x), wave_sum(m), wave_sum(k), wave_sum(r)};
    const int lane = lane_id();
    const int w = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < 6; q++) shc[w * 6 + q] = v[q];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    int prev = 0;
    if (lane == 0) prev = atomicAdd(done, 1);
    prev = __builtin_amdgcn_readfirstlane(prev);
    if (prev != nw - 1) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    if (HASH) {
        for (int s = lane; s < (1 << BITS); s += kWave) {
            const uint32_t key = hk[s];
            if (key != 0xFFFFFFFFu) unsafeAtomicAdd(den + key, hv[s]);
        }
    }
    if (lane < 6) {
        int sum = 0;
        for (int q = 0; q < nw; q++) sum += shc[q * 6 + lane];
        part[(int64_t)blockIdx.x * 6 + lane] = sum;
    }
}

// K steps of one particle whose tet record t is in registers, each lane on its own: one
// iteration of the loop is one barycentric evaluation (preceded by the push when the lane
// starts a step); a lane whose particle moves loads the next tet's record for its next
// iteration, and a lane that finishes a step goes on with the next one while other lanes still
// walk.  A wave thus takes about max over its lanes of (K + moves) iterations instead of the
// sum over the steps of its slowest walk.  The arithmetic of every step is walk()'s after the
// push, in the same order.  Step k's weights go to key0 + k * per_step + node: the LDS hash
// (HASH) or direct atomics on den.  Counters as in k_push_soa.
template <int WALL, bool HASH, int BITS = kStepBits>
__device__ __forceinline__ void lane_steps(const TetBlock* __restrict__ tets, TetRegs& t, int32_t& c, bool& live,
                                           double& x, double& y, double& z, double& vx, double& vy, double& vz,
                                           double qm, double dt, int K, int cap, double* __restrict__ den,
                                           uint32_t key0, uint32_t per_step, uint32_t* hk, double* hv, int& n_act,
                                           int& n_lost, int& n_cross, int& n_moves, int& n_cap, int& n_refl) {
    // few live counters: steps started = k (+1 for a step that loses the particle), the moves
    // summed, crossings and reflections packed (crossings <= K < 256 in the low byte)
    int k = 0, moves = 0, cross_refl = 0;
    bool start = true, lost = false, capped = false;
    WalkStats st{0, 0};
    bool active = live && K > 0;
    while (active) {
        if (start) {
            vx = vx + (qm * t.ex) * dt;
            vy = vy + (qm * t.ey) * dt;
            vz = vz + (qm * t.ez) * dt;
            x = x + vx * dt;
            y = y + vy * dt;
            z = z + vz * dt;
            st.moves = 0;
            st.refl = 0;
            start = false;
        }
        const StartTet& g = t.g;
        const double l0 = bary(g.a0, g.b0, g.g0, g.d0, x, y, z, g.vol);
        const double l1 = bary(g.a1, g.b1, g.g1, g.d1, x, y, z, g.vol);
        const double l2 = bary(g.a2, g.b2, g.g2, g.d2, x, y, z, g.vol);
        const double l3 = bary(g.a3, g.b3, g.g3, g.d3, x, y, z, g.vol);
        const bool out = (l0 < 0 || l0 > 1.0) || (l1 < 0 || l1 > 1.0) || (l2 < 0 || l2 > 1.0) ||
                         (l3 < 0 || l3 > 1.0);
        if (!out) {
            const uint32_t base = key0 + (uint32_t)k * per_step;
            if (HASH) {
                step_insert<BITS>(hk, hv, den, base + (uint32_t)t.m0, l0);
                step_insert<BITS>(hk, hv, den, base + (uint32_t)t.m1, l1);
                step_insert<BITS>(hk, hv, den, base + (uint32_t)t.m2, l2);
                step_insert<BITS>(hk, hv, den, base + (uint32_t)t.m3, l3);
            } else {
                unsafeAtomicAdd(den + base + t.m0, l0);
                unsafeAtomicAdd(den + base + t.m1, l1);
                unsafeAtomicAdd(den + base + t.m2, l2);
                unsafeAtomicAdd(den + base + t.m3, l3);
            }
            moves += st.moves;
            cross_refl += ((st.moves - st.refl) > 0) + (st.refl << 8);
            k++;
            start = true;
            active = k < K;
            continue;
        }
        int mi = 0;
        double ml = l0;
        if (l1 < ml) { ml = l1; mi = 1; }
        if (l2 < ml) { ml = l2; mi = 2; }
        if (l3 < ml) { ml = l3; mi = 3; }
        const int32_t nb = mi == 0 ? g.n0 : mi == 1 ? g.n1 : mi == 2 ? g.n2 : g.n3;
        capped = st.moves >= cap;
        if (!capped) st.moves++;
        if (!capped && nb >= 0) {
            c = nb;
            load_tet_regs(tet_at(tets, nb), t);
            continue;
        }
        if (capped || WALL == 0) {
            moves += st.moves;
            cross_refl += ((st.moves - st.refl) > 0) + (st.refl << 8);
            lost = true;
            live = false;
            active = false;
            continue;
        }
        // reflection about the face plane, as in walk()
        const double b = mi == 0 ? g.b0 : mi == 1 ? g.b1 : mi == 2 ? g.b2 : g.b3;
        const double gm = mi == 0 ? g.g0 : mi == 1 ? g.g1 : mi == 2 ? g.g2 : g.g3;
        const double d = mi == 0 ? g.d0 : mi == 1 ? g.d1 : mi == 2 ? g.d2 : g.d3;
        const double gx = -b, gy = gm, gz = -d;
        const double gg = (gx * gx + gy * gy) + gz * gz;
        const double sc = ((2.0 * ml) * (6.0 * g.vol)) / gg;
        x = x - sc * gx;
        y = y - sc * gy;
        z = z - sc * gz;
        const double w = (2.0 * ((vx * gx + vy * gy) + vz * gz)) / gg;
        vx = vx - w * gx;
        vy = vy - w * gy;
        vz = vz - w * gz;
        st.refl++;
    }
    n_act += k + lost;
    n_lost += lost;
    n_cap += lost && capped;
    n_moves += moves;
    n_cross += cross_refl & 0xFF;
    n_refl += cross_refl >> 8;
}

}  // namespace mpd
