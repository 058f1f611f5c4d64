// HIP kernels of the particle mover for gfx950 (CDNA4, wave64).
//
//   k_push_soa     push + XtoLtet walk + scatter over the resident SoA particles
//                  (src/mini-pic.cpp:287-308), atomic or deterministic-record scatter
//   k_push_aos     the same over the reference's 88 B Particle records (drop-in path)
//   k_normalise    den = raw * (spwt / node volume), raw cleared (src/mini-pic.cpp:319)
//   k_seg_bounds / k_fold  deterministic scatter: after sorting (node,pid) records, a
//                  serial fold per node in ascending pid, then the normalisation
//   k_cell_hist / k_cell_scatter   counting sort of resident particles by cell >> shift
//   k_scan_*       exclusive prefix sums (counting sort offsets, stable compaction)
//   k_flag_aos / k_compact_aos     stable erase of cell_index < 0 (src/mini-pic.cpp:316)
#include <algorithm>

#include "mp_common.h"
#include "mp_launch.h"

namespace mpd {

// Sums per-block partials [nb][kTally] into the counters: each of up to 64 workgroups folds
// a strided share of the partials and adds 6 totals.
__global__ void __launch_bounds__(1024) k_tally_finish(const int32_t* __restrict__ part, int64_t nb,
                                                       unsigned long long* __restrict__ ctr) {
    __shared__ unsigned long long sh[1024 / kWave][kTally];
    unsigned long long s[kTally] = {0, 0, 0, 0, 0, 0};
    for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += (int64_t)gridDim.x * blockDim.x)
#pragma unroll
        for (int k = 0; k < kTally; k++) s[k] += (unsigned long long)(uint32_t)part[b * kTally + k];
#pragma unroll
    for (int k = 0; k < kTally; k++)
        for (int o = 32; o > 0; o >>= 1) s[k] += __shfl_xor(s[k], o, kWave);
    const int w = threadIdx.x / kWave;
    if (lane_id() == 0)
        for (int k = 0; k < kTally; k++) sh[w][k] = s[k];
    __syncthreads();
    if (threadIdx.x < kTally) {
        unsigned long long t = 0;
        for (int q = 0; q < (int)(blockDim.x / kWave); q++) t += sh[q][threadIdx.x];
        if (t) atomicAdd(ctr + threadIdx.x, t);
    }
}

#ifndef MP_PUSH_WAVES
#define MP_PUSH_WAVES 5
#endif

// Push + XtoLtet walk + scatter of the resident particles, one step.  AGG: 0 = one atomic per
// weight, 2 = LDS hash per workgroup.  Each thread takes ppt particles, one per round, a round
// being blockDim.x consecutive particles; MULTI == false fixes ppt = 1 at compile time (the
// round loop's live counters cost 8 VGPRs, one wave per SIMD of occupancy).  The placing push
// runs fixed workgroups whatever the block size option: DET 3 256 threads writing PlacedRec
// records into keys (pid slabs), DET 4 512 threads writing pids and weights into keys and vals.
template <int WALL, int DET, int AGG, bool MULTI>
__global__ void __launch_bounds__(DET == 4 ? 512 : 256) __attribute__((amdgpu_waves_per_eu(MP_PUSH_WAVES, 8)))
k_push_soa(const TetBlock* __restrict__ tets, Soa s, int64_t n, double qm, double dt,
           double* __restrict__ den, uint32_t* __restrict__ keys, double* __restrict__ vals,
           int32_t* __restrict__ part, int cap, int ppt_arg, bool rec_by_slot, PlaceArgs pl) {
    const int ppt = MULTI ? ppt_arg : 1;
    const int64_t b0 = xcd_block() * ppt;
    int n_act = 0, n_lost = 0, n_cross = 0, n_moves = 0, n_cap = 0, n_refl = 0;
    for (int k = 0; k < ppt; k++) {
        const int64_t i = (b0 + k) * blockDim.x + threadIdx.x;
        bool ok = false;
        int32_t fc = -1;
        uint32_t pid = 0;
        double l0 = 0.0, l1 = 0.0, l2 = 0.0, l3 = 0.0;
        if (i < n) {
            // The particle's state is loaded with its cell, not after it: only the start tet
            // depends on the cell.  Nothing branches on the cell before the walk (a lost slot,
            // cell < 0, evaluates tet 0 with a cap of 0 and its results are dropped), so the
            // compiler cannot sink the state or start-tet loads behind a wait.
            const int32_t c_in = s.cell[i];
            if (DET >= 3) pid = (uint32_t)s.pid[i];
            double vx = ld(s.vx + i), vy = ld(s.vy + i), vz = ld(s.vz + i);
            const double x0 = ld(s.x + i), y0 = ld(s.y + i), z0 = ld(s.z + i);
            // keep the six state loads issued ahead of the cell's first use (the scheduler
            // would otherwise place them after the wait for the cell)
            __builtin_amdgcn_sched_barrier(0);
            const bool live = c_in >= 0;
            int32_t c = live ? c_in : 0;
            const TetView t = tet_at(tets, c);
            StartTet g0;
            load_start(t, g0);
            const double ex = t.ex(), ey = t.ey(), ez = t.ez();
            vx = vx + (qm * ex) * dt;
            vy = vy + (qm * ey) * dt;
            vz = vz + (qm * ez) * dt;
            double x = x0 + vx * dt;
            double y = y0 + vy * dt;
            double z = z0 + vz * dt;
            WalkStats st{0, 0};
            const int r = walk_start<WALL>(tets, g0, c, x, y, z, vx, vy, vz, l0, l1, l2, l3, live ? cap : 0, st);
            if (live) {
                n_act++;
                n_cross += (st.moves - st.refl) > 0;
                n_moves += st.moves;
                n_refl += st.refl;
                if (r == kInside) {
                    s.cell[i] = c;
                    ok = true;
                    fc = c;
                    if (DET == 1) put_records(keys, vals, rec_by_slot ? i : (int64_t)s.pid[i], tet_at(tets, c), l0, l1, l2, l3);
                } else {
                    s.cell[i] = -1;
                    n_lost++;
                    n_cap += (r == kCapped);
                }
                st_(s.x + i, x); st_(s.y + i, y); st_(s.z + i, z);
                st_(s.vx + i, vx); st_(s.vy + i, vy); st_(s.vz + i, vz);
            }
        }
        if constexpr (DET >= 3)
            place_records<DET == 4 ? 512 : 256, DET == 3>(tets, ok, fc, pid, l0, l1, l2, l3, pl, keys, vals,
                                                          MULTI && k > 0);
        else
            scatter_any<DET, AGG>(tets, den, ok, fc, l0, l1, l2, l3);
    }
    tally_counts(part, n_act, n_lost, n_cross, n_moves, n_cap, n_refl);
}

// ---- several steps per launch under a frozen field ------------------------------------------
// K consecutive steps of k_push_soa in one launch.  Each thread keeps its particle's state and
// its current tet's record (coefficients, neighbours, node ids, E) in registers from one step
// to the next: the state crosses HBM once per launch instead of twice per step, and a step in
// which the particle stays in its tet loads nothing.  Step k scatters into den + k * nn.  The
// per-particle arithmetic is k_push_soa's, in the same order, so K launches of k_push_soa and
// one of these leave the same bits (atomic density sums aside).  Atomic scatters only.
// 5 waves per SIMD: the absorbing-wall kernels fit 96 VGPRs, the reflecting ones spill a few
// bytes per lane
#ifndef MP_STEPS_WAVES
#define MP_STEPS_WAVES 5
#endif
// threads per workgroup of k_push_steps, and log2 of its (step, node) LDS hash slots
#ifndef MP_STEPS_THREADS
#define MP_STEPS_THREADS 256
#endif
#ifndef MP_STEPS_BITS
#define MP_STEPS_BITS 11
#endif
// AGG != 0: the (step, node)-keyed LDS hash of step_insert / steps_finish; AGG 0: direct
// atomics.  The lanes run their steps independently (lane_steps); K <= kStepMaxLaunch.
// GATHER: the launch completes a cell sort whose rank pass (k_cell_rank) wrote only the
// permutation: particle i is read from slot perm[i] of src (pid included) and written to slot
// i of s, so the sorted order costs a 4 B index instead of a 56 B record copy.
template <int WALL, int AGG, bool GATHER>
__global__ void __launch_bounds__(MP_STEPS_THREADS) __attribute__((amdgpu_waves_per_eu(MP_STEPS_WAVES, 8)))
k_push_steps(const TetBlock* __restrict__ tets, Soa s, Soa src, const int32_t* __restrict__ perm, int64_t n,
             double qm, double dt, double* __restrict__ den, int32_t nn, int K, int32_t* __restrict__ part,
             int cap) {
    const int64_t i = xcd_block() * blockDim.x + threadIdx.x;
    int n_act = 0, n_lost = 0, n_cross = 0, n_moves = 0, n_cap = 0, n_refl = 0;
    int32_t c_in = -1, pid = 0;
    double x = 0.0, y = 0.0, z = 0.0, vx = 0.0, vy = 0.0, vz = 0.0;
    if (i < n) {
        const int64_t j = GATHER ? (int64_t)perm[i] : i;
        const Soa& r = GATHER ? src : s;
        c_in = r.cell[j];
        vx = ld(r.vx + j);
        vy = ld(r.vy + j);
        vz = ld(r.vz + j);
        x = ld(r.x + j);
        y = ld(r.y + j);
        z = ld(r.z + j);
        if (GATHER) pid = r.pid[j];
    }
    __builtin_amdgcn_sched_barrier(0);
    bool live = c_in >= 0;
    int32_t c = live ? c_in : 0;
    TetRegs t;
    load_tet_regs(tet_at(tets, c), t);
    if (GATHER && i < n) s.pid[i] = pid;
    constexpr bool kHash = AGG != 0;
    __shared__ uint32_t hk[kHash ? (1 << MP_STEPS_BITS) : 1];
    __shared__ double hv[kHash ? (1 << MP_STEPS_BITS) : 1];
    __shared__ int32_t shc[MP_STEPS_THREADS / kWave * 6];
    __shared__ int32_t done;
    if (threadIdx.x == 0) done = 0;
    if (kHash) {
        for (int q = threadIdx.x; q < (1 << MP_STEPS_BITS); q += blockDim.x) {
            hk[q] = 0xFFFFFFFFu;
            hv[q] = 0.0;
        }
    }
    __syncthreads();
    lane_steps<WALL, kHash, MP_STEPS_BITS>(tets, t, c, live, x, y, z, vx, vy, vz, qm, dt, K, cap, den, 0u,
                                           (uint32_t)nn, hk, hv, n_act, n_lost, n_cross, n_moves, n_cap, n_refl);
    if (c_in >= 0) {
        s.cell[i] = live ? c : -1;
        st_(s.x + i, x); st_(s.y + i, y); st_(s.z + i, z);
        st_(s.vx + i, vx); st_(s.vy + i, vy); st_(s.vz + i, vz);
    }
    // no barrier: the last wave of the workgroup flushes the hash (3.8% faster than a
    // workgroup-wide flush at 20M particles, 16 steps per launch)
    steps_finish<MP_STEPS_BITS, kHash>(hk, hv, den, part, shc, &done, n_act, n_lost, n_cross, n_moves, n_cap,
                                       n_refl);
}

// den[k * nn + j] *= spwt / nvol[j] for the K steps of a k_push_steps launch (k_normalise's
// expression; the raw sums are replaced in place).
__global__ void k_normalise_steps(double* __restrict__ den, const double* __restrict__ nvol, int32_t nn,
                                  int64_t total, double spwt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    den[i] = den[i] * (spwt / nvol[i % nn]);
}

// Drop-in kernel over the reference's AoS Particle (88 B: pos@0 vel@24 lc@48 cell@80).
// Absorbing walls, lc written back as the reference kernel does; pid = array index.
template <int DET>
__global__ void __launch_bounds__(256)
k_push_aos(const TetBlock* __restrict__ tets, mp_particle* __restrict__ parts, int64_t n,
           double qm, double dt, double* __restrict__ den, uint32_t* __restrict__ keys,
           double* __restrict__ vals, int32_t* __restrict__ part, int cap) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool act = false, lost = false, crossed = false, capped = false;
    WalkStats st{0, 0};
    if (i < n) {
        mp_particle* p = parts + i;
        int32_t c = p->cell_index;
        act = true;
        const TetView t = tet_at(tets, c);
        double vx = p->vel[0] + (qm * t.ex()) * dt;
        double vy = p->vel[1] + (qm * t.ey()) * dt;
        double vz = p->vel[2] + (qm * t.ez()) * dt;
        double x = p->pos[0] + vx * dt;
        double y = p->pos[1] + vy * dt;
        double z = p->pos[2] + vz * dt;
        double l0, l1, l2, l3;
        const int r = walk<0>(tets, c, x, y, z, vx, vy, vz, l0, l1, l2, l3, cap, st);
        crossed = st.moves > 0;
        p->pos[0] = x; p->pos[1] = y; p->pos[2] = z;
        p->vel[0] = vx; p->vel[1] = vy; p->vel[2] = vz;
        p->lc[0] = l0; p->lc[1] = l1; p->lc[2] = l2; p->lc[3] = l3;
        if (r == kInside) {
            p->cell_index = c;
            const TetView tf = tet_at(tets, c);
            if (DET) {
                put_records(keys, vals, i, tf, l0, l1, l2, l3);
            } else {
                unsafeAtomicAdd(den + tf.con(0), l0);
                unsafeAtomicAdd(den + tf.con(1), l1);
                unsafeAtomicAdd(den + tf.con(2), l2);
                unsafeAtomicAdd(den + tf.con(3), l3);
            }
        } else {
            p->cell_index = -1;
            lost = true;
            capped = (r == kCapped);
        }
    }
    tally(part, act, lost, crossed, st.moves, capped, st.refl);
}

__global__ void k_normalise(double* __restrict__ raw, double* __restrict__ out,
                            const double* __restrict__ nvol, int32_t nn, double spwt) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    out[i] = raw[i] * (spwt / nvol[i]);
    raw[i] = 0.0;
}

__global__ void k_fill_i32(int32_t* __restrict__ p, int64_t n, int32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_f64(double* __restrict__ p, int64_t n, double v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Record range [start[node], end[node]) of each node in the sorted records (keys are node
// ids; sentinels sort last).  start is pre-filled with -1: a node with no record keeps it.
// Segment bounds of the sorted keys: start[node] = first record, end[node] = one past the
// last.  Each thread takes 4 consecutive keys (one 16-byte load) and the two around them.
__global__ void k_seg_bounds(const uint32_t* __restrict__ keys, int64_t m, int32_t nn,
                             int32_t* __restrict__ start, int32_t* __restrict__ end) {
    const int64_t j0 = 4 * ((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (j0 >= m) return;
    uint32_t k[4];
    if (j0 + 4 <= m) {
        const uint4 q = *reinterpret_cast<const uint4*>(keys + j0);  // keys is 16-byte aligned
        k[0] = q.x; k[1] = q.y; k[2] = q.z; k[3] = q.w;
    } else {
#pragma unroll
        for (int i = 0; i < 4; i++) k[i] = j0 + i < m ? keys[j0 + i] : 0xFFFFFFFFu;
    }
    const uint32_t prev = j0 > 0 ? keys[j0 - 1] : 0u;
    const uint32_t next = j0 + 4 < m ? keys[j0 + 4] : 0u;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int64_t j = j0 + i;
        const uint32_t node = k[i];
        if (j >= m || node >= (uint32_t)nn) continue;
        const uint32_t left = i == 0 ? prev : k[i - 1];
        const uint32_t right = i == 3 ? next : k[i + 1];
        if (j == 0 || left != node) start[node] = (int32_t)j;
        if (j == m - 1 || right != node) end[node] = (int32_t)(j + 1);
    }
}

// Serial fold per node in ascending pid: ((0 + v1) + v2) + ..., as a serial MoveParticles
// adds into den; then den *= spwt / node volume.  The fold is a dependent chain per node, so
// the work is laid out to run many chains at once:
//  - k_fold: one lane per node; each lane folds its own segment with kFoldUnroll independent
//    loads in flight per round (a node of a 60^3 box at 20M particles has ~350 records);
//  - segments longer than kFoldLong (dense clusters: 10^4..10^5 records on one node) are
//    appended to a list instead and k_fold_long folds each with a whole wave, one wave per
//    node: coalesced 64-record loads, four ahead of the chain, broadcast lane by lane.  In
//    k_fold a lane's long segment would hold its wave for len / kFoldUnroll memory round
//    trips, and neighbouring hot nodes (consecutive ids) would serialise in one wave.
constexpr int kFoldUnroll = 16;
constexpr int kFoldLong = 1024;

__device__ __forceinline__ void fold_out(double s, int64_t n, double spwt, const double* __restrict__ run_spwt,
                                         int32_t nn_mesh, const double* __restrict__ nvol,
                                         double* __restrict__ out) {
    // batch: key n = run * nn_mesh + node, per-run spwt (stride 3 in the run table)
    const double sp = run_spwt ? run_spwt[3 * (n / nn_mesh)] : spwt;
    out[n] = s * (sp / nvol[n % nn_mesh]);
}

__global__ void __launch_bounds__(256)
k_fold(const double* __restrict__ vals, const int32_t* __restrict__ start, const int32_t* __restrict__ end,
       const double* __restrict__ nvol, int32_t nn, double spwt, const double* __restrict__ run_spwt,
       int32_t nn_mesh, double* __restrict__ out, int32_t* __restrict__ long_list,
       int32_t* __restrict__ long_count) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = n < nn;
    const int64_t j0 = valid ? start[n] : -1;
    const int64_t j1 = j0 >= 0 ? (int64_t)end[n] : j0;
    const bool longseg = j1 - j0 > kFoldLong;
    const unsigned long long lm = __ballot(longseg);
    if (lm) {  // wave-aggregated append of the long segments
        const int lane = lane_id();
        int32_t base = 0;
        const int leader = __ffsll((long long)lm) - 1;
        if (lane == leader) base = atomicAdd(long_count, __popcll(lm));
        base = __shfl(base, leader, kWave);
        if (longseg) long_list[base + __popcll(lm & ((1ull << lane) - 1))] = (int32_t)n;
    }
    if (!valid || longseg) return;
    double s = 0.0;
    int64_t j = j0;
    for (; j + kFoldUnroll <= j1; j += kFoldUnroll) {
        double v[kFoldUnroll];
#pragma unroll
        for (int q = 0; q < kFoldUnroll; q++) v[q] = vals[j + q];
#pragma unroll
        for (int q = 0; q < kFoldUnroll; q++) s = s + v[q];
    }
    for (; j < j1; j++) s = s + vals[j];
    fold_out(s, n, spwt, run_spwt, nn_mesh, nvol, out);
}

// One wave per listed node; every wave strides over the list and leaves when it is done.  Per
// 256 records: coalesced loads (the next chunk's issued before this one is folded), the wave
// writes them to its LDS slice, and its first lane adds them in order from LDS, the reads
// running ahead of the dependent adds (one v_add_f64 per record on the chain).
constexpr int kLongChunk = 4 * kWave;

__global__ void __launch_bounds__(256)
k_fold_long(const double* __restrict__ vals, const int32_t* __restrict__ start,
            const int32_t* __restrict__ end, const double* __restrict__ nvol, double spwt,
            const double* __restrict__ run_spwt, int32_t nn_mesh, double* __restrict__ out,
            const int32_t* __restrict__ long_list, const int32_t* __restrict__ long_count) {
    __shared__ double buf[256 / kWave][kLongChunk];
    const int lane = lane_id();
    double* lb = buf[threadIdx.x / kWave];
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x / kWave);
    const int32_t count = *long_count;
    // the wave's index as a scalar: every loop below has a wave-uniform bound, so the compiler
    // keeps the lanes together around the lane-0 fold
    const int64_t w0 = (int64_t)blockIdx.x * (blockDim.x / kWave) +
                       __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
    for (int64_t i = w0; i < count; i += waves) {
        const int64_t n = long_list[i];
        const int64_t b0 = start[n], b1 = end[n];
        double acc = 0.0;
        double v[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int64_t j = b0 + q * kWave + lane;
            v[q] = j < b1 ? vals[j] : 0.0;
        }
        for (int64_t base = b0; base < b1; base += kLongChunk) {
#pragma unroll
            for (int q = 0; q < 4; q++) lb[q * kWave + lane] = v[q];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // the next chunk's loads fly while lane 0 folds this one
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int64_t j = base + kLongChunk + q * kWave + lane;
                v[q] = j < b1 ? vals[j] : 0.0;
            }
            if (lane == 0) {
                const int64_t left = b1 - base;
                if (left >= kLongChunk) {
#pragma unroll 32
                    for (int r = 0; r < kLongChunk; r++) acc = acc + lb[r];
                } else {
                    for (int r = 0; r < (int)left; r++) acc = acc + lb[r];
                }
            }
            // lane 0's reads complete before the next chunk overwrites the slice
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        if (lane == 0) fold_out(acc, n, spwt, run_spwt, nn_mesh, nvol, out);
    }
}

// ---- counting sort by cell bucket (cell >> shift) ---------------------------------------

// Workgroup-level grouping of bucket keys: a workgroup takes a tile of kSortTile consecutive
// slots (kSortItems per thread, strided by the workgroup size so that each load instruction
// of a wave reads consecutive slots), inserts their keys into an LDS hash (kSortHash >= 2 x
// tile slots) and takes a rank within each key.  The tile's particles of one bucket then
// occupy one contiguous range claimed with a single global atomic.  Tiles are mapped to XCDs
// in contiguous eighths (xcd_block), so the tiles that fill one bucket's range, neighbours
// in near-sorted input, run on one XCD at about the same time and its L2 merges their
// partial lines.
constexpr int kSortThreads = 512;
constexpr int kSortItems = 2;
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kSortHash = 2048;
static_assert(kSortHash % kSortThreads == 0, "whole hash slots per thread");
static_assert(kSortHash >= 2 * kSortTile, "the LDS hash of a tile stays at most half full");

struct SortSlots {
    int32_t key[kSortHash];
    int32_t cnt[kSortHash];  // count, then (k_cell_scatter) the bucket's tile-local offset
    int32_t base[kSortHash];
};

__device__ __forceinline__ int sort_slot(SortSlots& t, int32_t key) {
    uint32_t h = ((uint32_t)key * 0x9E3779B1u) >> (32 - 11);
    for (;;) {
        const int32_t prev = atomicCAS(&t.key[h], -1, key);
        if (prev == -1 || prev == key) return (int)h;
        h = (h + 1) & (kSortHash - 1);
    }
}

// Bucket keys of a thread's kSortItems slots (-1: lost or past n); every load is issued
// before the first key is formed.
// The sort position of cell c: c itself, or its rank along a space-filling curve (trank).
__device__ __forceinline__ int32_t sort_pos(const int32_t* __restrict__ trank, int32_t c) {
    return trank ? trank[c] : c;
}

__device__ __forceinline__ void tile_keys(const int32_t* __restrict__ cell, const uint16_t* __restrict__ run,
                                          const int32_t* __restrict__ trank, int64_t n, int shift, int32_t nb,
                                          int64_t i0, int32_t* key) {
    int32_t c[kSortItems], r[kSortItems];
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
        const int64_t i = i0 + j * kSortThreads;
        c[j] = (i < n) ? cell[i] : -1;
        r[j] = (run && i < n) ? (int32_t)run[i] : 0;
    }
#pragma unroll
    for (int j = 0; j < kSortItems; j++) key[j] = c[j] >= 0 ? r[j] * nb + (sort_pos(trank, c[j]) >> shift) : -1;
}

__device__ __forceinline__ void sort_slots_clear(SortSlots& t) {
    for (int q = threadIdx.x; q < kSortHash; q += kSortThreads) {
        t.key[q] = -1;
        t.cnt[q] = 0;
    }
}

// Slot and rank of this lane's key (-1: none) among the workgroup's keys.  Called by every
// lane of the wave: the lanes that hold the wave's first valid key (most lanes, on cell-sorted
// input) share the leader's slot lookup and one LDS atomic, ranked by lane; the others take
// their own.
__device__ __forceinline__ void sort_take(SortSlots& t, int32_t key, int& slot, int& rank) {
    slot = -1;
    rank = 0;
    const unsigned long long valid = __ballot(key >= 0);
    if (valid == 0) return;
    const int lane = lane_id();
    const int lead = __ffsll((long long)valid) - 1;
    const int32_t k0 = __builtin_amdgcn_readlane(key, lead);
    const unsigned long long same = __ballot(key == k0);
    int s0 = 0, r0 = 0;
    if (lane == lead) {
        s0 = sort_slot(t, k0);
        r0 = atomicAdd(&t.cnt[s0], __popcll(same));
    }
    s0 = __builtin_amdgcn_readlane(s0, lead);
    r0 = __builtin_amdgcn_readlane(r0, lead);
    if (key == k0) {
        slot = s0;
        rank = r0 + __popcll(same & ((1ull << lane) - 1ull));
    } else if (key >= 0) {
        slot = sort_slot(t, key);
        rank = atomicAdd(&t.cnt[slot], 1);
    }
}

// The bucket key refined by a pid digit (db > 0): the top db bits of (pid * mul) >> 32, a
// 16-bit digit non-decreasing in the pid (mul = floor(2^48 / P), every pid < P).
__device__ __forceinline__ int32_t pid_refine(int32_t key, int32_t pid, unsigned long long mul, int db) {
    if (db == 0 || key < 0) return key;
    const uint32_t d = (uint32_t)(((unsigned long long)(uint32_t)pid * mul) >> 32);
    return (key << db) | (int32_t)(d >> (16 - db));
}

// The histogram takes one slot per thread in 1024-thread workgroups and no wave-level
// grouping: on near-sorted input this is faster than the tiled form (79-94 against 95-127 us
// at 20M particles).
constexpr int kHistThreads = 1024;
static_assert(kSortHash >= 2 * kHistThreads, "the LDS hash of a workgroup stays at most half full");

__global__ void __launch_bounds__(kHistThreads)
k_cell_hist(const int32_t* __restrict__ cell, const uint16_t* __restrict__ run, const int32_t* __restrict__ trank,
            int64_t n, int shift, int32_t nb, int32_t* __restrict__ cnt, const int32_t* __restrict__ pid,
            unsigned long long mul, int db) {
    __shared__ SortSlots t;
    for (int q = threadIdx.x; q < kSortHash; q += kHistThreads) {
        t.key[q] = -1;
        t.cnt[q] = 0;
    }
    const int64_t i = (int64_t)blockIdx.x * kHistThreads + threadIdx.x;
    const int32_t c = (i < n) ? cell[i] : -1;
    const int32_t r = (run && i < n) ? (int32_t)run[i] : 0;
    const int32_t p = (db && i < n) ? pid[i] : 0;
    __syncthreads();
    if (c >= 0) atomicAdd(&t.cnt[sort_slot(t, pid_refine(r * nb + (sort_pos(trank, c) >> shift), p, mul, db))], 1);
    __syncthreads();
    for (int q = threadIdx.x; q < kSortHash; q += kHistThreads)
        if (t.key[q] >= 0) atomicAdd(cnt + t.key[q], t.cnt[q]);
}

// One field of the tile through LDS: each thread puts its particles' values at their
// tile-local positions lp (bucket by bucket), and after the barrier stores the positions
// q = tid + j * kSortThreads to their destinations gq, so consecutive lanes store consecutive
// slots of a bucket's range.  buf alternates between two halves from field to field: a half
// is rewritten only after the barrier that follows every read of its previous field.
template <class Put>
__device__ __forceinline__ void stage_field(uint64_t* __restrict__ buf, const int* lp, const uint64_t* v,
                                            const int* gq, int m, Put put) {
#pragma unroll
    for (int j = 0; j < kSortItems; j++)
        if (lp[j] >= 0) buf[lp[j]] = v[j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
        const int q = threadIdx.x + j * kSortThreads;
        if (q < m) put(gq[j], buf[q]);
    }
}

// Moves every particle still in the domain into its bucket's range (cursor = bucket offsets,
// advanced).  The tile's records are loaded with their cells, before the workgroup's hash
// and atomics, so the HBM reads overlap the atomics' round trip instead of following it; they
// leave through LDS in bucket order (stage_field), field by field.
__global__ void __launch_bounds__(kSortThreads)
k_cell_scatter(Soa src, const int32_t* __restrict__ trank, int64_t n, int shift, int32_t nb,
               int32_t* __restrict__ cursor, Soa dst, unsigned long long mul, int db) {
    __shared__ SortSlots t;
    __shared__ int32_t gidx[kSortTile];
    __shared__ uint64_t buf[2][kSortTile];
    __shared__ int32_t n_live;
    const int64_t i0 = xcd_block() * kSortTile + threadIdx.x;
    int32_t c[kSortItems], pid[kSortItems];
    uint16_t run[kSortItems];
    double x[kSortItems], y[kSortItems], z[kSortItems], vx[kSortItems], vy[kSortItems], vz[kSortItems];
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
        const int64_t i = i0 + j * kSortThreads;
        c[j] = -1;
        pid[j] = 0;
        run[j] = 0;
        x[j] = y[j] = z[j] = vx[j] = vy[j] = vz[j] = 0.0;
        if (i < n) {
            c[j] = src.cell[i];
            x[j] = ld(src.x + i); y[j] = ld(src.y + i); z[j] = ld(src.z + i);
            vx[j] = ld(src.vx + i); vy[j] = ld(src.vy + i); vz[j] = ld(src.vz + i);
            pid[j] = src.pid[i];
            if (src.run) run[j] = src.run[i];
        }
    }
    sort_slots_clear(t);
    if (threadIdx.x == 0) n_live = 0;
    __syncthreads();
    int slot[kSortItems], rank[kSortItems];
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
        const int32_t key = pid_refine(
            c[j] >= 0 ? (src.run ? (int32_t)run[j] * nb : 0) + (sort_pos(trank, c[j]) >> shift) : -1, pid[j], mul, db);
        sort_take(t, key, slot[j], rank[j]);
    }
    __syncthreads();
    // each bucket of the tile claims its global range and a tile-local one, a thread's claims
    // all issued before any is waited for
    {
        constexpr int kClaims = kSortHash / kSortThreads;
        int32_t key[kClaims], k[kClaims], gb[kClaims], lb[kClaims];
#pragma unroll
        for (int u = 0; u < kClaims; u++) {
            const int q = threadIdx.x + u * kSortThreads;
            key[u] = t.key[q];
            k[u] = key[u] >= 0 ? t.cnt[q] : 0;
        }
#pragma unroll
        for (int u = 0; u < kClaims; u++) {
            if (key[u] >= 0) {
                gb[u] = atomicAdd(cursor + key[u], k[u]);
                lb[u] = atomicAdd(&n_live, k[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < kClaims; u++) {
            if (key[u] >= 0) {
                const int q = threadIdx.x + u * kSortThreads;
                t.base[q] = gb[u];
                t.cnt[q] = lb[u];
            }
        }
    }
    __syncthreads();
    int lp[kSortItems];
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
        lp[j] = -1;
        if (c[j] >= 0) {
            lp[j] = t.cnt[slot[j]] + rank[j];
            gidx[lp[j]] = t.base[slot[j]] + rank[j];
        }
    }
    uint64_t v[kSortItems];
#pragma unroll
    for (int j = 0; j < kSortItems; j++) v[j] = (uint64_t)__double_as_longlong(x[j]);
#pragma unroll
    for (int j = 0; j < kSortItems; j++)
        if (lp[j] >= 0) buf[0][lp[j]] = v[j];
    __syncthreads();
    const int m = n_live;
    int gq[kSortItems];
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
        const int q = threadIdx.x + j * kSortThreads;
        gq[j] = q < m ? gidx[q] : 0;
        if (q < m) dst.x[gq[j]] = __longlong_as_double((long long)buf[0][q]);
    }
#define MP_STAGE_F64(H, F)                                                                             \
    do {                                                                                               \
        for (int j = 0; j < kSortItems; j++) v[j] = (uint64_t)__double_as_longlong(F[j]);              \
        stage_field(buf[H], lp, v, gq, m,                                                              \
                    [&](int d, uint64_t b) { dst.F[d] = __longlong_as_double((long long)b); });        \
    } while (0)
    MP_STAGE_F64(1, y);
    MP_STAGE_F64(0, z);
    MP_STAGE_F64(1, vx);
    MP_STAGE_F64(0, vy);
    MP_STAGE_F64(1, vz);
#undef MP_STAGE_F64
    for (int j = 0; j < kSortItems; j++) v[j] = (uint64_t)(uint32_t)c[j] | ((uint64_t)(uint32_t)pid[j] << 32);
    stage_field(buf[0], lp, v, gq, m, [&](int d, uint64_t b) {
        dst.cell[d] = (int32_t)(uint32_t)b;
        dst.pid[d] = (int32_t)(uint32_t)(b >> 32);
    });
    if (src.run) {
        for (int j = 0; j < kSortItems; j++) v[j] = run[j];
        stage_field(buf[1], lp, v, gq, m, [&](int d, uint64_t b) { dst.run[d] = (uint16_t)b; });
    }
}

// Claims each of the tile's buckets a range of its global cursor (one returning atomic per
// bucket per tile).
__device__ __forceinline__ void sort_claim(SortSlots& t, int32_t* __restrict__ cursor) {
    constexpr int kClaims = kSortHash / kSortThreads;
    int32_t key[kClaims], got[kClaims];
#pragma unroll
    for (int u = 0; u < kClaims; u++) {
        const int q = threadIdx.x + u * kSortThreads;
        key[u] = t.key[q];
        got[u] = key[u] >= 0 ? t.cnt[q] : 0;
    }
    // all of a thread's claims in flight before any is waited for
#pragma unroll
    for (int u = 0; u < kClaims; u++)
        if (key[u] >= 0) got[u] = atomicAdd(cursor + key[u], got[u]);
#pragma unroll
    for (int u = 0; u < kClaims; u++)
        if (key[u] >= 0) t.base[threadIdx.x + u * kSortThreads] = got[u];
}

// The rank pass of a gather sort: k_cell_scatter's bucket ranges, but only the particle's
// source slot is written (perm[d] = i).  The launch that consumes perm moves the records.
__global__ void __launch_bounds__(kSortThreads)
k_cell_rank(const int32_t* __restrict__ cell, const uint16_t* __restrict__ run, const int32_t* __restrict__ trank,
            int64_t n, int shift, int32_t nb, int32_t* __restrict__ cursor, int32_t* __restrict__ perm) {
    __shared__ SortSlots t;
    const int64_t i0 = xcd_block() * kSortTile + threadIdx.x;
    int32_t key[kSortItems];
    tile_keys(cell, run, trank, n, shift, nb, i0, key);
    sort_slots_clear(t);
    __syncthreads();
    int slot[kSortItems], rank[kSortItems];
#pragma unroll
    for (int j = 0; j < kSortItems; j++) sort_take(t, key[j], slot[j], rank[j]);
    __syncthreads();
    sort_claim(t, cursor);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kSortItems; j++)
        if (key[j] >= 0) perm[(int64_t)t.base[slot[j]] + rank[j]] = (int32_t)(i0 + j * kSortThreads);
}

// ---- exclusive scan (int32), 3 phases --------------------------------------------------

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ int32_t block_excl_scan(int32_t v, int32_t* lds, int32_t& total) {
    const int lane = lane_id();
    const int w = threadIdx.x / kWave;
    int32_t inc = v;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int32_t t = __shfl_up(inc, o, kWave);
        if (lane >= o) inc += t;
    }
    if (lane == kWave - 1) lds[w] = inc;
    __syncthreads();
    int32_t wbase = 0;
    total = 0;
    for (int k = 0; k < kScanThreads / kWave; k++) {
        if (k < w) wbase += lds[k];
        total += lds[k];
    }
    __syncthreads();
    return wbase + inc - v;
}

__global__ void __launch_bounds__(kScanThreads)
k_scan_reduce(const int32_t* __restrict__ in, int64_t n, int32_t* __restrict__ part) {
    __shared__ int32_t lds[kScanThreads / kWave];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++)
        if (base + k < n) s += in[base + k];
    int32_t total;
    block_excl_scan(s, lds, total);
    if (threadIdx.x == 0) part[blockIdx.x] = total;
}

__global__ void __launch_bounds__(kScanThreads)
k_scan_down(const int32_t* __restrict__ in, int64_t n, const int32_t* __restrict__ part_off,
            int32_t* __restrict__ out) {
    __shared__ int32_t lds[kScanThreads / kWave];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int32_t v[kScanItems];
    int32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        v[k] = (base + k < n) ? in[base + k] : 0;
        s += v[k];
    }
    int32_t total;
    int32_t run = block_excl_scan(s, lds, total) + (part_off ? part_off[blockIdx.x] : 0);
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
}

// ---- stable compaction of the AoS drop-in array ---------------------------------------

// first_bad = the lowest index whose cell_index is outside [0, n_tets) (left at its initial
// value when every particle is in range)
__global__ void k_check_cells_aos(const mp_particle* __restrict__ parts, int64_t n, int32_t n_tets,
                                  int32_t* __restrict__ first_bad) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t c = parts[i].cell_index;
    if (c < 0 || c >= n_tets) atomicMin(first_bad, (int32_t)i);
}

__global__ void k_flag_aos(const mp_particle* __restrict__ parts, int64_t n, int32_t* __restrict__ flag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = parts[i].cell_index >= 0 ? 1 : 0;
}

__global__ void k_compact_aos(const mp_particle* __restrict__ in, int64_t n,
                              const int32_t* __restrict__ off, mp_particle* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (in[i].cell_index >= 0) out[off[i]] = in[i];
}

// ---- stable compaction and pid order of the resident particles ---------------------------
// flag[i] = 1 for a particle still in the domain.
__global__ void k_flag_soa(const int32_t* __restrict__ cell, int64_t n, int32_t* __restrict__ flag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = cell[i] >= 0 ? 1 : 0;
}

// Live slots (cell >= 0), one atomic per wave; count was cleared by k_fill_u64.
// One atomic per workgroup: same-address atomics are performed one after another at the
// memory side, so one per wave (800 at 50k particles) took longer than the count itself.
__global__ void __launch_bounds__(256) k_count_live(const int32_t* __restrict__ cell, int64_t n,
                                                    unsigned long long* __restrict__ count) {
    __shared__ unsigned long long part[256 / kWave];
    unsigned long long k = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        k += cell[i] >= 0;
    for (int o = 32; o > 0; o >>= 1) k += __shfl_xor(k, o, kWave);
    if (lane_id() == 0) part[threadIdx.x / kWave] = k;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < 256 / kWave; w++) t += part[w];
        if (t) atomicAdd(count, t);
    }
}

__global__ void k_fill_u64(unsigned long long* __restrict__ p, unsigned long long v) { *p = v; }

// Copy particle i to slot dst_slot(i): off[i] (stable compaction) or off[pid] (pid order).
template <bool BY_PID>
__global__ void k_move_soa(Soa src, int64_t n, const int32_t* __restrict__ off, Soa dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t c = src.cell[i];
    if (c < 0) return;
    const int32_t pid = src.pid[i];
    const int64_t d = off[BY_PID ? (int64_t)pid : i];
    dst.x[d] = src.x[i]; dst.y[d] = src.y[i]; dst.z[d] = src.z[i];
    dst.vx[d] = src.vx[i]; dst.vy[d] = src.vy[i]; dst.vz[d] = src.vz[i];
    dst.cell[d] = c;
    dst.pid[d] = pid;
}

// present[pid] = 1 for every particle still in the domain (present pre-cleared).
__global__ void k_mark_pid(const int32_t* __restrict__ cell, const int32_t* __restrict__ pid, int64_t n,
                           int32_t* __restrict__ present) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && cell[i] >= 0) present[pid[i]] = 1;
}

// ---- field / particle staging --------------------------------------------------------

__global__ void k_set_field(TetBlock* __restrict__ tets, const double* __restrict__ ef, int32_t nt) {
    const int32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nt) return;
    TetChunk* q = tet_chunks(tets, e);
    q[kChVolEx * kTetBlock].d[1] = ef[3 * (int64_t)e + 0];
    q[kChEyEz * kTetBlock].d[0] = ef[3 * (int64_t)e + 1];
    q[kChEyEz * kTetBlock].d[1] = ef[3 * (int64_t)e + 2];
}

__global__ void k_unpack(const double* __restrict__ pos, const double* __restrict__ vel,
                         const int32_t* __restrict__ cell, int64_t n, int64_t pid0, Soa s, int64_t at) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t d = at + i;
    s.x[d] = pos[3 * i + 0]; s.y[d] = pos[3 * i + 1]; s.z[d] = pos[3 * i + 2];
    s.vx[d] = vel[3 * i + 0]; s.vy[d] = vel[3 * i + 1]; s.vz[d] = vel[3 * i + 2];
    s.cell[d] = cell[i];
    s.pid[d] = (int32_t)(pid0 + i);
}

__global__ void k_pack(Soa s, int64_t n, double* __restrict__ pos, double* __restrict__ vel) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    pos[3 * i + 0] = s.x[i]; pos[3 * i + 1] = s.y[i]; pos[3 * i + 2] = s.z[i];
    vel[3 * i + 0] = s.vx[i]; vel[3 * i + 1] = s.vy[i]; vel[3 * i + 2] = s.vz[i];
}

// ---- launchers ------------------------------------------------------------------------

static inline unsigned grid_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

int steps_block() { return MP_STEPS_THREADS; }

hipError_t launch_push_steps(const PushArgs& a, int32_t nn, int k_steps, hipStream_t st) {
    if (a.n <= 0 || k_steps <= 0) return hipSuccess;
    // (step, node) keys are 32-bit offsets below the empty-slot marker
    if ((int64_t)nn * k_steps >= (int64_t)0xFFFFFFFFu) return hipErrorInvalidValue;
    const unsigned g = (grid_for(a.n, MP_STEPS_THREADS) + kXcds - 1) / kXcds * kXcds;
#define MP_STEPS(W, A, G)                                                                                 \
    hipLaunchKernelGGL((k_push_steps<W, A, G>), dim3(g), dim3(MP_STEPS_THREADS), 0, st, a.tets, a.soa, a.src,    \
                       a.perm, a.n, a.qm, a.dt, a.den, nn, k_steps, a.part, a.cap)
#define MP_STEPS_W(W, A)                  \
    do {                                  \
        if (a.perm) MP_STEPS(W, A, true); \
        else MP_STEPS(W, A, false);       \
    } while (0)
    if (k_steps > kStepMaxLaunch) return hipErrorInvalidValue;
    if (a.wall == 0) {
        if (a.agg == 0) MP_STEPS_W(0, 0);
        else MP_STEPS_W(0, 2);
    } else {
        if (a.agg == 0) MP_STEPS_W(1, 0);
        else MP_STEPS_W(1, 2);
    }
#undef MP_STEPS_W
#undef MP_STEPS
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_tally_finish(a.part, g, a.ctr, st);
}

hipError_t launch_normalise_steps(double* den, const double* nvol, int32_t nn, int k_steps, double spwt,
                                  hipStream_t st) {
    const int64_t total = (int64_t)nn * k_steps;
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_normalise_steps, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, den, nvol, nn,
                       total, spwt);
    return hipGetLastError();
}

hipError_t launch_push_soa(const PushArgs& a, hipStream_t st) {
    if (a.n <= 0) return hipSuccess;
    // the placing push's workgroup size is fixed (its LDS hash is sized for it)
    const int bs = a.det == 4 ? 512 : a.det == 3 ? 256 : a.block;
    const int ppt = a.ppt > 0 ? a.ppt : 1;
    const unsigned g = (grid_for(a.n, (int64_t)bs * ppt) + kXcds - 1) / kXcds * kXcds;
#define MP_PUSH(W, D, G)                                                                                  \
    do {                                                                                                  \
        if (ppt == 1)                                                                                     \
            hipLaunchKernelGGL((k_push_soa<W, D, G, false>), dim3(g), dim3(bs), 0, st, a.tets, a.soa, a.n, \
                               a.qm, a.dt, a.den, a.keys, a.vals, a.part, a.cap, 1, a.rec_by_slot, a.place); \
        else                                                                                              \
            hipLaunchKernelGGL((k_push_soa<W, D, G, true>), dim3(g), dim3(bs), 0, st, a.tets, a.soa, a.n,  \
                               a.qm, a.dt, a.den, a.keys, a.vals, a.part, a.cap, ppt, a.rec_by_slot,     \
                               a.place);                                                                  \
    } while (0)
#define MP_PUSH_W(W)                                      \
    do {                                                  \
        if (a.det == 4) MP_PUSH(W, 4, 0);                 \
        else if (a.det == 3) MP_PUSH(W, 3, 0);            \
        else if (a.det == 2 && a.agg == 2) MP_PUSH(W, 2, 2); \
        else if (a.det == 2) MP_PUSH(W, 2, 0);            \
        else if (a.det) MP_PUSH(W, 1, 0);                 \
        else if (a.agg == 2) MP_PUSH(W, 0, 2);            \
        else MP_PUSH(W, 0, 0);                            \
    } while (0)
    if (a.wall == 0) MP_PUSH_W(0);
    else MP_PUSH_W(1);
#undef MP_PUSH_W
#undef MP_PUSH
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_tally_finish(a.part, g, a.ctr, st);
}

hipError_t launch_tally_finish(const int32_t* part, int64_t nb, unsigned long long* ctr, hipStream_t st) {
    const unsigned g = (unsigned)std::min<int64_t>(64, std::max<int64_t>(1, (nb + 4095) / 4096));
    hipLaunchKernelGGL(k_tally_finish, dim3(g), dim3(1024), 0, st, part, nb, ctr);
    return hipGetLastError();
}

// Tally rows for n particles: one per workgroup of `block` threads, or one per wave; sized
// for the finest split (a wave of 64 lanes, four lanes per particle) so any block size,
// threads-per-particle and aggregation fit.
int64_t tally_ints(int64_t n, int block) {
    (void)block;
    return ((4 * n + kWave - 1) / kWave + kXcds * (256 / kWave)) * kTally;
}

hipError_t launch_push_aos(const TetBlock* tets, mp_particle* parts, int64_t n, double qm, double dt,
                           double* den, uint32_t* keys, double* vals, int32_t* part,
                           unsigned long long* ctr, int cap, bool det, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const unsigned g = grid_for(n, 256);
    if (det)
        hipLaunchKernelGGL((k_push_aos<1>), dim3(g), dim3(256), 0, st, tets, parts, n, qm, dt, den,
                           keys, vals, part, cap);
    else
        hipLaunchKernelGGL((k_push_aos<0>), dim3(g), dim3(256), 0, st, tets, parts, n, qm, dt, den,
                           keys, vals, part, cap);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_tally_finish(part, g, ctr, st);
}

hipError_t launch_normalise(double* raw, double* out, const double* nvol, int32_t nn, double spwt,
                            hipStream_t st) {
    hipLaunchKernelGGL(k_normalise, dim3(grid_for(nn, 256)), dim3(256), 0, st, raw, out, nvol, nn, spwt);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t st) {
    return launch_fill_i32((int32_t*)p, n, (int32_t)v, st);
}

hipError_t launch_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n, 256)), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t launch_fill_f64(double* p, int64_t n, double v, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(n, 256)), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t launch_det_fold(const uint32_t* keys, const double* vals, int64_t m, int32_t* bounds,
                           const double* nvol, int32_t nn, double spwt, double* out, hipStream_t st,
                           const double* run_spwt, int32_t nn_mesh) {
    if (nn_mesh <= 0) nn_mesh = nn;
    int32_t* start = bounds;
    int32_t* end = bounds + nn;
    int32_t* long_list = bounds + 2 * (int64_t)nn;
    int32_t* long_count = bounds + 3 * (int64_t)nn;
    hipError_t e = launch_fill_i32(start, nn, -1, st);
    if (e == hipSuccess) e = launch_fill_i32(long_count, 1, 0, st);
    if (e != hipSuccess) return e;
    if (m > 0)
        hipLaunchKernelGGL(k_seg_bounds, dim3(grid_for((m + 3) / 4, 256)), dim3(256), 0, st, keys, m, nn, start,
                           end);
    return launch_fold_nodes(vals, start, end, nvol, nn, spwt, run_spwt, nn_mesh, out, long_list, long_count, st);
}

hipError_t launch_fold_nodes(const double* vals, const int32_t* start, const int32_t* end, const double* nvol,
                             int32_t nn, double spwt, const double* run_spwt, int32_t nn_mesh, double* out,
                             int32_t* long_list, int32_t* long_count, hipStream_t st) {
    hipLaunchKernelGGL(k_fold, dim3(grid_for(nn, 256)), dim3(256), 0, st, vals, start, end, nvol, nn, spwt,
                       run_spwt, nn_mesh, out, long_list, long_count);
    // long segments exist only where a few thousand nodes hold most records: one wave each, so
    // hot nodes do not queue behind each other (with 1024 waves a wave took three hot nodes in
    // turn: 14 ms per step at 200M clustered particles)
    // (up to 4096 workgroups = 16384 waves, one per listed node while the list is that long; the
    // dispatcher hands the queued waves out as others finish)
    const unsigned gl = (unsigned)std::min<int64_t>(4096, std::max<int64_t>(1024, ((int64_t)nn + 3) / 4));
    hipLaunchKernelGGL(k_fold_long, dim3(gl), dim3(256), 0, st, vals, start, end, nvol, spwt, run_spwt,
                       nn_mesh, out, long_list, long_count);
    return hipGetLastError();
}


hipError_t launch_fold_list(const double* vals, const int32_t* start, const int32_t* end, const double* nvol,
                            double spwt, double* out, const int32_t* list, const int32_t* count, int32_t max_count,
                            hipStream_t st) {
    const unsigned gl = (unsigned)std::min<int64_t>(4096, std::max<int64_t>(1, ((int64_t)max_count + 3) / 4));
    // nodes of one mesh: node n's volume is nvol[n % INT32_MAX] = nvol[n]
    hipLaunchKernelGGL(k_fold_long, dim3(gl), dim3(256), 0, st, vals, start, end, nvol, spwt, (const double*)nullptr,
                       (int32_t)0x7FFFFFFF, out, list, count);
    return hipGetLastError();
}

int64_t scan_scratch_ints(int64_t n) {
    int64_t total = 0;
    while (n > kScanTile) {
        n = (n + kScanTile - 1) / kScanTile;
        total += 2 * n;
    }
    return total + 16;
}

hipError_t launch_excl_scan(const int32_t* in, int64_t n, int32_t* out, int32_t* scratch, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (n <= kScanTile) {
        hipLaunchKernelGGL(k_scan_down, dim3(1), dim3(kScanThreads), 0, st, in, n, (const int32_t*)nullptr, out);
        return hipGetLastError();
    }
    const int64_t nb = (n + kScanTile - 1) / kScanTile;
    int32_t* part = scratch;
    int32_t* part_off = scratch + nb;
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(kScanThreads), 0, st, in, n, part);
    hipError_t e = launch_excl_scan(part, nb, part_off, scratch + 2 * nb, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_scan_down, dim3((unsigned)nb), dim3(kScanThreads), 0, st, in, n,
                       (const int32_t*)part_off, out);
    return hipGetLastError();
}

// tiles of kSortTile slots, padded to a multiple of the XCD count (xcd_block)
static unsigned sort_grid(int64_t n) {
    const int64_t tiles = (n + kSortTile - 1) / kSortTile;
    return (unsigned)((tiles + kXcds - 1) / kXcds * kXcds);
}

hipError_t launch_cell_hist(const int32_t* cell, const uint16_t* run, int64_t n, int shift, int32_t nb,
                            int32_t* cnt, hipStream_t st, const int32_t* trank, const int32_t* pid,
                            unsigned long long mul, int db) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_cell_hist, dim3(grid_for(n, kHistThreads)), dim3(kHistThreads), 0, st, cell, run, trank, n,
                       shift, nb, cnt, pid, mul, db);
    return hipGetLastError();
}

hipError_t launch_cell_scatter(const Soa& src, int64_t n, int shift, int32_t nb, int32_t* cursor,
                               const Soa& dst, hipStream_t st, const int32_t* trank, unsigned long long mul, int db) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_cell_scatter, dim3(sort_grid(n)), dim3(kSortThreads), 0, st, src, trank, n, shift,
                       nb, cursor, dst, mul, db);
    return hipGetLastError();
}

hipError_t launch_cell_rank(const int32_t* cell, const uint16_t* run, int64_t n, int shift, int32_t nb,
                            int32_t* cursor, int32_t* perm, hipStream_t st, const int32_t* trank) {
    if (n <= 0) return hipSuccess;
    if (n > (int64_t)INT32_MAX) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_cell_rank, dim3(sort_grid(n)), dim3(kSortThreads), 0, st, cell, run, trank, n, shift,
                       nb, cursor, perm);
    return hipGetLastError();
}

hipError_t launch_count_live(const int32_t* cell, int64_t n, unsigned long long* count, hipStream_t st) {
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(1), 0, st, count, 0ull);
    if (n > 0) {
        // at least 4096 cells per workgroup, at most 1024 workgroups
        const unsigned g = (unsigned)std::min<int64_t>(1024, (n + 4095) / 4096);
        hipLaunchKernelGGL(k_count_live, dim3(g), dim3(256), 0, st, cell, n, count);
    }
    return hipGetLastError();
}

hipError_t launch_flag_soa(const int32_t* cell, int64_t n, int32_t* flag, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_flag_soa, dim3(grid_for(n, 256)), dim3(256), 0, st, cell, n, flag);
    return hipGetLastError();
}

hipError_t launch_move_soa(const Soa& src, int64_t n, const int32_t* off, bool by_pid, const Soa& dst,
                           hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (by_pid)
        hipLaunchKernelGGL(k_move_soa<true>, dim3(grid_for(n, 256)), dim3(256), 0, st, src, n, off, dst);
    else
        hipLaunchKernelGGL(k_move_soa<false>, dim3(grid_for(n, 256)), dim3(256), 0, st, src, n, off, dst);
    return hipGetLastError();
}

hipError_t launch_mark_pid(const int32_t* cell, const int32_t* pid, int64_t n, int32_t* present, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mark_pid, dim3(grid_for(n, 256)), dim3(256), 0, st, cell, pid, n, present);
    return hipGetLastError();
}

hipError_t launch_check_cells_aos(const mp_particle* p, int64_t n, int32_t n_tets, int32_t* first_bad,
                                  hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_check_cells_aos, dim3(grid_for(n, 256)), dim3(256), 0, st, p, n, n_tets, first_bad);
    return hipGetLastError();
}

hipError_t launch_flag_aos(const mp_particle* p, int64_t n, int32_t* flag, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_flag_aos, dim3(grid_for(n, 256)), dim3(256), 0, st, p, n, flag);
    return hipGetLastError();
}

hipError_t launch_compact_aos(const mp_particle* in, int64_t n, const int32_t* off, mp_particle* out,
                              hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_compact_aos, dim3(grid_for(n, 256)), dim3(256), 0, st, in, n, off, out);
    return hipGetLastError();
}

hipError_t launch_set_field(TetBlock* tets, const double* ef, int32_t nt, hipStream_t st) {
    hipLaunchKernelGGL(k_set_field, dim3(grid_for(nt, 256)), dim3(256), 0, st, tets, ef, nt);
    return hipGetLastError();
}

hipError_t launch_unpack(const double* pos, const double* vel, const int32_t* cell, int64_t n, int64_t pid0,
                         const Soa& s, int64_t at, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_unpack, dim3(grid_for(n, 256)), dim3(256), 0, st, pos, vel, cell, n, pid0, s, at);
    return hipGetLastError();
}

hipError_t launch_pack(const Soa& s, int64_t n, double* pos, double* vel, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pack, dim3(grid_for(n, 256)), dim3(256), 0, st, s, n, pos, vel);
    return hipGetLastError();
}

}  // namespace mpd
