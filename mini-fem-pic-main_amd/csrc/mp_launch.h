// Host-side launch wrappers of the mover kernels (implemented in mp_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mp_abi.h"
#include "mp_device.h"

namespace mpd {

enum CounterIndex : int {
    kCtrParticleSteps = 0,
    kCtrLost = 1,
    kCtrCrossed = 2,
    kCtrMoves = 3,
    kCtrCapHits = 4,
    kCtrReflections = 5,
    kNumCounters = 8,
};

struct PushArgs {
    const TetBlock* tets;
    Soa soa;
    int64_t n;
    double qm;
    double dt;
    double* den;
    uint32_t* keys;  // deterministic records: node of slot 4*pid+v
    double* vals;    //                        weight of slot 4*pid+v
    int32_t* part;  // per-block or per-wave tallies, tally_ints(n, block) ints
    unsigned long long* ctr;
    int cap;
    int wall;
    int det;  // 0 atomic scatter; 1 records only; 2 no records: each node's count of particles
              // ending in its tets added to den (unit weights), the records being rebuilt from
              // the pushed state afterwards (launch_seg_fold); 3 and 4 the records placed into
              // the slab segments of `place`: 3 by 256-thread workgroups as PlacedRec records
              // in keys (vals unused), 4 by 512-thread ones as pids in keys, weights in vals
    int agg;
    int block;
    int ppt;  // particles per thread (rounds of blockDim.x consecutive particles)
    bool rec_by_slot;  // deterministic records at 4*slot+v (slots in pid order), else 4*pid+v
    PlaceArgs place{};  // det 3 only
    // launch_push_steps only: with perm set, particle i is read from slot perm[i] of src and
    // written to slot i of soa (the gather half of a cell sort)
    Soa src{};
    const int32_t* perm = nullptr;
};

hipError_t launch_push_soa(const PushArgs& a, hipStream_t st);
// k_steps <= kStepMaxLaunch steps in one launch (atomic scatter; aggregation 0 = direct
// atomics, any other = the (step, node) LDS hash); step k scatters into a.den + k * nn.
hipError_t launch_push_steps(const PushArgs& a, int32_t nn, int k_steps, hipStream_t st);
// workgroup size of launch_push_steps (its per-block tallies: tally_ints(n, steps_block()))
int steps_block();
hipError_t launch_normalise_steps(double* den, const double* nvol, int32_t nn, int k_steps, double spwt,
                                  hipStream_t st);
hipError_t launch_push_aos(const TetBlock* tets, mp_particle* parts, int64_t n, double qm, double dt,
                           double* den, uint32_t* keys, double* vals, int32_t* part,
                           unsigned long long* ctr, int cap, bool det, hipStream_t st);
// Step counters: the push kernels leave per-block partials; this folds them into ctr.
hipError_t launch_tally_finish(const int32_t* part, int64_t nb, unsigned long long* ctr, hipStream_t st);
int64_t tally_ints(int64_t n, int block);
hipError_t launch_normalise(double* raw, double* out, const double* nvol, int32_t nn, double spwt,
                            hipStream_t st);
hipError_t launch_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t st);
hipError_t launch_fill_f64(double* p, int64_t n, double v, hipStream_t st);
// Deterministic fold over nn keys; with run_spwt (batch) key k = run * nn_mesh + node and the
// run's spwt is run_spwt[3 * run].
// bounds: 3 * nn + 1 ints of scratch (segment start and end per node, long-segment list, count).
hipError_t launch_det_fold(const uint32_t* keys, const double* vals, int64_t m, int32_t* bounds,
                           const double* nvol, int32_t nn, double spwt, double* out, hipStream_t st,
                           const double* run_spwt = nullptr, int32_t nn_mesh = 0);
hipError_t launch_fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t st);
// Deterministic scatter by node segments (mp_segfold.hip): after a det = 2 push, every
// particle's records rebuilt from the pushed state, grouped by node (hot nodes cut into slabs by
// pid range) and folded in ascending pid into out (normalised by spwt / node volume).
struct SegFoldArgs {
    const TetBlock* tets;
    Soa soa;             // the pushed particles (cell < 0: none)
    int64_t n;           // slots
    uint32_t* tmp_pid;   // 4 * n entries of scratch (the fallback sort's second buffer)
    double* tmp_val;
    double* raw;         // per-node record counts summed by the push; cleared here
    int32_t nn;
    int32_t* tab;        // seg_table_ints(nn, 4 * n) ints
    int32_t* scan_tmp;   // scan_scratch_ints(seg_scan_len(nn, 4 * n)) ints
    uint32_t* seg_pid;   // 4 * n entries: the records grouped by slab
    double* seg_val;
    const double* nvol;
    double spwt;
    double* out;
    int pid_bits;            // bits of the largest pid
    unsigned long long mul;  // floor(2^48 / P), P = the next pid to issue (>= 1)
};
int64_t seg_table_ints(int32_t nn, int64_t m);
// the longest prefix sum launch_seg_fold takes
int64_t seg_scan_len(int32_t nn, int64_t m);
hipError_t launch_seg_fold(const SegFoldArgs& a, hipStream_t st);
// Records placed by the push (det_algo 2), mp_segfold.hip.  The placement keeps the slab layout
// of the last exact step (launch_seg_fold's table tab for nn nodes and m records): a cold node's
// one slab, a hot node's 2^b pid slabs.  Its arrays have smax + 1 = seg_slab_max(nn, m) + 1
// entries (lim, cap, base, cur; the entries past the layout's slabs stay 0).
int64_t seg_slab_max(int32_t nn, int64_t m);
// (first slab, b) per node, as the placing push keys its records by.
const int2* seg_slab_info(const int32_t* tab, int32_t nn);
// Where launch_seg_fold leaves its count of hot nodes (one int) and of slabs (one int).
const int32_t* seg_hot_count(const int32_t* tab, int32_t nn);
const int32_t* seg_slab_count(const int32_t* tab, int32_t nn);
// After an exact step: each slab's limit (lim: 256 for a hot node's slab, 0 = none for a cold
// node's), capacity (cap) and segment start (base, smax + 1 entries, base[smax] the total).
hipError_t launch_place_layout(const int32_t* tab, int32_t nn, int64_t m, int32_t* lim, int32_t* cap, int32_t* base,
                               int32_t* scan_tmp, hipStream_t st);
// After a placing push (the cursors cur are the slab counts): flags[1] = a slab over its limit,
// cap = the next step's capacities; the table's fallback count cleared for the placed fold.
hipError_t launch_place_check(const int32_t* tab, int32_t nn, int64_t smax, const int32_t* cur, const int32_t* lim,
                              int32_t* flags, int32_t* cap, hipStream_t st);
// The node counts into raw as FP64, the exact path's input, for a placed step that is rebuilt
// (scur: smax + 1 ints of scratch).
hipError_t launch_place_to_raw(const int32_t* tab, int32_t nn, int64_t smax, const int32_t* cur, int32_t* scur,
                               int32_t* scan_tmp, double* raw, hipStream_t st);
// One wave per node: the placed records of each slab in pid order, folded into out; nothing
// when flags[0] or flags[1] is set.  A cold node over 512 records is sorted in (spid, sval) by a
// workgroup (tmp_pid / tmp_val: scratch at the same offsets; PlacedRec records are unpacked
// there first) and folded by one wave.
struct PlaceFoldArgs {
    const int32_t* tab;  // launch_seg_fold's table of the layout's exact step
    int32_t nn;
    const uint32_t* rec;  // the placed records as PlacedRec, or nullptr: placed in (spid, sval)
    uint32_t* spid;       // as many entries as the layout has records
    double* sval;
    uint32_t* tmp_pid;
    double* tmp_val;
    const int32_t* base;  // the step's segment starts
    const int32_t* cur;   // the slab counts
    const int32_t* flags;
    const double* nvol;
    double spwt;
    double* out;
    int pid_bits;         // bits of the largest pid
};
hipError_t launch_place_fold(const PlaceFoldArgs& a, hipStream_t st);
// One wave's serial fold of each listed node's range [start[n], end[n]) of vals, normalised
// into out (at most max_count nodes listed; *count read on the device).
hipError_t launch_fold_list(const double* vals, const int32_t* start, const int32_t* end, const double* nvol,
                            double spwt, double* out, const int32_t* list, const int32_t* count, int32_t max_count,
                            hipStream_t st);
// The serial fold of every node's range [start[n], end[n]) of vals (start[n] < 0: none), in
// order, normalised into out; ranges over 1024 are listed (long_list, long_count: zeroed by
// the caller) and folded one wave per node.
hipError_t launch_fold_nodes(const double* vals, const int32_t* start, const int32_t* end, const double* nvol,
                             int32_t nn, double spwt, const double* run_spwt, int32_t nn_mesh, double* out,
                             int32_t* long_list, int32_t* long_count, hipStream_t st);
int64_t scan_scratch_ints(int64_t n);
hipError_t launch_excl_scan(const int32_t* in, int64_t n, int32_t* out, int32_t* scratch, hipStream_t st);
// stable compaction / pid reordering of the resident particles
hipError_t launch_flag_soa(const int32_t* cell, int64_t n, int32_t* flag, hipStream_t st);
// *count = number of cell[i] >= 0, i < n (count is cleared first)
hipError_t launch_count_live(const int32_t* cell, int64_t n, unsigned long long* count, hipStream_t st);
hipError_t launch_move_soa(const Soa& src, int64_t n, const int32_t* off, bool by_pid, const Soa& dst,
                           hipStream_t st);
hipError_t launch_mark_pid(const int32_t* cell, const int32_t* pid, int64_t n, int32_t* present, hipStream_t st);
// Counting sort by key = run * nb + (pos(cell) >> shift) (run term only when the Soa has runs),
// pos(c) = trank[c] when a tet order is given, else c; with db > 0 the key is refined to
// (key << db) | (top db bits of the pid digit (pid * mul) >> 32), nb << db keys in all.
hipError_t launch_cell_hist(const int32_t* cell, const uint16_t* run, int64_t n, int shift, int32_t nb,
                            int32_t* cnt, hipStream_t st, const int32_t* trank = nullptr,
                            const int32_t* pid = nullptr, unsigned long long mul = 0, int db = 0);
hipError_t launch_cell_scatter(const Soa& src, int64_t n, int shift, int32_t nb, int32_t* cursor,
                               const Soa& dst, hipStream_t st, const int32_t* trank = nullptr,
                               unsigned long long mul = 0, int db = 0);
// The rank pass alone: perm[d] = source slot of the particle that belongs at slot d (cursor =
// bucket offsets, advanced); lost particles get no slot.  n < 2^31.
hipError_t launch_cell_rank(const int32_t* cell, const uint16_t* run, int64_t n, int shift, int32_t nb,
                            int32_t* cursor, int32_t* perm, hipStream_t st, const int32_t* trank = nullptr);
hipError_t launch_flag_aos(const mp_particle* p, int64_t n, int32_t* flag, hipStream_t st);
hipError_t launch_check_cells_aos(const mp_particle* p, int64_t n, int32_t n_tets, int32_t* first_bad,
                                  hipStream_t st);
hipError_t launch_compact_aos(const mp_particle* in, int64_t n, const int32_t* off, mp_particle* out,
                              hipStream_t st);
hipError_t launch_set_field(TetBlock* tets, const double* ef, int32_t nt, hipStream_t st);
hipError_t launch_unpack(const double* pos, const double* vel, const int32_t* cell, int64_t n, int64_t pid0,
                         const Soa& s, int64_t at, hipStream_t st);
hipError_t launch_pack(const Soa& s, int64_t n, double* pos, double* vel, hipStream_t st);

// Stable radix sort of (key, value) pairs on key bits [0, end_bit) (mp_sort.hip).
size_t radix_sort_temp_bytes(int64_t n, int end_bit);
hipError_t radix_sort_pairs(void* temp, size_t temp_bytes, const uint32_t* keys_in, uint32_t* keys_out,
                            const double* vals_in, double* vals_out, int64_t n, int end_bit, hipStream_t st);

}  // namespace mpd
