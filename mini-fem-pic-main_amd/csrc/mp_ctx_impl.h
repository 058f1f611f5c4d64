// Private: the context behind mp_ctx* and the helpers shared by mp_ctx.hip and
// mp_batch.hip.  Not part of the C ABI.
#pragma once
#include <string>
#include <vector>

#include "mp_abi.h"
#include "mp_launch.h"
#include "mp_staging.h"

using mpd::TetBlock;
using mpd::Soa;

// A pinned host buffer that a small upload is copied through, so the call can return before
// the device has read it; `done` is recorded after the copy out of it, and the buffer is
// written again only once that event has completed.
struct PinStage {
    void* host = nullptr;
    size_t bytes = 0;
    hipEvent_t done = nullptr;
    bool pending = false;
};

struct mp_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int32_t n_tets = 0, n_nodes = 0;
    TetBlock* tets = nullptr;  // tet_blocks(n_tets) blocks
    double* node_vol = nullptr;
    double* den_raw = nullptr;
    double* den_out = nullptr;
    unsigned long long* ctr = nullptr;
    bool have_field = false;
    // resident particles, double-buffered for the cell sort
    void* soa_mem[2] = {nullptr, nullptr};
    Soa soa[2] = {};
    int cur = 0;
    int64_t cap = 0;
    int64_t n_slots = 0;
    int64_t next_pid = 0;
    int64_t since_sort = 0;
    bool need_sort = false;  // a large injection is sorted in before the next step
    bool pid_ordered = true;    // slots hold the particles in ascending pid (no cell sort since the set)
    int32_t* cmp_buf = nullptr;  // flags and offsets of the stable compaction / pid reordering
    int64_t cmp_n = 0;
    int32_t* perm = nullptr;  // gather sort: source slot of each sorted slot
    int64_t perm_n = 0;
    double* ef_stage = nullptr;  // the field's H2D staging (3 * n_tets), kept between uploads
    int32_t* tet_rank = nullptr;  // Morton rank of each tet's centroid (sort_order 1)
    unsigned long long* live_cnt = nullptr;  // mp_num_active's device count
    // counting-sort and scan scratch
    int32_t* cell_cnt = nullptr;
    int32_t* cell_cur = nullptr;
    int64_t cell_cap = 0;  // entries of cell_cnt and cell_cur
    int32_t* scan_tmp = nullptr;
    int64_t scan_tmp_n = 0;
    // deterministic-scatter scratch
    uint32_t* keys[2] = {nullptr, nullptr};
    double* vals[2] = {nullptr, nullptr};
    int64_t rec_cap = 0;
    void* sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    int32_t* seg_bounds = nullptr;
    int32_t* seg_tab = nullptr;  // node-segment scatter: counts, slab tables, node lists
    int64_t seg_tab_n = 0;
    // records placed by the push (det_algo 2), in the slab layout of the last exact step: flags
    // (8), cursors, capacities, limits and two segment layouts (the step's and the next step's),
    // place_smax + 1 ints each
    int32_t* place_mem = nullptr;
    int64_t place_mem_n = 0;    // ints allocated
    int64_t place_m = 0;        // records of the exact step the slab layout comes from
    int64_t place_smax = 0;     // seg_slab_max(nn, place_m)
    int place_buf = 0;          // which layout the next placing push uses
    bool place_ready = false;   // that layout holds the last step's counts
    bool place_pending = false; // its hot count and total are still on their way (place_ev)
    bool place_hot = false;     // the slab layout has hot nodes (records keyed by pid slab)
    int64_t place_total = 0;    // its size in records
    uint32_t* place_rec = nullptr;  // the placed records (PlacedRec: 3 words each)
    int64_t place_rec_n = 0;        // records it holds
    hipEvent_t place_ev = nullptr;
    int64_t det_placed = 0, det_rebuilt = 0;  // mp_get_det_paths
    // drop-in AoS scratch
    mp_particle* aos = nullptr;
    mp_particle* aos_out = nullptr;
    int32_t* flags = nullptr;
    int32_t* offs = nullptr;
    int64_t aos_cap = 0;
    int32_t* pinned_i32 = nullptr;
    mpstage::Stager stage;  // pinned chunks + copy threads for large host arrays
    // staging of injected particles (mp_add_particles)
    void* inj_tmp = nullptr;
    size_t inj_bytes = 0;
    // pinned host buffers of the small uploads (field, injected particles)
    PinStage pin_ef, pin_inj;
    // per-step densities of the last mp_step_many, [steps][n_nodes]
    double* den_steps = nullptr;
    int64_t den_steps_n = 0;
    int32_t den_steps_k = 0;
    // per-block tallies of the push kernels
    int32_t* tally = nullptr;
    int64_t tally_n = 0;
    // push-kernel timing (events recorded around each push launch while enabled)
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    mp_options opt{};
    std::string err;
};

namespace mpctx {

int set_err(mp_ctx* c, int code, const std::string& msg);
int bit_width(uint64_t v);
void dfree(void* p);
int reserve_scan(mp_ctx* c, int64_t n);
// Deterministic-scatter records for m record slots (keys[0..1], vals[0..1], sort scratch and
// segment table), and the sort + ordered fold of m records into den_out (normalised by spwt).
int reserve_records(mp_ctx* c, int64_t m, int end_bit, bool need_sort = true);
int det_finish(mp_ctx* c, int64_t m, double spwt);
hipError_t alloc_soa(void** mem, Soa* s, int64_t n);
// Uploads up to kPinStageMax bytes go through a PinStage: pin_acquire waits for the buffer's
// previous copy and grows it to `bytes`; pin_release records the copy just enqueued.
constexpr size_t kPinStageMax = size_t(8) << 20;
hipError_t pin_acquire(PinStage& ps, size_t bytes);
hipError_t pin_release(PinStage& ps, hipStream_t st);
void pin_free(PinStage& ps);

template <typename T>
hipError_t dalloc(T** p, int64_t n) {
    *p = nullptr;
    if (n <= 0) n = 1;
    return hipMalloc((void**)p, sizeof(T) * (size_t)n);
}

}  // namespace mpctx

namespace mpna {
// Takes the no-alloc step's scratch address on `device` (current), once per device
// (mp_noalloc.hip); mp_create calls it.
hipError_t scratch_setup(int device);
}  // namespace mpna

#define HIP_TRY(ctx, expr)                                                                  \
    do {                                                                                    \
        hipError_t e__ = (expr);                                                            \
        if (e__ != hipSuccess)                                                              \
            return mpctx::set_err((ctx), MP_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e__)); \
    } while (0)
