// mp_step_aos_noalloc: the drop-in MoveParticles step over Particle records that the caller
// already holds in device memory, for callers that cannot allocate device memory.  Every
// scratch buffer of the step (keep flags, their prefix sums and the scan's partials, the
// per-workgroup step tallies, the first-bad-cell word) lives in one file-scope device array,
// `scratch`, of fixed size; the call itself allocates nothing.  The context's own buffers
// (tet table, raw density, counters, one pinned host word) were allocated by mp_create.
//
// The array is one per GPU and is held until the process ends.  Its device address on each
// GPU is taken once, by mp_create with that GPU current (mpna::scratch_setup), and kept per
// device: host code passes that address to the kernels, never the array's name.  Calls that
// overlap on one GPU (two host threads, two contexts) share the array and race.
#include <mutex>

#include "mp_ctx_impl.h"

using namespace mpctx;

namespace {

constexpr int64_t kMaxParticles = MP_NOALLOC_MAX_PARTICLES;
constexpr int kMaxDevices = 64;
// Word layout of the scratch array (int32 words):
//   flags [kMaxParticles + 1]   1 = particle kept (the last word is the scan's 0 sentinel)
//   offs  [kMaxParticles + 1]   exclusive prefix sum of flags: destination of each survivor
//   scan  [kScanInts]           partials of the multi-level scan
//   tally [kTallyInts]          per-workgroup step counters of the push (6 per 256 particles)
//   bad   [64]                  first particle with an out-of-range cell (word 0)
constexpr int64_t kFlagInts = kMaxParticles + 1;
constexpr int64_t kScanInts = kMaxParticles / 256 + 256;
constexpr int64_t kTallyInts = 6 * ((kMaxParticles + 255) / 256 + 8);
constexpr int64_t kOffFlags = 0;
constexpr int64_t kOffOffs = ((kOffFlags + kFlagInts + 63) / 64) * 64;
constexpr int64_t kOffScan = ((kOffOffs + kFlagInts + 63) / 64) * 64;
constexpr int64_t kOffTally = ((kOffScan + kScanInts + 63) / 64) * 64;
constexpr int64_t kOffBad = ((kOffTally + kTallyInts + 63) / 64) * 64;
constexpr int64_t kScratchInts = kOffBad + 64;
static_assert(kScratchInts * 4 < (int64_t)48 << 20, "the scratch array stays under 48 MiB");

}  // namespace

// The step's scratch memory on each GPU (8,503,808 words, 32.4 MiB).  Host code uses the
// address that hipGetSymbolAddress returned for the current device (g_scratch), not this name.
static __device__ __align__(256) int32_t scratch[kScratchInts];

namespace {

std::mutex g_mu;
int32_t* g_scratch[kMaxDevices] = {};

}  // namespace

namespace mpna {

// Looks up the scratch array's address on `device`, which must be current; once per device.
hipError_t scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[device]) return hipSuccess;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch[device] = (int32_t*)p;
    return hipSuccess;
}

}  // namespace mpna

namespace {

int32_t* scratch_of(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    return g_scratch[device];
}

bool overlaps(const void* a, const void* b, size_t bytes) {
    const uintptr_t x = (uintptr_t)a, y = (uintptr_t)b;
    return x < y + bytes && y < x + bytes;
}

}  // namespace

extern "C" int mp_step_aos_noalloc(mp_ctx* c, mp_particle* parts, int64_t n, mp_particle* kept, int64_t* n_kept,
                                   double q_over_m, double dt, double spwt, double* den_out, int32_t mode) {
    if (!c) return MP_ERR_ARG;
    if (!n_kept || !den_out || (n > 0 && (!parts || !kept)))
        return set_err(c, MP_ERR_ARG, "mp_step_aos_noalloc: null argument");
    if (mode != MP_SCATTER_ATOMIC)
        return set_err(c, MP_ERR_ARG, "mp_step_aos_noalloc: atomic scatter only (the deterministic scatter needs "
                                      "per-particle record storage beyond the fixed scratch)");
    if (n < 0 || n > kMaxParticles)
        return set_err(c, MP_ERR_ARG, "mp_step_aos_noalloc: " + std::to_string(n) +
                                          " particles; the fixed scratch holds at most " +
                                          std::to_string(kMaxParticles));
    if (n > 0 && overlaps(parts, kept, sizeof(mp_particle) * (size_t)n))
        return set_err(c, MP_ERR_ARG, "mp_step_aos_noalloc: kept must not overlap parts");
    if (!c->have_field) return set_err(c, MP_ERR_STATE, "mp_step_aos_noalloc: no field uploaded");
    if (mpd::scan_scratch_ints(n + 1) > kScanInts)
        return set_err(c, MP_ERR_STATE, "mp_step_aos_noalloc: scan scratch layout too small");
    HIP_TRY(c, hipSetDevice(c->device));
    int32_t* base = scratch_of(c->device);
    if (!base) return set_err(c, MP_ERR_STATE, "mp_step_aos_noalloc: scratch not set up on this device");
    int32_t* flags = base + kOffFlags;
    int32_t* offs = base + kOffOffs;
    int32_t* scan = base + kOffScan;
    int32_t* tally = base + kOffTally;
    int32_t* bad = base + kOffBad;
    if (n == 0) {  // nothing to move: the density is all zero
        HIP_TRY(c, mpd::launch_fill_f64(den_out, c->n_nodes, 0.0, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        *n_kept = 0;
        return MP_OK;
    }
    // every cell_index must name a tet before the walk indexes the table with it; the word
    // keeps the last call's value, so it is reset first
    HIP_TRY(c, mpd::launch_fill_i32(bad, 1, INT32_MAX, c->stream));
    HIP_TRY(c, mpd::launch_check_cells_aos(parts, n, c->n_tets, bad, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->pinned_i32, bad, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (*c->pinned_i32 != INT32_MAX)
        return set_err(c, MP_ERR_ARG, "mp_step_aos_noalloc: particle " + std::to_string(*c->pinned_i32) +
                                          " has cell_index out of range");
    // push + walk + scatter in place (every workgroup writes its tally row), normalise into the
    // caller's density (the raw sums are cleared for the next step), stable erase into kept
    HIP_TRY(c, mpd::launch_push_aos(c->tets, parts, n, q_over_m, dt, c->den_raw, nullptr, nullptr, tally, c->ctr,
                                    c->opt.walk_cap, false, c->stream));
    HIP_TRY(c, mpd::launch_normalise(c->den_raw, den_out, c->node_vol, c->n_nodes, spwt, c->stream));
    HIP_TRY(c, mpd::launch_flag_aos(parts, n, flags, c->stream));
    HIP_TRY(c, mpd::launch_fill_i32(flags + n, 1, 0, c->stream));
    HIP_TRY(c, mpd::launch_excl_scan(flags, n + 1, offs, scan, c->stream));
    HIP_TRY(c, mpd::launch_compact_aos(parts, n, offs, kept, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->pinned_i32, offs + n, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    *n_kept = *c->pinned_i32;
    return MP_OK;
}
