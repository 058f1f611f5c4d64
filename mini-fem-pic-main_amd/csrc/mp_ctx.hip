// C ABI of the particle mover (declared in include/mp_abi.h): context, device buffers,
// per-step orchestration of the kernels in mp_kernels.hip.
//
// Reference behaviour mirrored:
//   DeviceData (src/mini-pic.cpp:39-53, 322-355): one device context owning the Tetra table,
//     field, density and particle buffers; particle capacity grows to 2n on overflow.
//   MoveParticles (src/mini-pic.cpp:268-320): mp_step_aos does copy-in, kernel, copy-out,
//     stable erase of cell_index < 0 and den *= spwt / node volume, blocking like q.wait().
// The device-resident path (mp_step) keeps particles in HBM as structure of arrays with a
// particle id (pid) per particle; pid order is the order a serial MoveParticles would see.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "mp_ctx_impl.h"

namespace {
thread_local std::string g_create_err;
}  // namespace

namespace mpctx {

int set_err(mp_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    else g_create_err = msg;
    return code;
}

int bit_width(uint64_t v) {
    int b = 0;
    while (v) {
        b++;
        v >>= 1;
    }
    return b;
}

void dfree(void* p) {
    if (p) (void)hipFree(p);
}

}  // namespace mpctx

namespace mpctx {

// Lays out one SoA buffer of capacity n in a single allocation.
hipError_t alloc_soa(void** mem, Soa* s, int64_t n) {
    const size_t a = 256;
    auto up = [&](size_t b) { return (b + a - 1) / a * a; };
    const size_t d = up(sizeof(double) * (size_t)n), i = up(sizeof(int32_t) * (size_t)n);
    hipError_t e = hipMalloc(mem, 6 * d + 2 * i);
    if (e != hipSuccess) {
        *mem = nullptr;
        return e;
    }
    char* b = (char*)*mem;
    s->x = (double*)(b + 0 * d);
    s->y = (double*)(b + 1 * d);
    s->z = (double*)(b + 2 * d);
    s->vx = (double*)(b + 3 * d);
    s->vy = (double*)(b + 4 * d);
    s->vz = (double*)(b + 5 * d);
    s->cell = (int32_t*)(b + 6 * d);
    s->pid = (int32_t*)(b + 6 * d + i);
    s->run = nullptr;
    return hipSuccess;
}

int reserve_scan(mp_ctx* c, int64_t n) {
    const int64_t need = mpd::scan_scratch_ints(n);
    if (need <= c->scan_tmp_n) return MP_OK;
    dfree(c->scan_tmp);
    c->scan_tmp = nullptr;
    c->scan_tmp_n = 0;
    if (dalloc(&c->scan_tmp, need) != hipSuccess) return set_err(c, MP_ERR_ALLOC, "scan scratch allocation failed");
    c->scan_tmp_n = need;
    return MP_OK;
}

}  // namespace mpctx

namespace {
using namespace mpctx;

// Records an event on the context stream from the pool (grows on demand).
hipError_t record_event(mp_ctx* c) {
    if (c->ev_used == c->ev_pool.size()) {
        hipEvent_t e;
        hipError_t r = hipEventCreate(&e);
        if (r != hipSuccess) return r;
        c->ev_pool.push_back(e);
    }
    return hipEventRecord(c->ev_pool[c->ev_used++], c->stream);
}


mp_options default_options() {
    mp_options o;
    o.sort_interval = 10;
    o.wall_mode = MP_WALL_ABSORB;
    o.block_size = 256;
    o.particles_per_thread = 1;
    o.walk_cap = 4096;
    o.scatter_agg = 2;
    o.sort_shift = 0;
    o.keep_pid_order = 0;
    o.gather_sort = 0;
    o.sort_order = 0;
    o.det_algo = 2;
    o.sort_pid_bits = 0;
    return o;
}

int reserve_soa(mp_ctx* c, int64_t n, bool keep) {
    if (n <= c->cap) return MP_OK;
    const int64_t nc = std::max<int64_t>(2 * n, 1024);
    void* mem[2] = {nullptr, nullptr};
    Soa s[2];
    for (int k = 0; k < 2; k++) {
        if (alloc_soa(&mem[k], &s[k], nc) != hipSuccess) {
            dfree(mem[0]);
            (void)hipGetLastError();
            return set_err(c, MP_ERR_ALLOC, "Failed to allocate " + std::to_string(nc) + " particles on the device");
        }
    }
    if (keep && c->n_slots > 0) {
        const Soa& o = c->soa[c->cur];
        const size_t nd = sizeof(double) * (size_t)c->n_slots, ni = sizeof(int32_t) * (size_t)c->n_slots;
        HIP_TRY(c, hipMemcpyAsync(s[0].x, o.x, nd, hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(c, hipMemcpyAsync(s[0].y, o.y, nd, hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(c, hipMemcpyAsync(s[0].z, o.z, nd, hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(c, hipMemcpyAsync(s[0].vx, o.vx, nd, hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(c, hipMemcpyAsync(s[0].vy, o.vy, nd, hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(c, hipMemcpyAsync(s[0].vz, o.vz, nd, hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(c, hipMemcpyAsync(s[0].cell, o.cell, ni, hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(c, hipMemcpyAsync(s[0].pid, o.pid, ni, hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    dfree(c->soa_mem[0]);
    dfree(c->soa_mem[1]);
    c->soa_mem[0] = mem[0];
    c->soa_mem[1] = mem[1];
    c->soa[0] = s[0];
    c->soa[1] = s[1];
    c->cur = 0;
    c->cap = nc;
    return MP_OK;
}

int reserve_tally(mp_ctx* c, int64_t n, int block) {
    const int64_t need = mpd::tally_ints(std::max<int64_t>(n, 1), block);
    if (need <= c->tally_n) return MP_OK;
    dfree(c->tally);
    c->tally = nullptr;
    c->tally_n = 0;
    const int64_t nc = need + need / 2;
    if (dalloc(&c->tally, nc) != hipSuccess) return set_err(c, MP_ERR_ALLOC, "tally buffer allocation failed");
    c->tally_n = nc;
    return MP_OK;
}

}  // namespace

namespace mpctx {

int reserve_records(mp_ctx* c, int64_t m, int end_bit, bool need_sort) {
    if (m > c->rec_cap) {
        for (int k = 0; k < 2; k++) {
            dfree(c->keys[k]);
            dfree(c->vals[k]);
            c->keys[k] = nullptr;
            c->vals[k] = nullptr;
        }
        c->rec_cap = 0;
        // doubling: the pid-indexed record count grows with every injection, and each
        // reallocation's free synchronises the device
        const int64_t nc = std::max<int64_t>(2 * m, 4096);
        for (int k = 0; k < 2; k++)
            if (dalloc(&c->keys[k], nc) != hipSuccess || dalloc(&c->vals[k], nc) != hipSuccess)
                return set_err(c, MP_ERR_ALLOC, "deterministic scatter buffers: allocation failed");
        c->rec_cap = nc;
    }
    const size_t need = need_sort ? mpd::radix_sort_temp_bytes(c->rec_cap, end_bit) : 0;
    if (need > c->sort_tmp_bytes) {
        dfree(c->sort_tmp);
        c->sort_tmp = nullptr;
        c->sort_tmp_bytes = 0;
        if (hipMalloc(&c->sort_tmp, need) != hipSuccess)
            return set_err(c, MP_ERR_ALLOC, "radix sort scratch allocation failed");
        c->sort_tmp_bytes = need;
    }
    if (!c->seg_bounds && dalloc(&c->seg_bounds, 3 * (int64_t)c->n_nodes + 1) != hipSuccess)
        return set_err(c, MP_ERR_ALLOC, "segment table allocation failed");
    return MP_OK;
}

hipError_t pin_acquire(PinStage& ps, size_t bytes) {
    hipError_t e = hipSuccess;
    if (ps.pending) {
        e = hipEventSynchronize(ps.done);
        if (e != hipSuccess) return e;
        ps.pending = false;
    }
    if (!ps.done && (e = hipEventCreateWithFlags(&ps.done, hipEventDisableTiming)) != hipSuccess) return e;
    if (bytes > ps.bytes) {
        if (ps.host) (void)hipHostFree(ps.host);
        ps.host = nullptr;
        ps.bytes = 0;
        const size_t nb = std::max(bytes + bytes / 2, size_t(64) << 10);
        if ((e = hipHostMalloc(&ps.host, nb, 0)) != hipSuccess) return e;
        ps.bytes = nb;
    }
    return hipSuccess;
}

hipError_t pin_release(PinStage& ps, hipStream_t st) {
    const hipError_t e = hipEventRecord(ps.done, st);
    ps.pending = e == hipSuccess;
    return e;
}

void pin_free(PinStage& ps) {
    if (ps.host) (void)hipHostFree(ps.host);
    if (ps.done) (void)hipEventDestroy(ps.done);
    ps = PinStage{};
}

// The pid digit multiplier of the slab cut, floor(2^48 / P) for the pids issued so far (below P).
unsigned long long pid_digit_mul(const mp_ctx* c) {
    return (1ull << 48) / (unsigned long long)std::max<int64_t>(c->next_pid, 1);
}

// Deterministic scatter by node segments: the m slot-ordered records of a det = 2 push (and
// the per-node counts it left in den_raw) grouped by node and folded in pid order.
int det_finish_seg(mp_ctx* c, int64_t m, double spwt) {
    const int64_t need = mpd::seg_table_ints(c->n_nodes, m);
    if (need > c->seg_tab_n) {
        dfree(c->seg_tab);
        c->seg_tab = nullptr;
        c->seg_tab_n = 0;
        // headroom: the record count grows with injections
        const int64_t nc = need + need / 4;
        if (dalloc(&c->seg_tab, nc) != hipSuccess) return set_err(c, MP_ERR_ALLOC, "segment table allocation failed");
        c->seg_tab_n = nc;
    }
    int rc = reserve_scan(c, mpd::seg_scan_len(c->n_nodes, m));
    if (rc) return rc;
    mpd::SegFoldArgs a;
    a.tets = c->tets;
    a.soa = c->soa[c->cur];
    a.n = c->n_slots;
    a.tmp_pid = c->keys[0];
    a.tmp_val = c->vals[0];
    a.raw = c->den_raw;
    a.nn = c->n_nodes;
    a.tab = c->seg_tab;
    a.scan_tmp = c->scan_tmp;
    a.seg_pid = c->keys[1];
    a.seg_val = c->vals[1];
    a.nvol = c->node_vol;
    a.spwt = spwt;
    a.out = c->den_out;
    const uint64_t p = (uint64_t)std::max<int64_t>(c->next_pid, 1);
    a.pid_bits = bit_width(p - 1 > 0 ? p - 1 : 1);
    a.mul = pid_digit_mul(c);
    HIP_TRY(c, mpd::launch_seg_fold(a, c->stream));
    return MP_OK;
}

// Records placed by the push (det_algo 2), in the slab layout of the last exact step.
// place_mem: flags (8), then cursors, capacities, limits and the two segment layouts, smax + 1
// ints each (smax = seg_slab_max of that step).
namespace {
struct PlaceBufs {
    int32_t* flags;
    int32_t* cur;
    int32_t* cap;
    int32_t* lim;
    int32_t* base[2];
};
PlaceBufs place_bufs(mp_ctx* c) {
    const int64_t l = c->place_smax + 1;
    PlaceBufs p;
    p.flags = c->place_mem;
    p.cur = p.flags + 8;
    p.cap = p.cur + l;
    p.lim = p.cap + l;
    p.base[0] = p.lim + l;
    p.base[1] = p.base[0] + l;
    return p;
}
// pinned_i32 slots of the placed path's flags: 8 no fit, 9 a slab over its limit, 10 the next
// layout's size, 11 the exact path's hot count
constexpr int kPinPlace = 8;
}  // namespace

int reserve_place(mp_ctx* c, int64_t smax) {
    const int64_t need = 8 + 5 * (smax + 1);
    if (need > c->place_mem_n) {
        dfree(c->place_mem);
        c->place_mem = nullptr;
        c->place_mem_n = 0;
        const int64_t nc = need + need / 4;  // headroom: the slab count grows with the records
        if (dalloc(&c->place_mem, nc) != hipSuccess)
            return set_err(c, MP_ERR_ALLOC, "placed-record tables: allocation failed");
        c->place_mem_n = nc;
    }
    if (!c->place_ev && hipEventCreateWithFlags(&c->place_ev, hipEventDisableTiming) != hipSuccess)
        return set_err(c, MP_ERR_HIP, "placed-record event creation failed");
    return reserve_scan(c, smax + 1);
}

// The placed-record array for a layout of `total` records (a quarter of headroom: the layouts
// follow the counts).
int reserve_place_rec(mp_ctx* c, int64_t total) {
    if (total <= c->place_rec_n) return MP_OK;
    dfree(c->place_rec);
    c->place_rec = nullptr;
    c->place_rec_n = 0;
    const int64_t nc = std::max<int64_t>(total + total / 4, 4096);
    if (dalloc(&c->place_rec, mpd::kPlacedWords * nc) != hipSuccess)
        return set_err(c, MP_ERR_ALLOC, "placed-record array allocation failed");
    c->place_rec_n = nc;
    return MP_OK;
}

// Whether the next step can place its records: a layout from the last step's counts exists and
// fits the record buffers.
bool place_usable(mp_ctx* c) {
    if (c->place_pending) {
        if (hipEventSynchronize(c->place_ev) != hipSuccess) return false;
        c->place_pending = false;
        c->place_total = c->pinned_i32[kPinPlace + 2];
        c->place_hot = c->pinned_i32[kPinPlace + 3] != 0;
    }
    return c->place_ready && c->place_total <= c->rec_cap;
}

// After an exact (det_algo 1 path) step of m records with det_algo 2: its slab layout kept for the
// next steps, and the next step's segments from its slab counts.
int place_prepare(mp_ctx* c, int64_t m) {
    const int32_t nn = c->n_nodes;
    const int64_t smax = mpd::seg_slab_max(nn, m);
    int rc = reserve_place(c, smax);
    if (rc) return rc;
    c->place_m = m;
    c->place_smax = smax;
    const PlaceBufs p = place_bufs(c);
    HIP_TRY(c, mpd::launch_place_layout(c->seg_tab, nn, m, p.lim, p.cap, p.base[c->place_buf], c->scan_tmp,
                                        c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->pinned_i32 + kPinPlace + 2, p.base[c->place_buf] + smax, sizeof(int32_t),
                              hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->pinned_i32 + kPinPlace + 3, mpd::seg_hot_count(c->seg_tab, nn), sizeof(int32_t),
                              hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipEventRecord(c->place_ev, c->stream));
    c->place_pending = true;
    c->place_ready = true;
    return MP_OK;
}

// After a placing push: the slab counts checked and the next layout scanned; the placed records
// folded (on the device, while the host waits for the check); a step whose records did not all
// fit, or with a hot node's slab over its limit, is rebuilt exactly, and its counts give the next
// layout.
int place_finish(mp_ctx* c, double spwt) {
    const PlaceBufs p = place_bufs(c);
    const int32_t nn = c->n_nodes;
    const int64_t smax = c->place_smax;
    int32_t* bnow = p.base[c->place_buf];
    int32_t* bnext = p.base[1 - c->place_buf];
    HIP_TRY(c, mpd::launch_place_check(c->seg_tab, nn, smax, p.cur, p.lim, p.flags, p.cap, c->stream));
    HIP_TRY(c, mpd::launch_excl_scan(p.cap, smax + 1, bnext, c->scan_tmp, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->pinned_i32 + kPinPlace, p.flags, 2 * sizeof(int32_t), hipMemcpyDeviceToHost,
                              c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->pinned_i32 + kPinPlace + 2, bnext + smax, sizeof(int32_t), hipMemcpyDeviceToHost,
                              c->stream));
    HIP_TRY(c, hipEventRecord(c->place_ev, c->stream));
    mpd::PlaceFoldArgs f;
    f.tab = c->seg_tab;
    f.nn = nn;
    f.rec = c->place_hot ? c->place_rec : nullptr;  // the records' layout: see mp_step
    f.spid = c->keys[1];
    f.sval = c->vals[1];
    f.tmp_pid = c->keys[0];
    f.tmp_val = c->vals[0];
    f.base = bnow;
    f.cur = p.cur;
    f.flags = p.flags;
    f.nvol = c->node_vol;
    f.spwt = spwt;
    f.out = c->den_out;
    const uint64_t np = (uint64_t)std::max<int64_t>(c->next_pid, 1);
    f.pid_bits = bit_width(np - 1 > 0 ? np - 1 : 1);
    HIP_TRY(c, mpd::launch_place_fold(f, c->stream));
    HIP_TRY(c, hipEventSynchronize(c->place_ev));
    const bool no_fit = c->pinned_i32[kPinPlace] != 0, over = c->pinned_i32[kPinPlace + 1] != 0;
    c->place_pending = false;
    if (no_fit || over) {
        c->det_rebuilt++;
        // the node counts from the slab counts (the capacities are scratch now), then the exact
        // path, whose counts give the next layout
        HIP_TRY(c, mpd::launch_place_to_raw(c->seg_tab, nn, smax, p.cur, p.cap, c->scan_tmp, c->den_raw, c->stream));
        const int64_t m = 4 * c->n_slots;
        int rc = det_finish_seg(c, m, spwt);
        return rc ? rc : place_prepare(c, m);
    }
    c->place_buf = 1 - c->place_buf;
    c->place_total = c->pinned_i32[kPinPlace + 2];
    c->det_placed++;
    return MP_OK;
}

// Deterministic scatter tail: stable sort of the m pid-ordered records on the node bits,
// then the serial fold per node.
int det_finish(mp_ctx* c, int64_t m, double spwt) {
    const int end_bit = bit_width((uint64_t)c->n_nodes);
    HIP_TRY(c, mpd::radix_sort_pairs(c->sort_tmp, c->sort_tmp_bytes, c->keys[0], c->keys[1], c->vals[0],
                                     c->vals[1], m, end_bit, c->stream));
    HIP_TRY(c, mpd::launch_det_fold(c->keys[1], c->vals[1], m, c->seg_bounds, c->node_vol, c->n_nodes, spwt,
                                    c->den_out, c->stream));
    return MP_OK;
}

}  // namespace mpctx

namespace {
using namespace mpctx;

// Counting sort of the resident particles by bucket = cell >> sort_shift, dropping the lost
// ones; order inside a bucket is whatever the atomics hand out (pid is carried along).
// Stable compaction of the resident particles, dropping the lost ones and keeping pid order
// (pid order is first restored by a scatter on pid ranks if a cell sort broke it).
int compact_now(mp_ctx* c) {
    c->since_sort = 0;
    c->need_sort = false;
    if (c->n_slots == 0) {
        c->pid_ordered = true;
        return MP_OK;
    }
    const bool by_pid = !c->pid_ordered;
    const int64_t m = by_pid ? c->next_pid : c->n_slots;  // flag domain: pids or slots
    if (2 * (m + 1) > c->cmp_n) {
        dfree(c->cmp_buf);
        c->cmp_buf = nullptr;
        c->cmp_n = 0;
        if (dalloc(&c->cmp_buf, 2 * (m + 1) + (m + 1) / 2) != hipSuccess)
            return set_err(c, MP_ERR_ALLOC, "compaction scratch allocation failed");
        c->cmp_n = 2 * (m + 1) + (m + 1) / 2;
    }
    int rc = reserve_scan(c, m + 1);
    if (rc) return rc;
    int32_t* flag = c->cmp_buf;
    int32_t* off = c->cmp_buf + (m + 1);
    const Soa& src = c->soa[c->cur];
    const Soa& dst = c->soa[1 - c->cur];
    if (by_pid) {
        HIP_TRY(c, mpd::launch_fill_i32(flag, m + 1, 0, c->stream));
        HIP_TRY(c, mpd::launch_mark_pid(src.cell, src.pid, c->n_slots, flag, c->stream));
    } else {
        HIP_TRY(c, mpd::launch_flag_soa(src.cell, c->n_slots, flag, c->stream));
        HIP_TRY(c, mpd::launch_fill_i32(flag + m, 1, 0, c->stream));
    }
    HIP_TRY(c, mpd::launch_excl_scan(flag, m + 1, off, c->scan_tmp, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->pinned_i32, off + m, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, mpd::launch_move_soa(src, c->n_slots, off, by_pid, dst, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->cur = 1 - c->cur;
    c->n_slots = *c->pinned_i32;
    c->pid_ordered = true;
    return MP_OK;
}

// The periodic sort is due (every sort_interval steps, or before the step after a large
// injection; never with sort_interval 0).
bool sort_due(const mp_ctx* c) {
    const int si = c->opt.sort_interval;
    return si > 0 && (c->need_sort || c->since_sort >= si);
}

// The tet order of the cell sort's key (sort_order 1), or none (cell order).
const int32_t* sort_rank(const mp_ctx* c) { return c->opt.sort_order == 1 ? c->tet_rank : nullptr; }

int sort_now(mp_ctx* c) {
    if (c->opt.keep_pid_order) return compact_now(c);
    c->since_sort = 0;
    c->need_sort = false;
    c->pid_ordered = false;
    if (c->n_slots == 0) return MP_OK;
    const int shift = c->opt.sort_shift;
    const int64_t nb = (((int64_t)c->n_tets - 1) >> shift) + 1;
    // with pid bits every bucket splits into 2^db pid ranges
    const int db = c->opt.sort_pid_bits;
    const int64_t nbk = nb << db;
    const int64_t nb1 = nbk + 1;
    int rc = reserve_scan(c, nb1);
    if (rc) return rc;
    if (nb1 > c->cell_cap) {
        dfree(c->cell_cnt);
        dfree(c->cell_cur);
        c->cell_cnt = c->cell_cur = nullptr;
        c->cell_cap = 0;
        if (dalloc(&c->cell_cnt, nb1) != hipSuccess || dalloc(&c->cell_cur, nb1) != hipSuccess)
            return set_err(c, MP_ERR_ALLOC, "cell sort count allocation failed");
        c->cell_cap = nb1;
    }
    const unsigned long long mul = pid_digit_mul(c);
    HIP_TRY(c, mpd::launch_fill_i32(c->cell_cnt, nb1, 0, c->stream));
    const Soa& src = c->soa[c->cur];
    const Soa& dst = c->soa[1 - c->cur];
    HIP_TRY(c, mpd::launch_cell_hist(src.cell, nullptr, c->n_slots, shift, (int32_t)nb, c->cell_cnt, c->stream,
                                     sort_rank(c), src.pid, mul, db));
    HIP_TRY(c, mpd::launch_excl_scan(c->cell_cnt, nb1, c->cell_cur, c->scan_tmp, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->pinned_i32, c->cell_cur + nbk, sizeof(int32_t), hipMemcpyDeviceToHost,
                              c->stream));
    HIP_TRY(c, mpd::launch_cell_scatter(src, c->n_slots, shift, (int32_t)nb, c->cell_cur, dst, c->stream,
                                        sort_rank(c), mul, db));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->cur = 1 - c->cur;
    c->n_slots = *c->pinned_i32;
    return MP_OK;
}

// The rank half of a cell sort (the records stay where they are): perm[d] = the current slot
// of the particle that belongs at slot d, lost particles dropped; *n_live = their count.  The
// caller's next launch moves the records (k_push_steps with perm) and then takes the sort's
// bookkeeping (gather_done), so a failed launch leaves the particles as they were.
int rank_now(mp_ctx* c, int64_t* n_live) {
    const int shift = c->opt.sort_shift;
    const int64_t nb = (((int64_t)c->n_tets - 1) >> shift) + 1;
    const int64_t nb1 = nb + 1;
    int rc = reserve_scan(c, nb1);
    if (rc) return rc;
    if (c->n_slots > c->perm_n) {
        dfree(c->perm);
        c->perm = nullptr;
        c->perm_n = 0;
        if (dalloc(&c->perm, c->cap) != hipSuccess)
            return set_err(c, MP_ERR_ALLOC, "sort permutation allocation failed");
        c->perm_n = c->cap;
    }
    const Soa& src = c->soa[c->cur];
    HIP_TRY(c, mpd::launch_fill_i32(c->cell_cnt, nb1, 0, c->stream));
    HIP_TRY(c, mpd::launch_cell_hist(src.cell, nullptr, c->n_slots, shift, (int32_t)nb, c->cell_cnt, c->stream,
                                     sort_rank(c)));
    HIP_TRY(c, mpd::launch_excl_scan(c->cell_cnt, nb1, c->cell_cur, c->scan_tmp, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->pinned_i32, c->cell_cur + nb, sizeof(int32_t), hipMemcpyDeviceToHost,
                              c->stream));
    HIP_TRY(c, mpd::launch_cell_rank(src.cell, nullptr, c->n_slots, shift, (int32_t)nb, c->cell_cur, c->perm,
                                     c->stream, sort_rank(c)));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    *n_live = *c->pinned_i32;
    return MP_OK;
}

void gather_done(mp_ctx* c, int64_t n_live) {
    c->cur = 1 - c->cur;
    c->n_slots = n_live;
    c->since_sort = 0;
    c->need_sort = false;
    c->pid_ordered = false;
}

bool good_options(const mp_options& o) {
    return o.sort_interval >= 0 && (o.wall_mode == MP_WALL_ABSORB || o.wall_mode == MP_WALL_REFLECT) &&
           (o.block_size == 64 || o.block_size == 128 || o.block_size == 256) &&
           (o.particles_per_thread == 1 || o.particles_per_thread == 2 || o.particles_per_thread == 4 ||
            o.particles_per_thread == 8) &&
           o.walk_cap >= 1 && (o.scatter_agg == 0 || o.scatter_agg == 2) && o.sort_shift >= 0 &&
           o.sort_shift <= 16 &&
           (o.keep_pid_order == 0 || o.keep_pid_order == 1) &&
           (o.gather_sort == 0 || o.gather_sort == 1) && (o.sort_order == 0 || o.sort_order == 1) &&
           (o.det_algo >= 0 && o.det_algo <= 2) && o.sort_pid_bits >= 0 && o.sort_pid_bits <= 12;
}

}  // namespace

extern "C" {

// Rank of every tet along a Morton (Z-order) curve through the tet centroids, 10 bits per
// axis over the centroids' bounding box, ties by tet id.  The centroid is the point where all
// four barycentrics are 1/4: lambda_i = ((alpha_i - x beta_i) + y gamma_i - z delta_i) / (6 V),
// so rows i = 1..3 give a 3 x 3 system, solved by Cramer's rule.
static std::vector<int32_t> morton_tet_rank(const mp_tetra* tets, int32_t n) {
    std::vector<double> cx((size_t)n), cy((size_t)n), cz((size_t)n);
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int32_t e = 0; e < n; e++) {
        const mp_tetra& t = tets[e];
        double a[3][3], b[3];
        for (int r = 0; r < 3; r++) {
            a[r][0] = -t.beta[r + 1];
            a[r][1] = t.gamma[r + 1];
            a[r][2] = -t.delta[r + 1];
            b[r] = 1.5 * t.volume - t.alpha[r + 1];
        }
        auto det3 = [](const double m[3][3]) {
            return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                   m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                   m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
        };
        const double d = det3(a);
        double x[3] = {0.0, 0.0, 0.0};
        if (d != 0.0) {
            for (int k = 0; k < 3; k++) {
                double m[3][3];
                for (int r = 0; r < 3; r++)
                    for (int q = 0; q < 3; q++) m[r][q] = q == k ? b[r] : a[r][q];
                x[k] = det3(m) / d;
            }
        }
        cx[e] = x[0];
        cy[e] = x[1];
        cz[e] = x[2];
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], x[k]);
            hi[k] = std::max(hi[k], x[k]);
        }
    }
    auto spread = [](uint32_t v) {
        v &= 0x3FFu;
        v = (v | (v << 16)) & 0x030000FFu;
        v = (v | (v << 8)) & 0x0300F00Fu;
        v = (v | (v << 4)) & 0x030C30C3u;
        v = (v | (v << 2)) & 0x09249249u;
        return v;
    };
    auto quant = [](double v, double l, double h) {
        const double f = h > l ? (v - l) / (h - l) : 0.0;
        return (uint32_t)std::min(1023.0, std::max(0.0, std::floor(f * 1023.0)));
    };
    std::vector<std::pair<uint32_t, int32_t>> code((size_t)n);
    for (int32_t e = 0; e < n; e++)
        code[e] = {spread(quant(cx[e], lo[0], hi[0])) | (spread(quant(cy[e], lo[1], hi[1])) << 1) |
                       (spread(quant(cz[e], lo[2], hi[2])) << 2),
                   e};
    std::sort(code.begin(), code.end());
    std::vector<int32_t> rank((size_t)n);
    for (int32_t k = 0; k < n; k++) rank[code[k].second] = k;
    return rank;
}

int mp_create(const mp_tetra* tets, int32_t n_tets, const double* node_volume, int32_t n_nodes, mp_ctx** out) {
    if (!out) return set_err(nullptr, MP_ERR_ARG, "mp_create: out is null");
    *out = nullptr;
    if (!tets || !node_volume || n_tets <= 0 || n_nodes <= 0)
        return set_err(nullptr, MP_ERR_ARG, "mp_create: empty mesh or null table");
    for (int32_t e = 0; e < n_tets; e++)
        for (int v = 0; v < 4; v++) {
            if (tets[e].con[v] < 0 || tets[e].con[v] >= n_nodes)
                return set_err(nullptr, MP_ERR_ARG, "mp_create: tet " + std::to_string(e) + " has a node id out of range");
            if (tets[e].cell_con[v] < -1 || tets[e].cell_con[v] >= n_tets)
                return set_err(nullptr, MP_ERR_ARG, "mp_create: tet " + std::to_string(e) + " has a neighbour id out of range");
        }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        (void)hipGetLastError();
        return set_err(nullptr, MP_ERR_NODEV, "mp_create: no HIP device");
    }
    int dev = 0;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return set_err(nullptr, MP_ERR_NODEV, "mp_create: cannot query the device");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return set_err(nullptr, MP_ERR_NODEV, std::string("mp_create: device ") + prop.gcnArchName +
                                                  " is not gfx950 (this library is built for MI355X only)");
    mp_ctx* c = new (std::nothrow) mp_ctx;
    if (!c) return set_err(nullptr, MP_ERR_ALLOC, "mp_create: out of host memory");
    c->device = dev;
    c->n_tets = n_tets;
    c->n_nodes = n_nodes;
    c->opt = default_options();
    // the device table (layout in mp_device.h); E starts at zero, padding tets are all zero
    const int64_t n_blk = mpd::tet_blocks(n_tets);
    std::vector<TetBlock> h((size_t)n_blk);
    std::memset((void*)h.data(), 0, sizeof(TetBlock) * h.size());
    for (int32_t e = 0; e < n_tets; e++) {
        mpd::TetChunk* q = mpd::tet_chunks(h.data(), e);
        for (int v = 0; v < 4; v++) {
            q[(2 * v) * mpd::kTetBlock].d[0] = tets[e].alpha[v];
            q[(2 * v) * mpd::kTetBlock].d[1] = tets[e].beta[v];
            q[(2 * v + 1) * mpd::kTetBlock].d[0] = tets[e].gamma[v];
            q[(2 * v + 1) * mpd::kTetBlock].d[1] = tets[e].delta[v];
            q[mpd::kChNbr * mpd::kTetBlock].i[v] = tets[e].cell_con[v];
            q[mpd::kChCon * mpd::kTetBlock].i[v] = tets[e].con[v];
        }
        q[mpd::kChVolEx * mpd::kTetBlock].d[0] = tets[e].volume;
    }
    auto bail = [&](const std::string& m) {
        mp_destroy(c);
        return set_err(nullptr, MP_ERR_ALLOC, m);
    };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return bail("mp_create: stream creation failed");
    if (dalloc(&c->tets, n_blk) != hipSuccess || dalloc(&c->node_vol, n_nodes) != hipSuccess ||
        dalloc(&c->den_raw, n_nodes) != hipSuccess || dalloc(&c->den_out, n_nodes) != hipSuccess ||
        dalloc(&c->ctr, mpd::kNumCounters) != hipSuccess || dalloc(&c->cell_cnt, (int64_t)n_tets + 1) != hipSuccess ||
        dalloc(&c->cell_cur, (int64_t)n_tets + 1) != hipSuccess)
        return bail("Failed to allocate the mesh on the device");
    c->cell_cap = (int64_t)n_tets + 1;
    if (hipHostMalloc((void**)&c->pinned_i32, 16 * sizeof(int32_t), 0) != hipSuccess)
        return bail("mp_create: pinned allocation failed");
    {
        // the fixed scratch of mp_step_aos_noalloc on this device: its address, once per device
        const hipError_t e = mpna::scratch_setup(dev);
        if (e != hipSuccess) {
            mp_destroy(c);
            return set_err(nullptr, MP_ERR_HIP, std::string("mp_create: no-alloc scratch lookup: ") +
                                                    hipGetErrorString(e));
        }
    }
    if (hipMemcpy(c->tets, h.data(), sizeof(TetBlock) * h.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->node_vol, node_volume, sizeof(double) * (size_t)n_nodes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(c->den_raw, 0, sizeof(double) * (size_t)n_nodes) != hipSuccess ||
        hipMemset(c->den_out, 0, sizeof(double) * (size_t)n_nodes) != hipSuccess ||
        hipMemset(c->ctr, 0, sizeof(unsigned long long) * mpd::kNumCounters) != hipSuccess)
        return bail("mp_create: upload failed");
    {
        const std::vector<int32_t> rank = morton_tet_rank(tets, n_tets);
        if (dalloc(&c->tet_rank, n_tets) != hipSuccess ||
            hipMemcpy(c->tet_rank, rank.data(), sizeof(int32_t) * rank.size(), hipMemcpyHostToDevice) != hipSuccess)
            return bail("mp_create: tet order upload failed");
    }
    *out = c;
    return MP_OK;
}

void mp_destroy(mp_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    void* bufs[] = {c->tets, c->node_vol, c->den_raw, c->den_out, c->ctr, c->soa_mem[0], c->soa_mem[1],
                    c->cell_cnt, c->cell_cur, c->scan_tmp, c->keys[0], c->keys[1], c->vals[0], c->vals[1],
                    c->sort_tmp, c->seg_bounds, c->aos, c->aos_out, c->flags, c->offs, c->tally,
                    c->inj_tmp, c->cmp_buf, c->den_steps, c->perm, c->ef_stage,
                    c->live_cnt, c->tet_rank, c->seg_tab, c->place_mem, c->place_rec};
    for (void* p : bufs) dfree(p);
    if (c->pinned_i32) (void)hipHostFree(c->pinned_i32);
    pin_free(c->pin_ef);
    pin_free(c->pin_inj);
    mpstage::stager_free(c->stage);
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    if (c->place_ev) (void)hipEventDestroy(c->place_ev);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* mp_last_error(const mp_ctx* c) { return c ? c->err.c_str() : g_create_err.c_str(); }

int mp_get_options(const mp_ctx* c, mp_options* o) {
    if (!c || !o) return MP_ERR_ARG;
    *o = c->opt;
    return MP_OK;
}

int mp_set_options(mp_ctx* c, const mp_options* o) {
    if (!c || !o) return MP_ERR_ARG;
    if (!good_options(*o)) return set_err(c, MP_ERR_ARG, "mp_set_options: unsupported option values");
    // the pid-refined bucket keys are int32
    if (((((int64_t)c->n_tets - 1) >> o->sort_shift) + 1) << o->sort_pid_bits >= ((int64_t)1 << 31))
        return set_err(c, MP_ERR_ARG, "mp_set_options: sort_pid_bits too large for this mesh and sort_shift");
    c->opt = *o;
    return MP_OK;
}

int mp_set_sort_interval(mp_ctx* c, int32_t k) {
    if (!c) return MP_ERR_ARG;
    if (k < 0) return set_err(c, MP_ERR_ARG, "mp_set_sort_interval: interval must be >= 0");
    c->opt.sort_interval = k;
    return MP_OK;
}

int mp_upload_field(mp_ctx* c, const double* ef, int32_t n_tets) {
    if (!c) return MP_ERR_ARG;
    if (!ef || n_tets != c->n_tets) return set_err(c, MP_ERR_ARG, "mp_upload_field: ef must be [n_tets][3]");
    HIP_TRY(c, hipSetDevice(c->device));
    // the staging buffer is kept: a free per upload would synchronise the device every step
    if (!c->ef_stage) HIP_TRY(c, dalloc(&c->ef_stage, 3 * (int64_t)n_tets));
    const size_t bytes = sizeof(double) * 3 * (size_t)n_tets;
    hipError_t e;
    if (bytes <= kPinStageMax) {
        // a small field (a coupled loop's mesh) goes through pinned memory, and the call returns
        // without waiting for the device: the step launched next queues behind the copy
        e = pin_acquire(c->pin_ef, bytes);
        if (e == hipSuccess) {
            std::memcpy(c->pin_ef.host, ef, bytes);
            e = hipMemcpyAsync(c->ef_stage, c->pin_ef.host, bytes, hipMemcpyHostToDevice, c->stream);
        }
        if (e == hipSuccess) e = pin_release(c->pin_ef, c->stream);
        if (e == hipSuccess) e = mpd::launch_set_field(c->tets, c->ef_stage, n_tets, c->stream);
    } else {
        e = hipMemcpyAsync(c->ef_stage, ef, bytes, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) e = mpd::launch_set_field(c->tets, c->ef_stage, n_tets, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    }
    if (e != hipSuccess) return set_err(c, MP_ERR_HIP, std::string("mp_upload_field: ") + hipGetErrorString(e));
    c->have_field = true;
    return MP_OK;
}

int mp_step_aos(mp_ctx* c, mp_particle* parts, size_t* n_inout, double q_over_m, double dt, double spwt,
                double* den_out, int32_t mode) {
    if (!c) return MP_ERR_ARG;
    if (!n_inout || !den_out || (*n_inout > 0 && !parts)) return set_err(c, MP_ERR_ARG, "mp_step_aos: null argument");
    if (mode != MP_SCATTER_ATOMIC && mode != MP_SCATTER_DETERMINISTIC)
        return set_err(c, MP_ERR_ARG, "mp_step_aos: unknown scatter mode");
    if (!c->have_field) return set_err(c, MP_ERR_STATE, "mp_step_aos: no field uploaded");
    const int64_t n = (int64_t)*n_inout;
    if (n > INT32_MAX / 4) return set_err(c, MP_ERR_ARG, "mp_step_aos: too many particles");
    HIP_TRY(c, hipSetDevice(c->device));
    if (n == 0) {  // nothing to move: the density is all zero
        for (int32_t i = 0; i < c->n_nodes; i++) den_out[i] = 0.0;
        return MP_OK;
    }
    if (n > c->aos_cap) {  // DeviceData::reserveParticles: grow to 2n
        dfree(c->aos); dfree(c->aos_out); dfree(c->flags); dfree(c->offs);
        c->aos = c->aos_out = nullptr;
        c->flags = c->offs = nullptr;
        c->aos_cap = 0;
        const int64_t nc = 2 * n;
        if (dalloc(&c->aos, nc) != hipSuccess || dalloc(&c->aos_out, nc) != hipSuccess ||
            dalloc(&c->flags, nc) != hipSuccess || dalloc(&c->offs, nc + 1) != hipSuccess)
            return set_err(c, MP_ERR_ALLOC, "Failed to allocate " + std::to_string(nc) + " particles on the device");
        c->aos_cap = nc;
    }
    const bool det = mode == MP_SCATTER_DETERMINISTIC;
    if (det) {
        int rc = reserve_records(c, 4 * n, bit_width((uint64_t)c->n_nodes));
        if (rc) return rc;
        HIP_TRY(c, mpd::launch_fill_u32(c->keys[0], 4 * n, 0xFFFFFFFFu, c->stream));
    }
    int rc = reserve_scan(c, n + 1);
    if (rc) return rc;
    const size_t bytes = sizeof(mp_particle) * (size_t)n;
    if (bytes >= mpstage::kStageMin) {
        HIP_TRY(c, mpstage::stager_init(c->stage));
        HIP_TRY(c, mpstage::upload(c->stage, c->aos, parts, bytes, c->stream));
    } else {
        HIP_TRY(c, hipMemcpyAsync(c->aos, parts, bytes, hipMemcpyHostToDevice, c->stream));
    }
    // every cell_index must name a tet before the walk indexes the table with it (checked on
    // the device, nothing written yet on failure)
    HIP_TRY(c, mpd::launch_fill_i32(c->flags, 1, INT32_MAX, c->stream));
    HIP_TRY(c, mpd::launch_check_cells_aos(c->aos, n, c->n_tets, c->flags, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->pinned_i32, c->flags, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (*c->pinned_i32 != INT32_MAX)
        return set_err(c, MP_ERR_ARG, "mp_step_aos: particle " + std::to_string(*c->pinned_i32) +
                                          " has cell_index out of range");
    rc = reserve_tally(c, n, 256);
    if (rc) return rc;
    HIP_TRY(c, mpd::launch_push_aos(c->tets, c->aos, n, q_over_m, dt, c->den_raw, det ? c->keys[0] : nullptr,
                                    det ? c->vals[0] : nullptr, c->tally, c->ctr, c->opt.walk_cap,
                                    det, c->stream));
    if (det) {
        rc = det_finish(c, 4 * n, spwt);
        if (rc) return rc;
    } else {
        HIP_TRY(c, mpd::launch_normalise(c->den_raw, c->den_out, c->node_vol, c->n_nodes, spwt, c->stream));
    }
    // stable erase of cell_index < 0: flags -> exclusive scan -> compaction
    HIP_TRY(c, mpd::launch_flag_aos(c->aos, n, c->flags, c->stream));
    HIP_TRY(c, mpd::launch_fill_i32(c->flags + n, 1, 0, c->stream));
    HIP_TRY(c, mpd::launch_excl_scan(c->flags, n + 1, c->offs, c->scan_tmp, c->stream));
    HIP_TRY(c, mpd::launch_compact_aos(c->aos, n, c->offs, c->aos_out, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->pinned_i32, c->offs + n, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipMemcpyAsync(den_out, c->den_out, sizeof(double) * (size_t)c->n_nodes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const int64_t kept = *c->pinned_i32;
    const size_t kept_bytes = sizeof(mp_particle) * (size_t)kept;
    if (kept_bytes >= mpstage::kStageMin) {
        HIP_TRY(c, mpstage::stager_init(c->stage));
        HIP_TRY(c, mpstage::download(c->stage, parts, c->aos_out, kept_bytes, c->stream));
    } else if (kept > 0) {
        HIP_TRY(c, hipMemcpy(parts, c->aos_out, kept_bytes, hipMemcpyDeviceToHost));
    }
    *n_inout = (size_t)kept;
    return MP_OK;
}

int mp_set_particles(mp_ctx* c, const double* pos, const double* vel, const int32_t* cell, int64_t n) {
    if (!c) return MP_ERR_ARG;
    if (n < 0 || (n > 0 && (!pos || !vel || !cell))) return set_err(c, MP_ERR_ARG, "mp_set_particles: bad arrays");
    if (n > INT32_MAX / 4) return set_err(c, MP_ERR_ARG, "mp_set_particles: too many particles");
    HIP_TRY(c, hipSetDevice(c->device));
    c->n_slots = 0;
    c->next_pid = 0;
    c->pid_ordered = true;
    c->place_ready = false;
    c->place_pending = false;
    c->det_placed = c->det_rebuilt = 0;
    HIP_TRY(c, hipMemsetAsync(c->ctr, 0, sizeof(unsigned long long) * mpd::kNumCounters, c->stream));
    return mp_add_particles(c, pos, vel, cell, n);
}

int mp_add_particles(mp_ctx* c, const double* pos, const double* vel, const int32_t* cell, int64_t n) {
    if (!c) return MP_ERR_ARG;
    if (n < 0 || (n > 0 && (!pos || !vel || !cell))) return set_err(c, MP_ERR_ARG, "mp_add_particles: bad arrays");
    if (c->next_pid + n > INT32_MAX / 4) return set_err(c, MP_ERR_ARG, "mp_add_particles: too many particles");
    for (int64_t i = 0; i < n; i++)
        if (cell[i] < -1 || cell[i] >= c->n_tets)
            return set_err(c, MP_ERR_ARG, "mp_add_particles: particle " + std::to_string(i) + " has cell out of range");
    HIP_TRY(c, hipSetDevice(c->device));
    int rc = reserve_soa(c, c->n_slots + n, true);
    if (rc) return rc;
    if (n > 0) {
        // staging buffer for the host arrays, kept for the next injection
        const size_t bp = sizeof(double) * 3 * (size_t)n, bc = sizeof(int32_t) * (size_t)n;
        if (2 * bp + bc > c->inj_bytes) {
            dfree(c->inj_tmp);
            c->inj_tmp = nullptr;
            c->inj_bytes = 0;
            if (hipMalloc(&c->inj_tmp, 2 * (2 * bp + bc)) != hipSuccess) {
                (void)hipGetLastError();
                return set_err(c, MP_ERR_ALLOC, "mp_add_particles: staging buffer allocation failed");
            }
            c->inj_bytes = 2 * (2 * bp + bc);
        }
        char* tmp = (char*)c->inj_tmp;
        const size_t all = 2 * bp + bc;
        hipError_t e;
        const bool small = all <= kPinStageMax;
        if (small) {
            // a per-step injection: one copy through pinned memory, no wait for the device
            e = pin_acquire(c->pin_inj, all);
            if (e == hipSuccess) {
                char* h = (char*)c->pin_inj.host;
                std::memcpy(h, pos, bp);
                std::memcpy(h + bp, vel, bp);
                std::memcpy(h + 2 * bp, cell, bc);
                e = hipMemcpyAsync(tmp, h, all, hipMemcpyHostToDevice, c->stream);
            }
            if (e == hipSuccess) e = pin_release(c->pin_inj, c->stream);
        } else {
            e = hipMemcpyAsync(tmp, pos, bp, hipMemcpyHostToDevice, c->stream);
            if (e == hipSuccess) e = hipMemcpyAsync(tmp + bp, vel, bp, hipMemcpyHostToDevice, c->stream);
            if (e == hipSuccess) e = hipMemcpyAsync(tmp + 2 * bp, cell, bc, hipMemcpyHostToDevice, c->stream);
        }
        if (e == hipSuccess)
            e = mpd::launch_unpack((double*)tmp, (double*)(tmp + bp), (int32_t*)(tmp + 2 * bp), n, c->next_pid,
                                   c->soa[c->cur], c->n_slots, c->stream);
        if (e == hipSuccess && !small) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return set_err(c, MP_ERR_HIP, std::string("mp_add_particles: ") + hipGetErrorString(e));
    }
    c->n_slots += n;
    c->next_pid += n;
    // a large batch (over an eighth of the set) is sorted in before the next step; small
    // injections join the next periodic sort
    if (8 * n > c->n_slots) c->need_sort = true;
    return MP_OK;
}

int mp_step(mp_ctx* c, double q_over_m, double dt, double spwt, int32_t mode) {
    if (!c) return MP_ERR_ARG;
    if (mode != MP_SCATTER_ATOMIC && mode != MP_SCATTER_DETERMINISTIC)
        return set_err(c, MP_ERR_ARG, "mp_step: unknown scatter mode");
    if (!c->have_field) return set_err(c, MP_ERR_STATE, "mp_step: no field uploaded");
    HIP_TRY(c, hipSetDevice(c->device));
    if (sort_due(c)) {
        int rc = sort_now(c);
        if (rc) return rc;
    }
    const bool det = mode == MP_SCATTER_DETERMINISTIC;
    const bool seg = det && (c->opt.det_algo == 1 || c->opt.det_algo == 2);
    // node segments: 4*slot+v record slots (rebuilt from the state, or placed by the push with
    // det_algo 2, the segments sized from the last step's counts); radix sort: 4*slot+v while
    // the slots are in pid order, else 4*pid+v for every pid issued (sentinels pre-filled)
    const int64_t n_rec = (seg || c->pid_ordered) ? 4 * c->n_slots : 4 * c->next_pid;
    if (det) {
        int rc = reserve_records(c, n_rec, bit_width((uint64_t)c->n_nodes), !seg);
        if (rc) return rc;
        if (!seg) HIP_TRY(c, mpd::launch_fill_u32(c->keys[0], n_rec, 0xFFFFFFFFu, c->stream));
    }
    const bool place = seg && c->opt.det_algo == 2 && place_usable(c);
    if (place && c->place_hot) {
        int rc = reserve_place_rec(c, c->place_total);
        if (rc) return rc;
    }
    {
        int rc = reserve_tally(c, c->n_slots, c->opt.block_size);
        if (rc) return rc;
    }
    mpd::PushArgs a;
    a.tets = c->tets;
    a.soa = c->soa[c->cur];
    a.n = c->n_slots;
    a.qm = q_over_m;
    a.dt = dt;
    a.den = c->den_raw;
    // placed records: PlacedRec records where hot nodes are cut into pid slabs (det 3), else
    // pids and weights in the grouped arrays (det 4)
    const bool place_aos = place && c->place_hot;
    a.keys = place_aos ? c->place_rec : place ? c->keys[1] : det ? c->keys[0] : nullptr;
    a.vals = place_aos ? nullptr : place ? c->vals[1] : det ? c->vals[0] : nullptr;
    a.part = c->tally;
    a.ctr = c->ctr;
    a.cap = c->opt.walk_cap;
    a.wall = c->opt.wall_mode;
    // placing: 512-thread workgroups where every node is one slab, 256 with pid slabs
    a.det = place ? (c->place_hot ? 3 : 4) : seg ? 2 : det ? 1 : 0;
    if (place) {
        const PlaceBufs p = place_bufs(c);
        a.place = mpd::PlaceArgs{p.base[c->place_buf], p.cur, p.flags,
                                 c->place_hot ? mpd::seg_slab_info(c->seg_tab, c->n_nodes) : nullptr,
                                 pid_digit_mul(c)};
        // flags and cursors (contiguous) start at zero
        HIP_TRY(c, mpd::launch_fill_i32(p.flags, 8 + c->place_smax + 1, 0, c->stream));
    }
    a.agg = c->opt.scatter_agg;
    a.block = c->opt.block_size;
    a.ppt = c->opt.particles_per_thread;
    a.rec_by_slot = c->pid_ordered;
    if (c->timing) HIP_TRY(c, record_event(c));
    HIP_TRY(c, mpd::launch_push_soa(a, c->stream));
    if (c->timing) HIP_TRY(c, record_event(c));
    if (place) {
        int rc = place_finish(c, spwt);
        if (rc) return rc;
    } else if (seg) {
        c->det_rebuilt++;
        int rc = det_finish_seg(c, n_rec, spwt);
        if (!rc && c->opt.det_algo == 2) rc = place_prepare(c, n_rec);
        if (rc) return rc;
    } else if (det) {
        int rc = det_finish(c, n_rec, spwt);
        if (rc) return rc;
    } else {
        HIP_TRY(c, mpd::launch_normalise(c->den_raw, c->den_out, c->node_vol, c->n_nodes, spwt, c->stream));
    }
    c->since_sort++;
    return MP_OK;
}

int mp_step_many(mp_ctx* c, double q_over_m, double dt, double spwt, int32_t mode, int32_t n_steps) {
    if (!c) return MP_ERR_ARG;
    if (mode != MP_SCATTER_ATOMIC && mode != MP_SCATTER_DETERMINISTIC)
        return set_err(c, MP_ERR_ARG, "mp_step_many: unknown scatter mode");
    if (n_steps < 1) return set_err(c, MP_ERR_ARG, "mp_step_many: n_steps must be >= 1");
    if (!c->have_field) return set_err(c, MP_ERR_STATE, "mp_step_many: no field uploaded");
    HIP_TRY(c, hipSetDevice(c->device));
    const int64_t nn = c->n_nodes;
    const int64_t need = nn * (int64_t)n_steps;
    if (need > c->den_steps_n) {
        dfree(c->den_steps);
        c->den_steps = nullptr;
        c->den_steps_n = 0;
        if (dalloc(&c->den_steps, need) != hipSuccess)
            return set_err(c, MP_ERR_ALLOC, "mp_step_many: density buffer allocation failed");
        c->den_steps_n = need;
    }
    c->den_steps_k = 0;
    if (mode == MP_SCATTER_DETERMINISTIC) {
        // the sorted-record reduction is per step; only the step loop is saved
        for (int32_t k = 0; k < n_steps; k++) {
            int rc = mp_step(c, q_over_m, dt, spwt, mode);
            if (rc) return rc;
            HIP_TRY(c, hipMemcpyAsync(c->den_steps + k * nn, c->den_out, sizeof(double) * (size_t)nn,
                                      hipMemcpyDeviceToDevice, c->stream));
        }
        c->den_steps_k = n_steps;
        return MP_OK;
    }
    HIP_TRY(c, mpd::launch_fill_f64(c->den_steps, need, 0.0, c->stream));
    int32_t done = 0;
    while (done < n_steps) {
        const int32_t si = c->opt.sort_interval;
        // gather sort: the rank pass now, the record moves inside this launch
        bool gather = false;
        int64_t n_live = 0;
        if (sort_due(c)) {
            gather = c->opt.gather_sort && !c->opt.keep_pid_order && c->n_slots > 0 &&
                     c->n_slots <= (int64_t)INT32_MAX;
            int rc = gather ? rank_now(c, &n_live) : sort_now(c);
            if (rc) return rc;
        }
        int32_t chunk = n_steps - done;
        if (si > 0) chunk = (int32_t)std::min<int64_t>(chunk, si - (gather ? 0 : c->since_sort));
        chunk = std::min<int32_t>(chunk, mpd::kStepMaxLaunch);
        int rc = reserve_tally(c, c->n_slots, mpd::steps_block());
        if (rc) return rc;
        mpd::PushArgs a;
        a.tets = c->tets;
        a.soa = c->soa[gather ? 1 - c->cur : c->cur];
        a.n = gather ? n_live : c->n_slots;
        if (gather) {
            a.src = c->soa[c->cur];
            a.perm = c->perm;
        }
        a.qm = q_over_m;
        a.dt = dt;
        a.den = c->den_steps + done * nn;
        a.keys = nullptr;
        a.vals = nullptr;
        a.part = c->tally;
        a.ctr = c->ctr;
        a.cap = c->opt.walk_cap;
        a.wall = c->opt.wall_mode;
        a.det = 0;
        a.agg = c->opt.scatter_agg;
        a.block = 256;
        a.ppt = 1;
        a.rec_by_slot = false;
        if (c->timing) HIP_TRY(c, record_event(c));
        HIP_TRY(c, mpd::launch_push_steps(a, c->n_nodes, chunk, c->stream));
        if (c->timing) HIP_TRY(c, record_event(c));
        if (gather) gather_done(c, n_live);
        c->since_sort += chunk;
        done += chunk;
    }
    HIP_TRY(c, mpd::launch_normalise_steps(c->den_steps, c->node_vol, c->n_nodes, n_steps, spwt, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->den_out, c->den_steps + (int64_t)(n_steps - 1) * nn, sizeof(double) * (size_t)nn,
                              hipMemcpyDeviceToDevice, c->stream));
    c->den_steps_k = n_steps;
    return MP_OK;
}

int mp_get_density_step(mp_ctx* c, int32_t k, double* den) {
    if (!c || !den) return MP_ERR_ARG;
    if (k < 0 || k >= c->den_steps_k)
        return set_err(c, MP_ERR_ARG, "mp_get_density_step: no such step in the last mp_step_many");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMemcpyAsync(den, c->den_steps + (int64_t)k * c->n_nodes, sizeof(double) * (size_t)c->n_nodes,
                              hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return MP_OK;
}

int mp_sync(mp_ctx* c) {
    if (!c) return MP_ERR_ARG;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return MP_OK;
}

int mp_sort_now(mp_ctx* c) {
    if (!c) return MP_ERR_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    return sort_now(c);
}

int mp_num_active(mp_ctx* c, int64_t* n_active) {
    if (!c || !n_active) return MP_ERR_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    // counted on the device: one 8 B read-back instead of the whole cell array
    if (!c->live_cnt) HIP_TRY(c, dalloc(&c->live_cnt, 1));
    HIP_TRY(c, mpd::launch_count_live(c->soa[c->cur].cell, c->n_slots, c->live_cnt, c->stream));
    unsigned long long* host = reinterpret_cast<unsigned long long*>(c->pinned_i32 + 2);
    HIP_TRY(c, hipMemcpyAsync(host, c->live_cnt, sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    *n_active = (int64_t)*host;
    return MP_OK;
}

int mp_get_slots(mp_ctx* c, int32_t* cell, int32_t* pid, int64_t* n_inout) {
    if (!c || !n_inout) return MP_ERR_ARG;
    const int64_t cap = *n_inout;
    *n_inout = c->n_slots;
    if ((cell || pid) && cap < c->n_slots) return set_err(c, MP_ERR_ARG, "mp_get_slots: capacity too small");
    HIP_TRY(c, hipSetDevice(c->device));
    const Soa& s = c->soa[c->cur];
    const size_t bytes = sizeof(int32_t) * (size_t)c->n_slots;
    if (cell && bytes) HIP_TRY(c, hipMemcpyAsync(cell, s.cell, bytes, hipMemcpyDeviceToHost, c->stream));
    if (pid && bytes) HIP_TRY(c, hipMemcpyAsync(pid, s.pid, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return MP_OK;
}

int mp_get_tet_order(mp_ctx* c, int32_t* rank) {
    if (!c || !rank) return MP_ERR_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMemcpyAsync(rank, c->tet_rank, sizeof(int32_t) * (size_t)c->n_tets, hipMemcpyDeviceToHost,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return MP_OK;
}

int mp_get_particles(mp_ctx* c, double* pos, double* vel, int32_t* cell, int64_t* pid, int64_t* n_inout) {
    if (!c || !n_inout) return MP_ERR_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    const int64_t n = c->n_slots;
    std::vector<double> hp(3 * (size_t)n), hv(3 * (size_t)n);
    std::vector<int32_t> hc((size_t)n), hid((size_t)n);
    if (n > 0) {
        double* tmp = nullptr;
        HIP_TRY(c, dalloc(&tmp, 6 * n));
        hipError_t e = mpd::launch_pack(c->soa[c->cur], n, tmp, tmp + 3 * n, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(hp.data(), tmp, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(hv.data(), tmp + 3 * n, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(hc.data(), c->soa[c->cur].cell, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(hid.data(), c->soa[c->cur].pid, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        dfree(tmp);
        if (e != hipSuccess) return set_err(c, MP_ERR_HIP, std::string("mp_get_particles: ") + hipGetErrorString(e));
    }
    // ascending pid: place slots by pid, then walk the pid range
    std::vector<int64_t> slot((size_t)c->next_pid, -1);
    int64_t live = 0;
    for (int64_t i = 0; i < n; i++)
        if (hc[i] >= 0) {
            slot[(size_t)hid[i]] = i;
            live++;
        }
    if (live > *n_inout) {
        *n_inout = live;
        return set_err(c, MP_ERR_ARG, "mp_get_particles: output capacity too small");
    }
    int64_t k = 0;
    for (int64_t p = 0; p < c->next_pid; p++) {
        const int64_t i = slot[(size_t)p];
        if (i < 0) continue;
        if (pos) for (int d = 0; d < 3; d++) pos[3 * k + d] = hp[3 * i + d];
        if (vel) for (int d = 0; d < 3; d++) vel[3 * k + d] = hv[3 * i + d];
        if (cell) cell[k] = hc[i];
        if (pid) pid[k] = p;
        k++;
    }
    *n_inout = k;
    return MP_OK;
}

int mp_get_density(mp_ctx* c, double* den) {
    if (!c || !den) return MP_ERR_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMemcpyAsync(den, c->den_out, sizeof(double) * (size_t)c->n_nodes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return MP_OK;
}

int mp_get_det_paths(mp_ctx* c, int64_t* placed, int64_t* rebuilt) {
    if (!c) return MP_ERR_ARG;
    if (placed) *placed = c->det_placed;
    if (rebuilt) *rebuilt = c->det_rebuilt;
    return MP_OK;
}

int mp_get_counters(mp_ctx* c, mp_counters* out) {
    if (!c || !out) return MP_ERR_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    unsigned long long h[mpd::kNumCounters];
    HIP_TRY(c, hipMemcpyAsync(h, c->ctr, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    out->particle_steps = (int64_t)h[mpd::kCtrParticleSteps];
    out->lost = (int64_t)h[mpd::kCtrLost];
    out->crossed = (int64_t)h[mpd::kCtrCrossed];
    out->moves = (int64_t)h[mpd::kCtrMoves];
    out->cap_hits = (int64_t)h[mpd::kCtrCapHits];
    out->reflections = (int64_t)h[mpd::kCtrReflections];
    return MP_OK;
}

int mp_reset_counters(mp_ctx* c) {
    if (!c) return MP_ERR_ARG;
    c->det_placed = c->det_rebuilt = 0;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMemsetAsync(c->ctr, 0, sizeof(unsigned long long) * mpd::kNumCounters, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return MP_OK;
}

void* mp_stream(mp_ctx* c) { return c ? (void*)c->stream : nullptr; }

int mp_set_timing(mp_ctx* c, int32_t enable) {
    if (!c) return MP_ERR_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->timing = enable != 0;
    c->ev_used = 0;
    return MP_OK;
}

int mp_get_timing(mp_ctx* c, double* push_ms, int64_t* push_launches) {
    if (!c || !push_ms || !push_launches) return MP_ERR_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    double total = 0.0;
    for (size_t k = 0; k + 1 < c->ev_used; k += 2) {
        float ms = 0.f;
        HIP_TRY(c, hipEventElapsedTime(&ms, c->ev_pool[k], c->ev_pool[k + 1]));
        total += ms;
    }
    *push_ms = total;
    *push_launches = (int64_t)(c->ev_used / 2);
    return MP_OK;
}

}  // extern "C"
