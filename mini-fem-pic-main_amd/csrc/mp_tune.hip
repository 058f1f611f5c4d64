// Autotuner of the device-resident step (mp_autotune): picks threads per particle, particles
// per thread, workgroup size, atomic-scatter aggregation, start-tet staging, crossing-particle
// compaction, pid-order compaction, cell-sort key shift and sort interval for the particles
// currently resident, by timing short windows of mp_step from one
// saved state.
//
// Every candidate gives bit-identical cells, positions and velocities (the options change only
// the schedule and the memory order), so the tuner only measures time.  Protocol per
// candidate with sort interval I:
//   restore the saved particles, sort them once (untimed), run I steps (untimed: the cell
//   order decays as it does between sorts in a long run), then time the next I steps; the
//   first of them sorts again, so the timed window holds exactly one sort per I steps, as the
//   steady state does.
// Coordinate descent, one knob at a time in the order sort shift, sort interval, aggregation
// (atomic mode), workgroup size, particles per thread, threads per particle, start-tet
// staging, crossing-particle compaction, pid order (deterministic mode), then the sort
// interval again under the chosen kernel.  The saved state, counters and density are restored
// before returning, so tuning leaves the run where it was; only the options change.
#include <algorithm>
#include <cstring>
#include <vector>

#include "mp_ctx_impl.h"

using namespace mpctx;

namespace {

struct Saved {
    void* mem = nullptr;
    Soa soa{};
    int64_t n_slots = 0, next_pid = 0, since_sort = 0;
    bool need_sort = false;
    bool pid_ordered = true;
    unsigned long long ctr[mpd::kNumCounters] = {};
    std::vector<double> den;
};

hipError_t copy_soa(const Soa& dst, const Soa& src, int64_t n, hipStream_t st) {
    const size_t nd = sizeof(double) * (size_t)n, ni = sizeof(int32_t) * (size_t)n;
    double* const dd[6] = {dst.x, dst.y, dst.z, dst.vx, dst.vy, dst.vz};
    const double* const sd[6] = {src.x, src.y, src.z, src.vx, src.vy, src.vz};
    for (int k = 0; k < 6; k++) {
        hipError_t e = hipMemcpyAsync(dd[k], sd[k], nd, hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) return e;
    }
    hipError_t e = hipMemcpyAsync(dst.cell, src.cell, ni, hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dst.pid, src.pid, ni, hipMemcpyDeviceToDevice, st);
    return e;
}

int save(mp_ctx* c, Saved& s) {
    s.n_slots = c->n_slots;
    s.next_pid = c->next_pid;
    s.since_sort = c->since_sort;
    s.need_sort = c->need_sort;
    s.pid_ordered = c->pid_ordered;
    if (alloc_soa(&s.mem, &s.soa, std::max<int64_t>(c->n_slots, 1)) != hipSuccess) {
        (void)hipGetLastError();
        return set_err(c, MP_ERR_ALLOC, "mp_autotune: no device memory for the saved particle state");
    }
    HIP_TRY(c, copy_soa(s.soa, c->soa[c->cur], c->n_slots, c->stream));
    HIP_TRY(c, hipMemcpyAsync(s.ctr, c->ctr, sizeof(s.ctr), hipMemcpyDeviceToHost, c->stream));
    s.den.resize((size_t)c->n_nodes);
    HIP_TRY(c, hipMemcpyAsync(s.den.data(), c->den_out, sizeof(double) * s.den.size(), hipMemcpyDeviceToHost,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return MP_OK;
}

int restore(mp_ctx* c, const Saved& s) {
    // both SoA buffers have capacity c->cap >= s.n_slots (steps never add particles)
    HIP_TRY(c, copy_soa(c->soa[c->cur], s.soa, s.n_slots, c->stream));
    c->n_slots = s.n_slots;
    c->next_pid = s.next_pid;
    c->since_sort = s.since_sort;
    c->need_sort = s.need_sort;
    c->pid_ordered = s.pid_ordered;
    HIP_TRY(c, hipMemcpyAsync(c->ctr, s.ctr, sizeof(s.ctr), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->den_out, s.den.data(), sizeof(double) * s.den.size(), hipMemcpyHostToDevice,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return MP_OK;
}

struct Trial {
    mp_ctx* c;
    const Saved* saved;
    double qm, dt, spwt;
    int32_t mode;
    bool multi;
    int32_t min_steps;
    hipEvent_t ev[2];
    std::vector<mp_tune_trial> log;

    // Time per step (ms) of options o, by the protocol in the file comment.
    int run(const mp_options& o, double* ms_per_step) {
        int rc = restore(c, *saved);
        if (rc) return rc;
        c->opt = o;
        rc = mp_sort_now(c);
        if (rc) return rc;
        const int interval = o.sort_interval > 0 ? o.sort_interval : min_steps;
        const int timed = std::max(interval, min_steps);
        if (multi) {
            if ((rc = mp_step_many(c, qm, dt, spwt, mode, interval))) return rc;
        } else {
            for (int k = 0; k < interval; k++)
                if ((rc = mp_step(c, qm, dt, spwt, mode))) return rc;
        }
        HIP_TRY(c, hipEventRecord(ev[0], c->stream));
        if (multi) {
            if ((rc = mp_step_many(c, qm, dt, spwt, mode, timed))) return rc;
        } else {
            for (int k = 0; k < timed; k++)
                if ((rc = mp_step(c, qm, dt, spwt, mode))) return rc;
        }
        HIP_TRY(c, hipEventRecord(ev[1], c->stream));
        HIP_TRY(c, hipEventSynchronize(ev[1]));
        float ms = 0.f;
        HIP_TRY(c, hipEventElapsedTime(&ms, ev[0], ev[1]));
        *ms_per_step = (double)ms / timed;
        log.push_back(mp_tune_trial{o, *ms_per_step});
        return MP_OK;
    }
};

bool valid(const mp_options&) { return true; }

}  // namespace

extern "C" int mp_autotune(mp_ctx* c, double q_over_m, double dt, double spwt, int32_t mode, int32_t min_steps,
                           mp_options* best, double* best_ms_per_step, mp_tune_trial* trials,
                           int32_t max_trials, int32_t* n_trials) {
    if (!c) return MP_ERR_ARG;
    const bool multi = (mode & MP_TUNE_MULTI_STEP) != 0;
    mode &= ~MP_TUNE_MULTI_STEP;
    if (mode != MP_SCATTER_ATOMIC && mode != MP_SCATTER_DETERMINISTIC)
        return set_err(c, MP_ERR_ARG, "mp_autotune: unknown scatter mode");
    if (min_steps < 1 || min_steps > 1000) return set_err(c, MP_ERR_ARG, "mp_autotune: min_steps must be 1..1000");
    if (max_trials < 0 || (max_trials > 0 && !trials)) return set_err(c, MP_ERR_ARG, "mp_autotune: bad trial buffer");
    if (!c->have_field) return set_err(c, MP_ERR_STATE, "mp_autotune: no field uploaded");
    if (c->n_slots == 0) return set_err(c, MP_ERR_STATE, "mp_autotune: no resident particles");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));

    const mp_options start = c->opt;
    const bool timing = c->timing;
    c->timing = false;
    Saved saved;
    int rc = save(c, saved);
    const bool have_saved = rc == MP_OK;
    Trial tr{c, &saved, q_over_m, dt, spwt, mode, multi, min_steps, {nullptr, nullptr}, {}};
    if (!rc && (hipEventCreate(&tr.ev[0]) != hipSuccess || hipEventCreate(&tr.ev[1]) != hipSuccess))
        rc = set_err(c, MP_ERR_HIP, "mp_autotune: hipEventCreate failed");

    mp_options cur = start;
    double cur_ms = 0.0;
    if (!rc) rc = tr.run(cur, &cur_ms);
    // one knob at a time: keep a candidate only if it beats the incumbent by more than the
    // run-to-run spread of a window (~1%), so noise does not walk the options away
    constexpr double kMargin = 0.99;
    auto sweep = [&](auto set, std::initializer_list<int> values) {
        for (int v : values) {
            if (rc) return;
            mp_options o = cur;
            set(o, v);
            if (!valid(o) || memcmp(&o, &cur, sizeof(o)) == 0) continue;
            double ms = 0.0;
            rc = tr.run(o, &ms);
            if (!rc && ms < kMargin * cur_ms) {
                cur = o;
                cur_ms = ms;
            }
        }
    };
    if (multi && mode == MP_SCATTER_ATOMIC) {
        // the multi-step kernel: aggregation 0 or the (step, node) hash, one particle per thread
        sweep([](mp_options& o, int v) { o.sort_shift = v; }, {0, 1, 2, 3, 4, 5, 6});
        sweep([](mp_options& o, int v) { o.sort_interval = v; }, {4, 6, 8, 10, 12, 16, 20, 24, 32});
        sweep([](mp_options& o, int v) { o.scatter_agg = v; }, {0, 2});
        sweep([](mp_options& o, int v) { o.gather_sort = v; }, {0, 1});
        sweep([](mp_options& o, int v) { o.sort_order = v; }, {0, 1});
        sweep([](mp_options& o, int v) { o.sort_shift = v; }, {0, 1, 2, 3, 4, 5, 6});
        sweep([](mp_options& o, int v) { o.sort_interval = v; }, {4, 6, 8, 10, 12, 16, 20, 24, 32});
    } else {
        sweep([](mp_options& o, int v) { o.sort_shift = v; }, {0, 1, 2, 3, 4, 5, 6});
        sweep([](mp_options& o, int v) { o.sort_interval = v; }, {4, 6, 8, 10, 14, 20, 30});
        if (mode == MP_SCATTER_ATOMIC) sweep([](mp_options& o, int v) { o.scatter_agg = v; }, {0, 2});
        sweep([](mp_options& o, int v) { o.sort_order = v; }, {0, 1});
        sweep([](mp_options& o, int v) { o.block_size = v; }, {64, 128, 256});
        sweep([](mp_options& o, int v) { o.particles_per_thread = v; }, {1, 2, 4, 8});
        if (mode == MP_SCATTER_DETERMINISTIC) {
            sweep([](mp_options& o, int v) { o.det_algo = v; }, {0, 1, 2});
            sweep([](mp_options& o, int v) { o.keep_pid_order = v; }, {0, 1});
            sweep([](mp_options& o, int v) { o.sort_pid_bits = v; }, {0, 2, 4, 6});
        }
        sweep([](mp_options& o, int v) { o.sort_interval = v; }, {4, 6, 8, 10, 14, 20, 30});
    }
    // the winner against the starting options once more, both timed again (best of two)
    if (!rc && memcmp(&cur, &start, sizeof(cur)) != 0) {
        double a0 = 0.0, a1 = 0.0, b0 = 0.0, b1 = 0.0;
        rc = tr.run(cur, &a0);
        if (!rc) rc = tr.run(start, &b0);
        if (!rc) rc = tr.run(cur, &a1);
        if (!rc) rc = tr.run(start, &b1);
        if (!rc) {
            const double a = std::min(a0, a1), b = std::min(b0, b1);
            if (b <= a) {
                cur = start;
                cur_ms = b;
            } else {
                cur_ms = a;
            }
        }
    }

    // leave the run as it was, with the chosen options; after a failed trial too (the state it
    // left is undone, and its error is the one reported)
    if (have_saved) {
        const int rr = restore(c, saved);
        if (!rc) rc = rr;
    }
    c->opt = rc ? start : cur;
    c->timing = timing;
    for (hipEvent_t e : tr.ev)
        if (e) (void)hipEventDestroy(e);
    dfree(saved.mem);
    if (rc) return rc;
    if (best) *best = cur;
    if (best_ms_per_step) *best_ms_per_step = cur_ms;
    const int32_t nt = (int32_t)tr.log.size();
    for (int32_t k = 0; k < std::min(nt, max_trials); k++) trials[k] = tr.log[(size_t)k];
    if (n_trials) *n_trials = nt;
    return MP_OK;
}
