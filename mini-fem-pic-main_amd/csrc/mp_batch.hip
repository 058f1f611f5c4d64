// Batched ensemble (declared in include/mp_abi.h, mp_batch_*): many independent runs on one
// mesh pushed in one launch per step.  The reference has no counterpart; each run obeys the
// MoveParticles semantics of src/mini-pic.cpp:268-320 on its own particles and density.
//
// Layout: one SoA particle array for all runs plus a uint16 run id per particle, kept
// grouped by (run, cell bucket) by the counting sort.  The push reads q/m and dt of the
// particle's run from a small table, scatters into den[run * n_nodes + node] through the
// per-workgroup LDS hash, and reads the live particle count from device memory, so the
// push + tally + normalisation sequence is captured once in a hipGraph per buffer parity
// and replayed every step.
#include <algorithm>
#include <cstring>
#include <vector>

#include "mp_common.h"
#include "mp_ctx_impl.h"

namespace mpd {

// The push of k_push_soa (same prologue: state, run and cell loaded together, no branch on
// the cell before the start tet is evaluated) with the run's q/m and dt from the run table.
// Persistent: the grid is what fits on the device at once (a multiple of 8), and each
// workgroup loops over 256-particle chunks of its XCD's contiguous eighth of the array, so one
// launch (and one captured graph) covers any live count up to the capacity; the count is read
// from device memory (n_dev) or given (i0, n_host: one run's range).
template <int WALL, int DET>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 8)))
k_push_batch(const TetBlock* __restrict__ tets, Soa s, int64_t i0, int64_t n_host,
             const int64_t* __restrict__ n_dev, const double* __restrict__ par, int32_t nn,
             double* __restrict__ den, uint32_t* __restrict__ keys, double* __restrict__ vals,
             int32_t* __restrict__ part, int cap) {
    const int64_t n = n_dev ? *n_dev : i0 + n_host;
    const int64_t chunks = (n - i0 + 255) / 256;
    const int64_t per_xcd = (chunks + kXcds - 1) / kXcds;
    const int64_t xcd = blockIdx.x % kXcds, stride = gridDim.x / kXcds;
    int n_act = 0, n_lost = 0, n_cross = 0, n_moves = 0, n_cap = 0, n_refl = 0;
    // block-uniform loop: every thread of the workgroup takes part in each chunk's scatter
    for (int64_t k = blockIdx.x / kXcds; k < per_xcd; k += stride) {
        const int64_t ch = xcd * per_xcd + k;
        if (ch >= chunks) break;
        const int64_t i = i0 + ch * 256 + threadIdx.x;
        bool ok = false;
        int32_t fc = -1, base = 0;
        double l0 = 0.0, l1 = 0.0, l2 = 0.0, l3 = 0.0;
        if (i < n) {
            const int32_t c_in = s.cell[i];
            const int32_t r = s.run[i];
            double vx = ld(s.vx + i), vy = ld(s.vy + i), vz = ld(s.vz + i);
            const double x0 = ld(s.x + i), y0 = ld(s.y + i), z0 = ld(s.z + i);
            __builtin_amdgcn_sched_barrier(0);
            const bool live = c_in >= 0;
            int32_t c = live ? c_in : 0;
            const double qm = par[3 * r], dt = par[3 * r + 1];
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
            const int rr = walk_start<WALL>(tets, g0, c, x, y, z, vx, vy, vz, l0, l1, l2, l3, live ? cap : 0, st);
            if (live) {
                n_act++;
                base = r * nn;
                n_cross += (st.moves - st.refl) > 0;
                n_moves += st.moves;
                n_refl += st.refl;
                if (rr == kInside) {
                    s.cell[i] = c;
                    ok = true;
                    fc = c;
                    if (DET) put_records(keys, vals, s.pid[i], tet_at(tets, c), l0, l1, l2, l3, (uint32_t)base);
                } else {
                    s.cell[i] = -1;
                    n_lost++;
                    n_cap += (rr == kCapped);
                }
                st_(s.x + i, x); st_(s.y + i, y); st_(s.z + i, z);
                st_(s.vx + i, vx); st_(s.vy + i, vy); st_(s.vz + i, vz);
            }
        }
        if (!DET) scatter_lds(tets, den, ok, fc, l0, l1, l2, l3, base);
    }
    tally_counts(part, n_act, n_lost, n_cross, n_moves, n_cap, n_refl);
}

// K <= kStepMaxLaunch steps of k_push_batch in one launch, as k_push_steps does for one run:
// the particle and its tet's record stay in registers, each lane runs its steps on its own
// (lane_steps); step k scatters into den + k * R * nn (run r's nodes at r * nn) through the
// (step, node)-keyed LDS hash.  Atomic scatter only.
// log2 of the (step, node) LDS hash slots of a workgroup: 4096 slots (48 KB, 3 workgroups
// per CU).  The runs of an ensemble are often sparse (a workgroup's particles spread over
// hundreds of tets), so a 2048-slot table overflows into memory-side atomics, which bound the
// launch; 4096 slots are 4% faster on 64 runs of 1e5..4e6 particles despite the lower occupancy.
#ifndef MP_BATCH_STEP_BITS
#define MP_BATCH_STEP_BITS 12
#endif
#ifndef MP_BATCH_STEPS_WAVES
#define MP_BATCH_STEPS_WAVES 3
#endif
// threads per workgroup of k_push_batch_steps (a multiple of 64, at most 1024)
#ifndef MP_BATCH_STEPS_THREADS
#define MP_BATCH_STEPS_THREADS 256
#endif
template <int WALL>
__global__ void __launch_bounds__(MP_BATCH_STEPS_THREADS) __attribute__((amdgpu_waves_per_eu(MP_BATCH_STEPS_WAVES, 8)))
k_push_batch_steps(const TetBlock* __restrict__ tets, Soa s, int64_t n, const double* __restrict__ par, int32_t nn,
                   int32_t R, double* __restrict__ den, int K, int32_t* __restrict__ part, int cap) {
    const int64_t i = xcd_block() * blockDim.x + threadIdx.x;
    int n_act = 0, n_lost = 0, n_cross = 0, n_moves = 0, n_cap = 0, n_refl = 0;
    int32_t c_in = -1, r = 0;
    double x = 0.0, y = 0.0, z = 0.0, vx = 0.0, vy = 0.0, vz = 0.0;
    if (i < n) {
        c_in = s.cell[i];
        r = s.run[i];
        vx = ld(s.vx + i);
        vy = ld(s.vy + i);
        vz = ld(s.vz + i);
        x = ld(s.x + i);
        y = ld(s.y + i);
        z = ld(s.z + i);
    }
    __builtin_amdgcn_sched_barrier(0);
    bool live = c_in >= 0;
    int32_t c = live ? c_in : 0;
    const double qm = par[3 * r], dt = par[3 * r + 1];
    TetRegs t;
    load_tet_regs(tet_at(tets, c), t);
    __shared__ uint32_t hk[1 << MP_BATCH_STEP_BITS];
    __shared__ double hv[1 << MP_BATCH_STEP_BITS];
    __shared__ int32_t shc[MP_BATCH_STEPS_THREADS / kWave * 6];
    __shared__ int32_t done;
    if (threadIdx.x == 0) done = 0;
    for (int q = threadIdx.x; q < (1 << MP_BATCH_STEP_BITS); q += blockDim.x) {
        hk[q] = 0xFFFFFFFFu;
        hv[q] = 0.0;
    }
    __syncthreads();
    lane_steps<WALL, true, MP_BATCH_STEP_BITS>(tets, t, c, live, x, y, z, vx, vy, vz, qm, dt, K, cap, den,
                                               (uint32_t)r * (uint32_t)nn, (uint32_t)R * (uint32_t)nn, hk, hv, n_act,
                                               n_lost, n_cross, n_moves, n_cap, n_refl);
    if (c_in >= 0) {
        s.cell[i] = live ? c : -1;
        st_(s.x + i, x); st_(s.y + i, y); st_(s.z + i, z);
        st_(s.vx + i, vx); st_(s.vy + i, vy); st_(s.vz + i, vz);
    }
    // no barrier: the last wave of the workgroup flushes the hash (1.7% faster on the
    // 64-run ensemble)
    steps_finish<MP_BATCH_STEP_BITS, true>(hk, hv, den, part, shc, &done, n_act, n_lost, n_cross, n_moves, n_cap,
                                           n_refl);
}

// den[q] *= spwt of the run / node volume over the K steps of a k_push_batch_steps launch,
// q = k * R * nn + run * nn + node (k_normalise_batch's expression, in place).
__global__ void k_normalise_batch_steps(double* __restrict__ den, const double* __restrict__ nvol,
                                        const double* __restrict__ par, int32_t nn, int64_t per_step,
                                        int64_t total) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= total) return;
    const int64_t rk = q % per_step;
    const int64_t r = rk / nn;
    den[q] = den[q] * (par[3 * r + 2] / nvol[rk - r * nn]);
}

// den[k] = raw[k] * (spwt of run k / nn) / node volume, raw cleared; k in [k0, k0 + count).
__global__ void k_normalise_batch(double* __restrict__ raw, double* __restrict__ out,
                                  const double* __restrict__ nvol, const double* __restrict__ par,
                                  int32_t nn, int64_t k0, int64_t count) {
    const int64_t k = k0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= k0 + count) return;
    const int64_t r = k / nn;
    out[k] = raw[k] * (par[3 * r + 2] / nvol[k - r * nn]);
    raw[k] = 0.0;
}

// dst[r] = src[r * stride] for r in [0, count): run starts from the bucket offsets.
__global__ void k_gather_stride(const int32_t* __restrict__ src, int64_t stride, int32_t count,
                                int32_t* __restrict__ dst) {
    const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < count) dst[r] = src[(int64_t)r * stride];
}

__global__ void k_set_count(const int32_t* __restrict__ src, int64_t* __restrict__ dst) {
    if (threadIdx.x == 0) *dst = *src;
}

__global__ void k_unpack_run(const double* __restrict__ pos, const double* __restrict__ vel,
                             const int32_t* __restrict__ cell, int64_t n, int64_t pid0, uint16_t run, Soa s,
                             int64_t at) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t d = at + i;
    s.x[d] = pos[3 * i + 0]; s.y[d] = pos[3 * i + 1]; s.z[d] = pos[3 * i + 2];
    s.vx[d] = vel[3 * i + 0]; s.vy[d] = vel[3 * i + 1]; s.vz[d] = vel[3 * i + 2];
    s.cell[d] = cell[i];
    s.pid[d] = (int32_t)(pid0 + i);
    s.run[d] = run;
}

static inline unsigned grid256(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace mpd

using namespace mpctx;

struct mp_batch {
    mp_ctx* ctx = nullptr;
    int32_t R = 0;
    int32_t nn = 0;
    std::vector<double> par;          // [3R]: q/m, dt, spwt
    std::vector<int64_t> pid_base;    // first global pid of each run, -1 = not set
    std::vector<int64_t> run_n;       // particles set per run
    std::vector<int32_t> run_start;   // [R+1] array ranges after the last sort
    void* soa_mem[2] = {nullptr, nullptr};
    uint16_t* run_mem[2] = {nullptr, nullptr};
    Soa soa[2] = {};
    int cur = 0;
    int64_t cap = 0, n_slots = 0, next_pid = 0, since_sort = 0;
    bool need_sort = true;
    int64_t* n_dev = nullptr;
    double* par_d = nullptr;
    double* den_raw = nullptr;
    double* den_out = nullptr;
    unsigned long long* ctr = nullptr;
    int32_t* tally = nullptr;
    int64_t tally_n = 0;
    int32_t* cnt = nullptr;
    int32_t* cursor = nullptr;
    int64_t cnt_n = 0;
    int32_t* runs_d = nullptr;
    int32_t* pinned = nullptr;
    // deterministic scratch
    uint32_t* keys[2] = {nullptr, nullptr};
    double* vals[2] = {nullptr, nullptr};
    int64_t rec_cap = 0;
    void* sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    int32_t* seg_bounds = nullptr;
    // graphs (one per buffer parity)
    bool use_graph = true;
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    int64_t graph_slots[2] = {-1, -1};
    // the options a graph's launch was captured with (the kernel instance and its walk cap)
    int32_t graph_wall[2] = {-1, -1};
    int32_t graph_cap[2] = {-1, -1};
    // per-step densities of the last mp_batch_step_many, [steps][R * nn]
    double* den_steps = nullptr;
    int64_t den_steps_n = 0;
    int32_t den_steps_k = 0;
};

namespace {

void drop_graphs(mp_batch* b) {
    for (int k = 0; k < 2; k++) {
        if (b->exec[k]) (void)hipGraphExecDestroy(b->exec[k]);
        if (b->graph[k]) (void)hipGraphDestroy(b->graph[k]);
        b->exec[k] = nullptr;
        b->graph[k] = nullptr;
        b->graph_slots[k] = -1;
    }
}

int reserve_particles(mp_batch* b, int64_t n) {
    mp_ctx* c = b->ctx;
    if (n <= b->cap) return MP_OK;
    const int64_t nc = std::max<int64_t>(2 * n, 1024);
    void* mem[2] = {nullptr, nullptr};
    uint16_t* rm[2] = {nullptr, nullptr};
    Soa s[2];
    for (int k = 0; k < 2; k++)
        if (alloc_soa(&mem[k], &s[k], nc) != hipSuccess || dalloc(&rm[k], nc) != hipSuccess) {
            dfree(mem[0]); dfree(mem[1]); dfree(rm[0]); dfree(rm[1]);
            (void)hipGetLastError();
            return set_err(c, MP_ERR_ALLOC, "mp_batch: particle allocation failed");
        }
    for (int k = 0; k < 2; k++) s[k].run = rm[k];
    if (b->n_slots > 0) {
        const Soa& o = b->soa[b->cur];
        const size_t nd = sizeof(double) * (size_t)b->n_slots, ni = sizeof(int32_t) * (size_t)b->n_slots;
        double* dsrc[6] = {o.x, o.y, o.z, o.vx, o.vy, o.vz};
        double* ddst[6] = {s[0].x, s[0].y, s[0].z, s[0].vx, s[0].vy, s[0].vz};
        for (int k = 0; k < 6; k++) HIP_TRY(c, hipMemcpyAsync(ddst[k], dsrc[k], nd, hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(c, hipMemcpyAsync(s[0].cell, o.cell, ni, hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(c, hipMemcpyAsync(s[0].pid, o.pid, ni, hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(c, hipMemcpyAsync(s[0].run, o.run, sizeof(uint16_t) * (size_t)b->n_slots, hipMemcpyDeviceToDevice,
                                  c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    for (int k = 0; k < 2; k++) {
        dfree(b->soa_mem[k]);
        dfree(b->run_mem[k]);
        b->soa_mem[k] = mem[k];
        b->run_mem[k] = rm[k];
        b->soa[k] = s[k];
    }
    b->cur = 0;
    b->cap = nc;
    drop_graphs(b);
    return MP_OK;
}

int reserve_tally_b(mp_batch* b, int64_t n) {
    const int64_t need = mpd::tally_ints(std::max<int64_t>(n, 1), 256);
    if (need <= b->tally_n) return MP_OK;
    dfree(b->tally);
    b->tally = nullptr;
    b->tally_n = 0;
    if (dalloc(&b->tally, need + need / 2) != hipSuccess) return set_err(b->ctx, MP_ERR_ALLOC, "mp_batch: tally allocation failed");
    b->tally_n = need + need / 2;
    drop_graphs(b);
    return MP_OK;
}

// Counting sort by (run, cell bucket); refreshes the device count and the host run ranges.
int batch_sort(mp_batch* b) {
    mp_ctx* c = b->ctx;
    b->since_sort = 0;
    b->need_sort = false;
    const int shift = c->opt.sort_shift;
    const int64_t nb = (((int64_t)c->n_tets - 1) >> shift) + 1;
    const int64_t nkeys = nb * b->R;
    if (nkeys + 1 > INT32_MAX) return set_err(c, MP_ERR_ARG, "mp_batch: too many sort buckets");
    if (nkeys + 1 > b->cnt_n) {
        dfree(b->cnt);
        dfree(b->cursor);
        b->cnt = b->cursor = nullptr;
        b->cnt_n = 0;
        if (dalloc(&b->cnt, nkeys + 1) != hipSuccess || dalloc(&b->cursor, nkeys + 1) != hipSuccess)
            return set_err(c, MP_ERR_ALLOC, "mp_batch: sort buffers allocation failed");
        b->cnt_n = nkeys + 1;
    }
    int rc = reserve_scan(c, nkeys + 1);
    if (rc) return rc;
    const Soa& src = b->soa[b->cur];
    const Soa& dst = b->soa[1 - b->cur];
    HIP_TRY(c, mpd::launch_fill_i32(b->cnt, nkeys + 1, 0, c->stream));
    const int32_t* trank = c->opt.sort_order == 1 ? c->tet_rank : nullptr;
    HIP_TRY(c, mpd::launch_cell_hist(src.cell, src.run, b->n_slots, shift, (int32_t)nb, b->cnt, c->stream, trank));
    HIP_TRY(c, mpd::launch_excl_scan(b->cnt, nkeys + 1, b->cursor, c->scan_tmp, c->stream));
    hipLaunchKernelGGL(mpd::k_gather_stride, dim3(mpd::grid256(b->R + 1)), dim3(256), 0, c->stream, b->cursor, nb,
                       b->R + 1, b->runs_d);
    HIP_TRY(c, hipGetLastError());
    hipLaunchKernelGGL(mpd::k_set_count, dim3(1), dim3(64), 0, c->stream, b->cursor + nkeys, b->n_dev);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipMemcpyAsync(b->pinned, b->runs_d, sizeof(int32_t) * (size_t)(b->R + 1), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, mpd::launch_cell_scatter(src, b->n_slots, shift, (int32_t)nb, b->cursor, dst, c->stream, trank));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    b->cur = 1 - b->cur;
    for (int32_t r = 0; r <= b->R; r++) b->run_start[r] = b->pinned[r];
    b->n_slots = b->run_start[b->R];
    return MP_OK;
}

// Grid of the persistent k_push_batch<W, D> for `span` particles: one workgroup per 256
// particles, at most as many as are resident on the device at once, a multiple of 8 (the
// kernel's XCD mapping).
template <int W, int D>
unsigned persistent_grid(int64_t span) {
    static int resident = 0;
    if (resident == 0) {
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, mpd::k_push_batch<W, D>, 256, 0) != hipSuccess ||
            cus < 1 || per_cu < 1) {
            (void)hipGetLastError();
            cus = 256;
            per_cu = 4;
        }
        resident = std::max(mpd::kXcds, cus * per_cu / mpd::kXcds * mpd::kXcds);
    }
    const int64_t need = ((span + 255) / 256 + mpd::kXcds - 1) / mpd::kXcds * mpd::kXcds;
    return (unsigned)std::min<int64_t>(need, resident);
}

int launch_atomic_step(mp_batch* b, const int64_t* n_dev, int64_t i0, int64_t n_host, int64_t k0, int64_t kcount) {
    mp_ctx* c = b->ctx;
    const int64_t span = n_dev ? b->n_slots : n_host;
    if (span > 0) {
        const unsigned g = c->opt.wall_mode == MP_WALL_ABSORB ? persistent_grid<0, 0>(span) : persistent_grid<1, 0>(span);
#define MP_B(W) hipLaunchKernelGGL((mpd::k_push_batch<W, 0>), dim3(g), dim3(256), 0, c->stream, c->tets, \
                                   b->soa[b->cur], i0, n_host, n_dev, b->par_d, b->nn, b->den_raw, \
                                   (uint32_t*)nullptr, (double*)nullptr, b->tally, c->opt.walk_cap)
        if (c->opt.wall_mode == MP_WALL_ABSORB) MP_B(0);
        else MP_B(1);
#undef MP_B
        HIP_TRY(c, hipGetLastError());
        HIP_TRY(c, mpd::launch_tally_finish(b->tally, g, b->ctr, c->stream));
    }
    if (kcount > 0) {
        hipLaunchKernelGGL(mpd::k_normalise_batch, dim3(mpd::grid256(kcount)), dim3(256), 0, c->stream, b->den_raw,
                           b->den_out, c->node_vol, b->par_d, b->nn, k0, kcount);
        HIP_TRY(c, hipGetLastError());
    }
    return MP_OK;
}

int det_step(mp_batch* b) {
    mp_ctx* c = b->ctx;
    const int64_t keys_n = (int64_t)b->R * b->nn;
    const int64_t m = 4 * b->next_pid;
    const int end_bit = bit_width((uint64_t)keys_n);
    if (m > b->rec_cap) {
        for (int k = 0; k < 2; k++) {
            dfree(b->keys[k]);
            dfree(b->vals[k]);
            b->keys[k] = nullptr;
            b->vals[k] = nullptr;
        }
        b->rec_cap = 0;
        for (int k = 0; k < 2; k++)
            if (dalloc(&b->keys[k], m) != hipSuccess || dalloc(&b->vals[k], m) != hipSuccess)
                return set_err(c, MP_ERR_ALLOC, "mp_batch: deterministic records allocation failed");
        b->rec_cap = m;
    }
    const size_t need = mpd::radix_sort_temp_bytes(b->rec_cap, end_bit);
    if (need > b->sort_tmp_bytes) {
        dfree(b->sort_tmp);
        b->sort_tmp = nullptr;
        b->sort_tmp_bytes = 0;
        if (hipMalloc(&b->sort_tmp, need) != hipSuccess) return set_err(c, MP_ERR_ALLOC, "mp_batch: sort scratch failed");
        b->sort_tmp_bytes = need;
    }
    if (!b->seg_bounds && dalloc(&b->seg_bounds, 3 * keys_n + 1) != hipSuccess)
        return set_err(c, MP_ERR_ALLOC, "mp_batch: segment table allocation failed");
    HIP_TRY(c, mpd::launch_fill_u32(b->keys[0], m, 0xFFFFFFFFu, c->stream));
    if (b->n_slots > 0) {
        const unsigned g = c->opt.wall_mode == MP_WALL_ABSORB ? persistent_grid<0, 1>(b->n_slots)
                                                              : persistent_grid<1, 1>(b->n_slots);
#define MP_B(W) hipLaunchKernelGGL((mpd::k_push_batch<W, 1>), dim3(g), dim3(256), 0, c->stream, c->tets, \
                                   b->soa[b->cur], (int64_t)0, b->n_slots, (const int64_t*)nullptr, b->par_d, \
                                   b->nn, (double*)nullptr, b->keys[0], b->vals[0], b->tally, c->opt.walk_cap)
        if (c->opt.wall_mode == MP_WALL_ABSORB) MP_B(0);
        else MP_B(1);
#undef MP_B
        HIP_TRY(c, hipGetLastError());
        HIP_TRY(c, mpd::launch_tally_finish(b->tally, g, b->ctr, c->stream));
    }
    HIP_TRY(c, mpd::radix_sort_pairs(b->sort_tmp, b->sort_tmp_bytes, b->keys[0], b->keys[1], b->vals[0], b->vals[1], m,
                                     end_bit, c->stream));
    HIP_TRY(c, mpd::launch_det_fold(b->keys[1], b->vals[1], m, b->seg_bounds, c->node_vol, (int32_t)keys_n, 0.0,
                                    b->den_out, c->stream, b->par_d + 2, b->nn));
    return MP_OK;
}

int check_ready(mp_batch* b) {
    for (int32_t r = 0; r < b->R; r++)
        if (b->pid_base[r] < 0) return set_err(b->ctx, MP_ERR_STATE, "mp_batch: run " + std::to_string(r) + " has no particles set");
    if (!b->ctx->have_field) return set_err(b->ctx, MP_ERR_STATE, "mp_batch: no field uploaded");
    return MP_OK;
}

int maybe_sort(mp_batch* b) {
    const int si = b->ctx->opt.sort_interval;
    if (b->need_sort || (si > 0 && b->since_sort >= si)) return batch_sort(b);
    return MP_OK;
}

}  // namespace

extern "C" {

int mp_batch_create(mp_ctx* c, int32_t n_runs, mp_batch** out) {
    if (!c || !out) return MP_ERR_ARG;
    *out = nullptr;
    if (n_runs < 1 || n_runs > 65535) return set_err(c, MP_ERR_ARG, "mp_batch_create: n_runs must be in 1..65535");
    if ((int64_t)n_runs * c->n_nodes > INT32_MAX / 2) return set_err(c, MP_ERR_ARG, "mp_batch_create: n_runs * n_nodes too large");
    HIP_TRY(c, hipSetDevice(c->device));
    mp_batch* b = new (std::nothrow) mp_batch;
    if (!b) return set_err(c, MP_ERR_ALLOC, "mp_batch_create: out of host memory");
    b->ctx = c;
    b->R = n_runs;
    b->nn = c->n_nodes;
    b->par.assign(3 * (size_t)n_runs, 0.0);
    b->pid_base.assign(n_runs, -1);
    b->run_n.assign(n_runs, 0);
    b->run_start.assign(n_runs + 1, 0);
    const int64_t nk = (int64_t)n_runs * c->n_nodes;
    if (dalloc(&b->n_dev, 1) != hipSuccess || dalloc(&b->par_d, 3 * (int64_t)n_runs) != hipSuccess ||
        dalloc(&b->den_raw, nk) != hipSuccess || dalloc(&b->den_out, nk) != hipSuccess ||
        dalloc(&b->ctr, mpd::kNumCounters) != hipSuccess || dalloc(&b->runs_d, n_runs + 1) != hipSuccess ||
        hipHostMalloc((void**)&b->pinned, sizeof(int32_t) * (size_t)(n_runs + 1), 0) != hipSuccess ||
        hipMemset(b->den_raw, 0, sizeof(double) * (size_t)nk) != hipSuccess ||
        hipMemset(b->den_out, 0, sizeof(double) * (size_t)nk) != hipSuccess ||
        hipMemset(b->ctr, 0, sizeof(unsigned long long) * mpd::kNumCounters) != hipSuccess ||
        hipMemset(b->n_dev, 0, sizeof(int64_t)) != hipSuccess) {
        mp_batch_destroy(b);
        return set_err(c, MP_ERR_ALLOC, "mp_batch_create: device allocation failed");
    }
    *out = b;
    return MP_OK;
}

void mp_batch_destroy(mp_batch* b) {
    if (!b) return;
    mp_ctx* c = b->ctx;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    drop_graphs(b);
    void* bufs[] = {b->soa_mem[0], b->soa_mem[1], b->run_mem[0], b->run_mem[1], b->n_dev, b->par_d, b->den_raw,
                    b->den_out, b->ctr, b->tally, b->cnt, b->cursor, b->runs_d, b->keys[0], b->keys[1], b->vals[0],
                    b->vals[1], b->sort_tmp, b->seg_bounds, b->den_steps};
    for (void* p : bufs) dfree(p);
    if (b->pinned) (void)hipHostFree(b->pinned);
    delete b;
}

int mp_batch_set_run(mp_batch* b, int32_t run, const double* pos, const double* vel, const int32_t* cell, int64_t n,
                     double q_over_m, double dt, double spwt) {
    if (!b) return MP_ERR_ARG;
    mp_ctx* c = b->ctx;
    if (run < 0 || run >= b->R) return set_err(c, MP_ERR_ARG, "mp_batch_set_run: run out of range");
    if (b->pid_base[run] >= 0) return set_err(c, MP_ERR_STATE, "mp_batch_set_run: run already set");
    if (n < 0 || (n > 0 && (!pos || !vel || !cell))) return set_err(c, MP_ERR_ARG, "mp_batch_set_run: bad arrays");
    if (b->next_pid + n > INT32_MAX / 4) return set_err(c, MP_ERR_ARG, "mp_batch_set_run: too many particles");
    for (int64_t i = 0; i < n; i++)
        if (cell[i] < -1 || cell[i] >= c->n_tets)
            return set_err(c, MP_ERR_ARG, "mp_batch_set_run: particle " + std::to_string(i) + " has cell out of range");
    HIP_TRY(c, hipSetDevice(c->device));
    int rc = reserve_particles(b, b->n_slots + n);
    if (rc) return rc;
    if (n > 0) {
        char* tmp = nullptr;
        const size_t bp = sizeof(double) * 3 * (size_t)n, bc = sizeof(int32_t) * (size_t)n;
        HIP_TRY(c, hipMalloc((void**)&tmp, 2 * bp + bc));
        hipError_t e = hipMemcpyAsync(tmp, pos, bp, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(tmp + bp, vel, bp, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(tmp + 2 * bp, cell, bc, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(mpd::k_unpack_run, dim3(mpd::grid256(n)), dim3(256), 0, c->stream, (const double*)tmp,
                               (const double*)(tmp + bp), (const int32_t*)(tmp + 2 * bp), n, b->next_pid,
                               (uint16_t)run, b->soa[b->cur], b->n_slots);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        dfree(tmp);
        if (e != hipSuccess) return set_err(c, MP_ERR_HIP, std::string("mp_batch_set_run: ") + hipGetErrorString(e));
    }
    b->pid_base[run] = b->next_pid;
    b->run_n[run] = n;
    b->next_pid += n;
    b->n_slots += n;
    b->par[3 * (size_t)run + 0] = q_over_m;
    b->par[3 * (size_t)run + 1] = dt;
    b->par[3 * (size_t)run + 2] = spwt;
    HIP_TRY(c, hipMemcpy(b->par_d, b->par.data(), sizeof(double) * b->par.size(), hipMemcpyHostToDevice));
    b->need_sort = true;
    return MP_OK;
}

int mp_batch_set_graph(mp_batch* b, int32_t enable) {
    if (!b) return MP_ERR_ARG;
    b->use_graph = enable != 0;
    return MP_OK;
}

int mp_batch_step(mp_batch* b, int32_t mode) {
    if (!b) return MP_ERR_ARG;
    mp_ctx* c = b->ctx;
    if (mode != MP_SCATTER_ATOMIC && mode != MP_SCATTER_DETERMINISTIC)
        return set_err(c, MP_ERR_ARG, "mp_batch_step: unknown scatter mode");
    int rc = check_ready(b);
    if (rc) return rc;
    HIP_TRY(c, hipSetDevice(c->device));
    rc = maybe_sort(b);
    if (rc) return rc;
    rc = reserve_tally_b(b, b->n_slots);
    if (rc) return rc;
    b->since_sort++;
    if (mode == MP_SCATTER_DETERMINISTIC) return det_step(b);
    const int64_t nk = (int64_t)b->R * b->nn;
    if (!b->use_graph) return launch_atomic_step(b, b->n_dev, 0, 0, 0, nk);
    const int k = b->cur;
    // a graph replays the launch as captured: capture again when the slots outgrew its grid or
    // mp_set_options changed the wall mode or the walk cap since
    if (!b->exec[k] || b->graph_slots[k] < b->n_slots || b->graph_wall[k] != c->opt.wall_mode ||
        b->graph_cap[k] != c->opt.walk_cap) {
        if (b->exec[k]) (void)hipGraphExecDestroy(b->exec[k]);
        if (b->graph[k]) (void)hipGraphDestroy(b->graph[k]);
        b->exec[k] = nullptr;
        b->graph[k] = nullptr;
        HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        rc = launch_atomic_step(b, b->n_dev, 0, 0, 0, nk);
        hipGraph_t g = nullptr;
        const hipError_t ee = hipStreamEndCapture(c->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        HIP_TRY(c, ee);
        b->graph[k] = g;
        HIP_TRY(c, hipGraphInstantiate(&b->exec[k], g, nullptr, nullptr, 0));
        b->graph_slots[k] = b->n_slots;
        b->graph_wall[k] = c->opt.wall_mode;
        b->graph_cap[k] = c->opt.walk_cap;
    }
    HIP_TRY(c, hipGraphLaunch(b->exec[k], c->stream));
    return MP_OK;
}

int mp_batch_step_many(mp_batch* b, int32_t mode, int32_t n_steps) {
    if (!b) return MP_ERR_ARG;
    mp_ctx* c = b->ctx;
    if (mode != MP_SCATTER_ATOMIC && mode != MP_SCATTER_DETERMINISTIC)
        return set_err(c, MP_ERR_ARG, "mp_batch_step_many: unknown scatter mode");
    if (n_steps < 1) return set_err(c, MP_ERR_ARG, "mp_batch_step_many: n_steps must be >= 1");
    int rc = check_ready(b);
    if (rc) return rc;
    HIP_TRY(c, hipSetDevice(c->device));
    const int64_t per_step = (int64_t)b->R * b->nn;
    const int64_t need = per_step * n_steps;
    if (need > b->den_steps_n) {
        dfree(b->den_steps);
        b->den_steps = nullptr;
        b->den_steps_n = 0;
        if (dalloc(&b->den_steps, need) != hipSuccess)
            return set_err(c, MP_ERR_ALLOC, "mp_batch_step_many: density buffer allocation failed");
        b->den_steps_n = need;
    }
    b->den_steps_k = 0;
    if (mode == MP_SCATTER_DETERMINISTIC) {
        for (int32_t k = 0; k < n_steps; k++) {
            rc = mp_batch_step(b, mode);
            if (rc) return rc;
            HIP_TRY(c, hipMemcpyAsync(b->den_steps + k * per_step, b->den_out, sizeof(double) * (size_t)per_step,
                                      hipMemcpyDeviceToDevice, c->stream));
        }
        b->den_steps_k = n_steps;
        return MP_OK;
    }
    HIP_TRY(c, mpd::launch_fill_f64(b->den_steps, need, 0.0, c->stream));
    int32_t done = 0;
    while (done < n_steps) {
        rc = maybe_sort(b);
        if (rc) return rc;
        const int32_t si = c->opt.sort_interval;
        int32_t chunk = n_steps - done;
        if (si > 0) chunk = (int32_t)std::min<int64_t>(chunk, si - b->since_sort);
        // (step, node) keys of one launch are 32-bit offsets below the empty-slot marker
        chunk = (int32_t)std::min<int64_t>(chunk, mpd::kStepMaxLaunch);
        chunk = (int32_t)std::min<int64_t>(chunk, std::max<int64_t>(1, (int64_t)0xFFFFFFFEu / per_step));
        if (per_step * chunk >= (int64_t)0xFFFFFFFFu)
            return set_err(c, MP_ERR_ARG, "mp_batch_step_many: n_runs * n_nodes too large for the step keys");
        rc = reserve_tally_b(b, b->n_slots);
        if (rc) return rc;
        if (b->n_slots > 0) {
            constexpr int kT = MP_BATCH_STEPS_THREADS;
            const unsigned g = (unsigned)(((b->n_slots + kT - 1) / kT + mpd::kXcds - 1) / mpd::kXcds * mpd::kXcds);
#define MP_BS(W) hipLaunchKernelGGL((mpd::k_push_batch_steps<W>), dim3(g), dim3(kT), 0, c->stream, c->tets, \
                                    b->soa[b->cur], b->n_slots, b->par_d, b->nn, b->R, \
                                    b->den_steps + (int64_t)done * per_step, chunk, b->tally, c->opt.walk_cap)
            if (c->opt.wall_mode == MP_WALL_ABSORB) MP_BS(0);
            else MP_BS(1);
#undef MP_BS
            HIP_TRY(c, hipGetLastError());
            HIP_TRY(c, mpd::launch_tally_finish(b->tally, g, b->ctr, c->stream));
        }
        b->since_sort += chunk;
        done += chunk;
    }
    hipLaunchKernelGGL(mpd::k_normalise_batch_steps, dim3((unsigned)((need + 255) / 256)), dim3(256), 0, c->stream,
                       b->den_steps, c->node_vol, b->par_d, b->nn, per_step, need);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipMemcpyAsync(b->den_out, b->den_steps + (int64_t)(n_steps - 1) * per_step,
                              sizeof(double) * (size_t)per_step, hipMemcpyDeviceToDevice, c->stream));
    b->den_steps_k = n_steps;
    return MP_OK;
}

int mp_batch_get_density_step(mp_batch* b, int32_t k, int32_t run, double* den) {
    if (!b || !den) return MP_ERR_ARG;
    mp_ctx* c = b->ctx;
    if (k < 0 || k >= b->den_steps_k)
        return set_err(c, MP_ERR_ARG, "mp_batch_get_density_step: no such step in the last mp_batch_step_many");
    if (run < 0 || run >= b->R) return set_err(c, MP_ERR_ARG, "mp_batch_get_density_step: bad run");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMemcpyAsync(den, b->den_steps + ((int64_t)k * b->R + run) * b->nn, sizeof(double) * (size_t)b->nn,
                              hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return MP_OK;
}

int mp_batch_step_sequential(mp_batch* b, int32_t mode) {
    if (!b) return MP_ERR_ARG;
    mp_ctx* c = b->ctx;
    if (mode != MP_SCATTER_ATOMIC) return set_err(c, MP_ERR_ARG, "mp_batch_step_sequential: atomic mode only");
    int rc = check_ready(b);
    if (rc) return rc;
    HIP_TRY(c, hipSetDevice(c->device));
    rc = maybe_sort(b);
    if (rc) return rc;
    rc = reserve_tally_b(b, b->n_slots);
    if (rc) return rc;
    b->since_sort++;
    for (int32_t r = 0; r < b->R; r++) {
        const int64_t i0 = b->run_start[r], n = b->run_start[r + 1] - b->run_start[r];
        rc = launch_atomic_step(b, nullptr, i0, n, (int64_t)r * b->nn, b->nn);
        if (rc) return rc;
    }
    return MP_OK;
}

int mp_batch_sync(mp_batch* b) {
    if (!b) return MP_ERR_ARG;
    HIP_TRY(b->ctx, hipStreamSynchronize(b->ctx->stream));
    return MP_OK;
}

int mp_batch_get_density(mp_batch* b, int32_t run, double* den) {
    if (!b || !den) return MP_ERR_ARG;
    mp_ctx* c = b->ctx;
    if (run < 0 || run >= b->R) return set_err(c, MP_ERR_ARG, "mp_batch_get_density: run out of range");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMemcpyAsync(den, b->den_out + (int64_t)run * b->nn, sizeof(double) * (size_t)b->nn,
                              hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return MP_OK;
}

int mp_batch_get_particles(mp_batch* b, int32_t run, double* pos, double* vel, int32_t* cell, int64_t* pid,
                           int64_t* n_inout) {
    if (!b || !n_inout) return MP_ERR_ARG;
    mp_ctx* c = b->ctx;
    if (run < 0 || run >= b->R) return set_err(c, MP_ERR_ARG, "mp_batch_get_particles: run out of range");
    HIP_TRY(c, hipSetDevice(c->device));
    const int64_t n = b->n_slots;
    const Soa& s = b->soa[b->cur];
    std::vector<double> h[6];
    std::vector<int32_t> hc((size_t)n), hid((size_t)n);
    std::vector<uint16_t> hr((size_t)n);
    double* src[6] = {s.x, s.y, s.z, s.vx, s.vy, s.vz};
    for (int k = 0; k < 6; k++) {
        h[k].resize((size_t)n);
        if (n) HIP_TRY(c, hipMemcpyAsync(h[k].data(), src[k], sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    }
    if (n) {
        HIP_TRY(c, hipMemcpyAsync(hc.data(), s.cell, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipMemcpyAsync(hid.data(), s.pid, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipMemcpyAsync(hr.data(), s.run, sizeof(uint16_t) * n, hipMemcpyDeviceToHost, c->stream));
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const int64_t base = b->pid_base[run], cnt = b->run_n[run];
    std::vector<int64_t> slot((size_t)std::max<int64_t>(cnt, 0), -1);
    int64_t live = 0;
    for (int64_t i = 0; i < n; i++)
        if (hr[i] == run && hc[i] >= 0) {
            slot[(size_t)(hid[i] - base)] = i;
            live++;
        }
    if (live > *n_inout) {
        *n_inout = live;
        return set_err(c, MP_ERR_ARG, "mp_batch_get_particles: output capacity too small");
    }
    int64_t k = 0;
    for (int64_t p = 0; p < cnt; p++) {
        const int64_t i = slot[(size_t)p];
        if (i < 0) continue;
        if (pos) for (int d = 0; d < 3; d++) pos[3 * k + d] = h[d][i];
        if (vel) for (int d = 0; d < 3; d++) vel[3 * k + d] = h[3 + d][i];
        if (cell) cell[k] = hc[i];
        if (pid) pid[k] = p;
        k++;
    }
    *n_inout = k;
    return MP_OK;
}

int mp_batch_get_counters(mp_batch* b, mp_counters* out) {
    if (!b || !out) return MP_ERR_ARG;
    mp_ctx* c = b->ctx;
    unsigned long long h[mpd::kNumCounters];
    HIP_TRY(c, hipMemcpyAsync(h, b->ctr, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    out->particle_steps = (int64_t)h[mpd::kCtrParticleSteps];
    out->lost = (int64_t)h[mpd::kCtrLost];
    out->crossed = (int64_t)h[mpd::kCtrCrossed];
    out->moves = (int64_t)h[mpd::kCtrMoves];
    out->cap_hits = (int64_t)h[mpd::kCtrCapHits];
    out->reflections = (int64_t)h[mpd::kCtrReflections];
    return MP_OK;
}

int mp_batch_reset_counters(mp_batch* b) {
    if (!b) return MP_ERR_ARG;
    mp_ctx* c = b->ctx;
    HIP_TRY(c, hipMemsetAsync(b->ctr, 0, sizeof(unsigned long long) * mpd::kNumCounters, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return MP_OK;
}

}  // extern "C"
