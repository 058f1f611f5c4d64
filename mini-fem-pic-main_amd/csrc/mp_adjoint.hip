// Adjoint of K push + walk + scatter steps under a frozen per-element field (mp_adjoint in
// include/mp_abi.h).  No reference counterpart: the forward pass is MoveParticles
// (src/mini-pic.cpp:287-308), repeated K times on the same particles, with absorbing or
// reflecting walls.
//
// With every particle's tet path c_0..c_{K-1} held fixed, one step is
//     v_{k+1} = v_k + (qm E[c_k]) dt,   x_{k+1} = x_k + v_{k+1} dt
// and den_K[n] = (spwt / V_n) sum over particles of lc_v(x_K) in the final tet, with
// grad lc_v = (-beta_v, gamma_v, -delta_v) / (6 V).  The reverse sweep keeps
//     lambda_x(k) = lambda_x(K) = sum_v dL/dlc_v grad lc_v,
//     lambda_v(k+1) = lambda_v(k+2) + lambda_x dt,   dL/dE[c_k] += qm dt lambda_v(k+1),
// and dL/d(drift) = lambda_v(0) . drift_dir.  A reflection maps x -> R x + b, v -> R v with
// R = I - 2 g g^T / |g|^2 (g the wall plane's normal, b fixed by the tet), so with reflecting
// walls the sweep first applies step k's reflections R_m .. R_1 (last first; R is symmetric)
// to lambda_x and lambda_v; it finds them by walking step k again from its start state, which
// the segment replay stores (48 B per particle and step).
//
// Memory: the state (x, v, cell) is copied every checkpoint_every steps during the forward
// pass; the backward pass restores each checkpoint into the second particle buffer, replays
// its segment recording c_k (4 B per particle per step), then sweeps the segment backwards.
// dL/dE contributions are summed per workgroup in an LDS hash keyed by tet (3 values per
// key) before the global FP64 atomics.
//
// Deterministic mode (MP_SCATTER_DETERMINISTIC): the last forward step scatters through the
// sorted records (the density is the ascending-pid serial fold, bit-identical to the serial
// loss), and dL/dE is summed exactly in 128-bit fixed point: each contribution v becomes
// x = v * 2^S (exact, a power of two) split into hi = floor(x) (int64) and lo = the fraction
// times 2^64 (uint64); lo words add with their carry into hi, so the sums are exact integers
// and do not depend on the order of the atomics.  S is the largest exponent with
// |x| * (particles * steps) <= 2^62 for the bound |v| <= |qm| dt (K dt) sqrt(3) max|lambda_x|
// (lambda_v grows by lambda_x dt per step back from zero).  The sums are converted to
// doubles once at the end: the same bits on every run.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "mp_common.h"
#include "mp_ctx_impl.h"

namespace mpd {

// One forward step without density scatter (SCATTER = 0) or with the LDS-hash scatter of
// the final step (SCATTER = 1).  path (optional) receives the cell at the start of the step.
// SCATTER = 2 writes the final step's deterministic records (node, weight) at 4 * pid + v.
// With states, the step's start state (x, y, z, vx, vy, vz) goes to states[d * n + i].
template <int SCATTER, int WALL>
__global__ void __launch_bounds__(256)
k_adj_step(const TetBlock* __restrict__ tets, Soa s, int64_t n, double qm, double dt, int cap,
           int32_t* __restrict__ path, double* __restrict__ states, double* __restrict__ den,
           uint32_t* __restrict__ keys = nullptr, double* __restrict__ vals = nullptr) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool ok = false;
    int32_t fc = -1;
    double l0 = 0.0, l1 = 0.0, l2 = 0.0, l3 = 0.0;
    if (i < n) {
        int32_t c = s.cell[i];
        if (path) path[i] = c;
        if (c >= 0) {
            if (states) {
                states[i] = s.x[i];
                states[n + i] = s.y[i];
                states[2 * n + i] = s.z[i];
                states[3 * n + i] = s.vx[i];
                states[4 * n + i] = s.vy[i];
                states[5 * n + i] = s.vz[i];
            }
            const TetView t = tet_at(tets, c);
            double vx = s.vx[i] + (qm * t.ex()) * dt;
            double vy = s.vy[i] + (qm * t.ey()) * dt;
            double vz = s.vz[i] + (qm * t.ez()) * dt;
            double x = s.x[i] + vx * dt;
            double y = s.y[i] + vy * dt;
            double z = s.z[i] + vz * dt;
            WalkStats st{0, 0};
            const int r = walk<WALL>(tets, c, x, y, z, vx, vy, vz, l0, l1, l2, l3, cap, st);
            s.x[i] = x; s.y[i] = y; s.z[i] = z;
            s.vx[i] = vx; s.vy[i] = vy; s.vz[i] = vz;
            if (r == kInside) {
                s.cell[i] = c;
                ok = true;
                fc = c;
                if (SCATTER == 2) put_records(keys, vals, (int64_t)s.pid[i], tet_at(tets, c), l0, l1, l2, l3);
            } else {
                s.cell[i] = -1;
            }
        }
    }
    if (SCATTER == 1) scatter_lds(tets, den, ok, fc, l0, l1, l2, l3);
}

// den = raw * spwt / V, g = 2 (den - target) spwt / V (dL/d raw weight), raw cleared;
// per-block partial sums of (den - target)^2 go to part[block].  With normalized, raw already
// holds den (the deterministic fold's output) and is left as it is.
__global__ void __launch_bounds__(256)
k_adj_loss(double* __restrict__ raw, const double* __restrict__ tgt, const double* __restrict__ nvol,
           int32_t nn, double spwt, double* __restrict__ g, double* __restrict__ part, bool normalized) {
    __shared__ double sh[256 / kWave];
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double r2 = 0.0;
    if (i < nn) {
        const double w = spwt / nvol[i];
        const double r = (normalized ? raw[i] : raw[i] * w) - tgt[i];
        g[i] = 2.0 * r * w;
        r2 = r * r;
        if (!normalized) raw[i] = 0.0;
    }
    r2 = wave_sum_f64(r2);
    if (lane_id() == 0) sh[threadIdx.x / kWave] = r2;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < (int)(blockDim.x / kWave); w++) s += sh[w];
        part[blockIdx.x] = s;
    }
}

// lambda_x of each particle in the domain after K steps; lambda_v cleared.
__global__ void k_adj_lambda(const TetBlock* __restrict__ tets, Soa s, int64_t n, const double* __restrict__ g,
                             double* __restrict__ lx, double* __restrict__ lv) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a = 0.0, b = 0.0, c3 = 0.0;
    const int32_t c = s.cell[i];
    if (c >= 0) {
        const TetView t = tet_at(tets, c);
        for (int q = 0; q < 4; q++) {
            const double f = g[t.con(q)] / (6.0 * t.vol());
            a += f * -t.c(q, 1);
            b += f * t.c(q, 2);
            c3 += f * -t.c(q, 3);
        }
    }
    lx[3 * i + 0] = a; lx[3 * i + 1] = b; lx[3 * i + 2] = c3;
    lv[3 * i + 0] = 0.0; lv[3 * i + 1] = 0.0; lv[3 * i + 2] = 0.0;
}

// LDS hash with three values per key (tet id), for the dL/dE sums.
__device__ __forceinline__ void lds_insert3(int32_t* hk, double* hv, int32_t key, double a, double b, double c) {
    uint32_t h = ((uint32_t)key * 0x9E3779B1u) >> (32 - 10);
    for (;;) {
        const int32_t prev = atomicCAS(&hk[h], -1, key);
        if (prev == -1 || prev == key) {
            atomicAdd(&hv[3 * h + 0], a);
            atomicAdd(&hv[3 * h + 1], b);
            atomicAdd(&hv[3 * h + 2], c);
            return;
        }
        h = (h + 1) & (kHashSlots - 1);
    }
}

// Largest |lambda_x| component of each block's particles, to part[block].
__global__ void __launch_bounds__(256)
k_adj_lmax(const double* __restrict__ lx, int64_t n, double* __restrict__ part) {
    __shared__ double sh[256 / kWave];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double m = 0.0;
    if (i < n) m = fmax(fmax(fabs(lx[3 * i]), fabs(lx[3 * i + 1])), fabs(lx[3 * i + 2]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, kWave));
    if (lane_id() == 0) sh[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0;
        for (int w = 0; w < (int)(blockDim.x / kWave); w++) r = fmax(r, sh[w]);
        part[blockIdx.x] = r;
    }
}

// 128-bit fixed point: value * 2^S = hi + lo * 2^-64 with hi = floor, lo in [0, 2^64).
struct Fixed {
    unsigned long long hi, lo;  // hi holds a two's complement int64
};

// For |x| >= 1 the fraction x - floor(x) and its product with 2^64 are exact; a negative x
// closer to zero than 2^-53 has a fraction that rounds to 1, taken as a carry into hi.
__device__ __forceinline__ Fixed to_fixed(double v, double scale) {
    const double x = v * scale;  // exact: scale is a power of two and |x| <= 2^62
    double f = floor(x);
    double frac = x - f;
    if (frac >= 1.0) {
        f += 1.0;
        frac = 0.0;
    }
    Fixed r;
    r.hi = (unsigned long long)(long long)f;
    r.lo = (unsigned long long)(frac * 18446744073709551616.0);
    return r;
}

// Exact accumulation: the lo word's carry goes into the hi word (LDS or global memory).
__device__ __forceinline__ void add_fixed(unsigned long long* lo, unsigned long long* hi, Fixed v) {
    const unsigned long long old = atomicAdd(lo, v.lo);
    const unsigned long long carry = (old + v.lo < old) ? 1ull : 0ull;
    atomicAdd(hi, v.hi + carry);
}

// dLdE[j] = (hi + lo * 2^-64) * 2^-S.
__global__ void k_adj_unfix(const unsigned long long* __restrict__ acc, int64_t m, double inv_scale,
                            double* __restrict__ out) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    const double hi = (double)(long long)acc[2 * j];
    const double lo = (double)acc[2 * j + 1] * 5.421010862427522e-20;  // 2^-64
    out[j] = (hi + lo) * inv_scale;
}

// lds_insert3 over (hi, lo) word pairs of 128-bit fixed-point sums.
__device__ __forceinline__ void lds_insert3_fixed(int32_t* hk, unsigned long long* hw, int32_t key, Fixed a,
                                                  Fixed b, Fixed c) {
    uint32_t h = ((uint32_t)key * 0x9E3779B1u) >> (32 - 10);
    for (;;) {
        const int32_t prev = atomicCAS(&hk[h], -1, key);
        if (prev == -1 || prev == key) {
            unsigned long long* w = hw + 6 * h;
            add_fixed(w + 1, w + 0, a);
            add_fixed(w + 3, w + 2, b);
            add_fixed(w + 5, w + 4, c);
            return;
        }
        h = (h + 1) & (kHashSlots - 1);
    }
}

// The reverse sweep of k_adj_sweep with dL/dE summed exactly in 128-bit fixed point:
// acc[2 * (3 * tet + d)] = hi word, acc[... + 1] = lo word, scale 2^S.
// The plane normals of one walk's reflections, in order (at most kMaxRefl per step).
constexpr int kMaxRefl = 4;
struct ReflRec {
    double g[kMaxRefl][3];
    int n = 0;
    bool over = false;
    __device__ __forceinline__ void operator()(double gx, double gy, double gz) {
        if (n == kMaxRefl) {
            over = true;
            return;
        }
        g[n][0] = gx;
        g[n][1] = gy;
        g[n][2] = gz;
        n++;
    }
};

// Reflecting walls: step k of particle i walked again from its start state st (k_adj_step's
// push and walk, the same bits), and its reflections applied to lambda_x (a) and lambda_v
// (v), the last one first.  *over is set if a step reflected the particle more than
// kMaxRefl times.
__device__ __forceinline__ void unreflect(const TetBlock* __restrict__ tets, const double* __restrict__ st,
                                          int64_t n, int64_t i, int32_t c, double qm, double dt, int cap,
                                          double& ax, double& ay, double& az, double& vx, double& vy,
                                          double& vz, int32_t* __restrict__ over) {
    const TetView t = tet_at(tets, c);
    double ux = st[3 * n + i] + (qm * t.ex()) * dt;
    double uy = st[4 * n + i] + (qm * t.ey()) * dt;
    double uz = st[5 * n + i] + (qm * t.ez()) * dt;
    double x = st[i] + ux * dt;
    double y = st[n + i] + uy * dt;
    double z = st[2 * n + i] + uz * dt;
    ReflRec rec;
    WalkStats ws{0, 0};
    double l0, l1, l2, l3;
    int32_t cc = c;
    (void)walk<1>(tets, cc, x, y, z, ux, uy, uz, l0, l1, l2, l3, cap, ws, rec);
    if (rec.over) *over = 1;
    for (int j = rec.n - 1; j >= 0; j--) {
        const double gx = rec.g[j][0], gy = rec.g[j][1], gz = rec.g[j][2];
        const double gg = (gx * gx + gy * gy) + gz * gz;
        const double pa = (2.0 * ((ax * gx + ay * gy) + az * gz)) / gg;
        ax = ax - pa * gx;
        ay = ay - pa * gy;
        az = az - pa * gz;
        const double pv = (2.0 * ((vx * gx + vy * gy) + vz * gz)) / gg;
        vx = vx - pv * gx;
        vy = vy - pv * gy;
        vz = vz - pv * gz;
    }
}

template <int WALL>
__global__ void __launch_bounds__(256)
k_adj_sweep_fixed(const int32_t* __restrict__ path, int32_t k0, int32_t k1, int64_t n, double* __restrict__ lx,
                  double* __restrict__ lv, double qm, double dt, unsigned long long* __restrict__ acc, double scale,
                  const TetBlock* __restrict__ tets, const double* __restrict__ states, int cap,
                  int32_t* __restrict__ over) {
    __shared__ int32_t hk[kHashSlots];
    __shared__ unsigned long long hw[6 * kHashSlots];  // per slot and axis: hi, lo
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < n;
    double ax = 0, ay = 0, az = 0, vx = 0, vy = 0, vz = 0;
    if (live) {
        ax = lx[3 * i]; ay = lx[3 * i + 1]; az = lx[3 * i + 2];
        vx = lv[3 * i]; vy = lv[3 * i + 1]; vz = lv[3 * i + 2];
    }
    const bool any = live && (ax != 0.0 || ay != 0.0 || az != 0.0);
    for (int q = threadIdx.x; q < kHashSlots; q += blockDim.x) {
        hk[q] = -1;
        for (int w = 0; w < 6; w++) hw[6 * q + w] = 0ull;
    }
    __syncthreads();
    for (int32_t k = k1 - 1; k >= k0; k--) {
        const int32_t c = any ? path[(int64_t)(k - k0) * n + i] : -1;
        if (WALL && c >= 0)
            unreflect(tets, states + (int64_t)(k - k0) * 6 * n, n, i, c, qm, dt, cap, ax, ay, az, vx, vy, vz, over);
        vx += ax * dt;
        vy += ay * dt;
        vz += az * dt;
        if (any) {
            if (c >= 0)
                lds_insert3_fixed(hk, hw, c, to_fixed((qm * dt) * vx, scale), to_fixed((qm * dt) * vy, scale),
                                  to_fixed((qm * dt) * vz, scale));
        }
        __syncthreads();
        for (int q = threadIdx.x; q < kHashSlots; q += blockDim.x) {
            const int32_t key = hk[q];
            if (key >= 0) {
                for (int d = 0; d < 3; d++) {
                    unsigned long long* g = acc + 2 * (3 * (int64_t)key + d);
                    add_fixed(g + 1, g + 0, Fixed{hw[6 * q + 2 * d], hw[6 * q + 2 * d + 1]});
                    hw[6 * q + 2 * d] = hw[6 * q + 2 * d + 1] = 0ull;
                }
                hk[q] = -1;
            }
        }
        __syncthreads();
    }
    if (live) {
        lv[3 * i] = vx; lv[3 * i + 1] = vy; lv[3 * i + 2] = vz;
        if (WALL) {
            lx[3 * i] = ax; lx[3 * i + 1] = ay; lx[3 * i + 2] = az;
        }
    }
}

// Reverse sweep of steps [k0, k1) of one segment, path[(k - k0) * n + i] = c_k of particle i.
template <int WALL>
__global__ void __launch_bounds__(256)
k_adj_sweep(const int32_t* __restrict__ path, int32_t k0, int32_t k1, int64_t n, double* __restrict__ lx,
            double* __restrict__ lv, double qm, double dt, double* __restrict__ dLdE, const TetBlock* __restrict__ tets,
            const double* __restrict__ states, int cap, int32_t* __restrict__ over) {
    __shared__ int32_t hk[kHashSlots];
    __shared__ double hv[3 * kHashSlots];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < n;
    double ax = 0, ay = 0, az = 0, vx = 0, vy = 0, vz = 0;
    if (live) {
        ax = lx[3 * i]; ay = lx[3 * i + 1]; az = lx[3 * i + 2];
        vx = lv[3 * i]; vy = lv[3 * i + 1]; vz = lv[3 * i + 2];
    }
    const bool any = live && (ax != 0.0 || ay != 0.0 || az != 0.0);
    for (int s = threadIdx.x; s < kHashSlots; s += blockDim.x) {
        hk[s] = -1;
        hv[3 * s] = hv[3 * s + 1] = hv[3 * s + 2] = 0.0;
    }
    __syncthreads();
    for (int32_t k = k1 - 1; k >= k0; k--) {
        const int32_t c = any ? path[(int64_t)(k - k0) * n + i] : -1;
        if (WALL && c >= 0)
            unreflect(tets, states + (int64_t)(k - k0) * 6 * n, n, i, c, qm, dt, cap, ax, ay, az, vx, vy, vz, over);
        vx += ax * dt;
        vy += ay * dt;
        vz += az * dt;
        if (any && c >= 0) lds_insert3(hk, hv, c, (qm * dt) * vx, (qm * dt) * vy, (qm * dt) * vz);
        __syncthreads();
        for (int s = threadIdx.x; s < kHashSlots; s += blockDim.x) {
            const int32_t key = hk[s];
            if (key >= 0) {
                unsafeAtomicAdd(dLdE + 3 * (int64_t)key + 0, hv[3 * s + 0]);
                unsafeAtomicAdd(dLdE + 3 * (int64_t)key + 1, hv[3 * s + 1]);
                unsafeAtomicAdd(dLdE + 3 * (int64_t)key + 2, hv[3 * s + 2]);
                hk[s] = -1;
                hv[3 * s] = hv[3 * s + 1] = hv[3 * s + 2] = 0.0;
            }
        }
        __syncthreads();
    }
    if (live) {
        lv[3 * i] = vx; lv[3 * i + 1] = vy; lv[3 * i + 2] = vz;
        if (WALL) {
            lx[3 * i] = ax; lx[3 * i + 1] = ay; lx[3 * i + 2] = az;
        }
    }
}

// Per-block partial sums of lambda_v . dir.
__global__ void __launch_bounds__(256)
k_adj_drift(const double* __restrict__ lv, int64_t n, double d0, double d1, double d2, double* __restrict__ part) {
    __shared__ double sh[256 / kWave];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (i < n) v = (lv[3 * i] * d0 + lv[3 * i + 1] * d1) + lv[3 * i + 2] * d2;
    v = wave_sum_f64(v);
    if (lane_id() == 0) sh[threadIdx.x / kWave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < (int)(blockDim.x / kWave); w++) s += sh[w];
        part[blockIdx.x] = s;
    }
}

// The same per particle, lambda_v . dir in 128-bit fixed point (scale 2^S), summed exactly per
// block: part[2 * block] = hi word, part[2 * block + 1] = lo word.
__global__ void __launch_bounds__(256)
k_adj_drift_fixed(const double* __restrict__ lv, int64_t n, double d0, double d1, double d2, double scale,
                  unsigned long long* __restrict__ part) {
    __shared__ unsigned long long sw[2];
    if (threadIdx.x == 0) sw[0] = sw[1] = 0ull;
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const double v = (lv[3 * i] * d0 + lv[3 * i + 1] * d1) + lv[3 * i + 2] * d2;
        add_fixed(&sw[1], &sw[0], to_fixed(v, scale));
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * (int64_t)blockIdx.x] = sw[0];
        part[2 * (int64_t)blockIdx.x + 1] = sw[1];
    }
}

static inline unsigned g256(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace mpd

using namespace mpctx;

namespace {

struct AdjBuffers {
    std::vector<void*> owned;
    template <typename T>
    T* get(int64_t n) {
        T* p = nullptr;
        if (dalloc(&p, n) != hipSuccess) return nullptr;
        owned.push_back(p);
        return p;
    }
    ~AdjBuffers() {
        for (void* p : owned) dfree(p);
    }
};

hipError_t copy_state(const Soa& src, const Soa& dst, int64_t n, hipStream_t st) {
    const size_t nd = sizeof(double) * (size_t)n, ni = sizeof(int32_t) * (size_t)n;
    const double* s6[6] = {src.x, src.y, src.z, src.vx, src.vy, src.vz};
    double* d6[6] = {dst.x, dst.y, dst.z, dst.vx, dst.vy, dst.vz};
    for (int k = 0; k < 6; k++) {
        hipError_t e = hipMemcpyAsync(d6[k], s6[k], nd, hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) return e;
    }
    return hipMemcpyAsync(dst.cell, src.cell, ni, hipMemcpyDeviceToDevice, st);
}

}  // namespace

extern "C" int mp_adjoint_ex(mp_ctx* c, const double* pos0, const double* vel0, const int32_t* cell0, int64_t n,
                             double q_over_m, double dt, double spwt, int32_t steps, const double* den_target,
                             const double drift_dir[3], int32_t checkpoint_every, double* loss, double* dL_dE,
                             double* dL_ddrift, int32_t mode) {
    if (!c) return MP_ERR_ARG;
    if (steps < 1 || checkpoint_every < 1 || !den_target || !drift_dir || !loss || !dL_dE || !dL_ddrift)
        return set_err(c, MP_ERR_ARG, "mp_adjoint: bad arguments");
    if (mode != MP_SCATTER_ATOMIC && mode != MP_SCATTER_DETERMINISTIC)
        return set_err(c, MP_ERR_ARG, "mp_adjoint: unknown scatter mode");
    const bool det = mode == MP_SCATTER_DETERMINISTIC;
    const bool wall = c->opt.wall_mode == MP_WALL_REFLECT;
    int rc = mp_set_particles(c, pos0, vel0, cell0, n);
    if (rc) return rc;
    // one cell sort for locality (pids are irrelevant to the summed gradients)
    {
        const int32_t keep = c->opt.sort_interval, keep_order = c->opt.keep_pid_order;
        c->opt.sort_interval = 1;
        c->opt.keep_pid_order = 0;
        rc = mp_sort_now(c);
        c->opt.sort_interval = keep;
        c->opt.keep_pid_order = keep_order;
        if (rc) return rc;
    }
    const int64_t m = c->n_slots;
    const int32_t nn = c->n_nodes, T = c->n_tets;
    const int32_t C = std::min(checkpoint_every, steps);
    const int32_t nseg = (steps + C - 1) / C;
    hipStream_t st = c->stream;
    AdjBuffers buf;
    std::vector<void*> ck_mem((size_t)nseg, nullptr);
    std::vector<Soa> cks((size_t)nseg);
    for (int32_t k = 0; k < nseg; k++) {
        if (alloc_soa(&ck_mem[k], &cks[k], std::max<int64_t>(m, 1)) != hipSuccess) {
            for (void* p : ck_mem) dfree(p);
            return set_err(c, MP_ERR_ALLOC, "mp_adjoint: checkpoint allocation failed");
        }
        buf.owned.push_back(ck_mem[k]);
    }
    const int64_t mm = std::max<int64_t>(m, 1);
    int32_t* path = buf.get<int32_t>((int64_t)C * mm);
    double* lx = buf.get<double>(3 * mm);
    double* lv = buf.get<double>(3 * mm);
    double* den = buf.get<double>(nn);
    double* tgt = buf.get<double>(nn);
    double* g = buf.get<double>(nn);
    double* dE = buf.get<double>(3 * (int64_t)T);
    const int64_t nb_loss = mpd::g256(nn), nb_part = mpd::g256(mm);
    double* part = buf.get<double>(std::max(nb_loss, nb_part));
    // deterministic mode: the 128-bit fixed-point dL/dE sums (hi, lo per component)
    unsigned long long* acc = det ? buf.get<unsigned long long>(6 * (int64_t)T) : nullptr;
    // reflecting walls: each replayed step's start state (6 x m doubles per step of a segment)
    // and the flag for a step that reflected a particle more often than the sweep records
    double* states = wall ? buf.get<double>((int64_t)C * 6 * mm) : nullptr;
    int32_t* over = buf.get<int32_t>(1);
    if (!path || !lx || !lv || !den || !tgt || !g || !dE || !part || (det && !acc) || (wall && !states) || !over)
        return set_err(c, MP_ERR_ALLOC, "mp_adjoint: workspace allocation failed");
    // deterministic mode: records of the last forward step at 4 * pid + v
    const int64_t n_rec = 4 * c->next_pid;
    if (det) {
        rc = reserve_records(c, n_rec, bit_width((uint64_t)nn));
        if (rc) return rc;
    }
    HIP_TRY(c, hipMemcpyAsync(tgt, den_target, sizeof(double) * (size_t)nn, hipMemcpyHostToDevice, st));
    HIP_TRY(c, mpd::launch_fill_f64(den, nn, 0.0, st));
    HIP_TRY(c, mpd::launch_fill_f64(dE, 3 * (int64_t)T, 0.0, st));
    HIP_TRY(c, mpd::launch_fill_i32(over, 1, 0, st));
    const Soa& S = c->soa[c->cur];
    const Soa& W = c->soa[1 - c->cur];
    const int cap = c->opt.walk_cap;
    const unsigned gm = mpd::g256(mm);
    // forward with checkpoints
    for (int32_t k = 0; k < steps; k++) {
        if (k % C == 0) HIP_TRY(c, copy_state(S, cks[k / C], m, st));
        if (m > 0) {
            const int sc = k < steps - 1 ? 0 : det ? 2 : 1;
            if (sc == 2) HIP_TRY(c, mpd::launch_fill_u32(c->keys[0], n_rec, 0xFFFFFFFFu, st));
#define MP_ADJ_FWD(SC, W)                                                                                     \
    hipLaunchKernelGGL((mpd::k_adj_step<SC, W>), dim3(gm), dim3(256), 0, st, c->tets, S, m, q_over_m, dt, cap, \
                       (int32_t*)nullptr, (double*)nullptr, sc == 1 ? den : (double*)nullptr,                  \
                       sc == 2 ? c->keys[0] : nullptr, sc == 2 ? c->vals[0] : nullptr)
            if (wall) {
                if (sc == 0) MP_ADJ_FWD(0, 1);
                else if (sc == 1) MP_ADJ_FWD(1, 1);
                else MP_ADJ_FWD(2, 1);
            } else {
                if (sc == 0) MP_ADJ_FWD(0, 0);
                else if (sc == 1) MP_ADJ_FWD(1, 0);
                else MP_ADJ_FWD(2, 0);
            }
#undef MP_ADJ_FWD
            HIP_TRY(c, hipGetLastError());
        }
    }
    double* den_k = den;
    if (det) {  // the ascending-pid fold of the last step's records, normalised
        if (m > 0) {
            rc = det_finish(c, n_rec, spwt);
            if (rc) return rc;
            den_k = c->den_out;
        } else {
            HIP_TRY(c, mpd::launch_fill_f64(c->den_out, nn, 0.0, st));
            den_k = c->den_out;
        }
    }
    hipLaunchKernelGGL(mpd::k_adj_loss, dim3((unsigned)nb_loss), dim3(256), 0, st, den_k, tgt, c->node_vol, nn, spwt, g,
                       part, det);
    HIP_TRY(c, hipGetLastError());
    std::vector<double> hp((size_t)std::max(nb_loss, nb_part));
    HIP_TRY(c, hipMemcpyAsync(hp.data(), part, sizeof(double) * (size_t)nb_loss, hipMemcpyDeviceToHost, st));
    HIP_TRY(c, hipStreamSynchronize(st));
    double L = 0.0;
    for (int64_t b = 0; b < nb_loss; b++) L += hp[(size_t)b];
    // backward: lambda_x from the final state, then segments in reverse
    double scale = 1.0, lmax_x = 0.0;
    if (m > 0) {
        hipLaunchKernelGGL(mpd::k_adj_lambda, dim3(gm), dim3(256), 0, st, c->tets, S, m, g, lx, lv);
        HIP_TRY(c, hipGetLastError());
    }
    if (det && m > 0) {
        // the fixed-point scale 2^S from the bound on every contribution
        hipLaunchKernelGGL(mpd::k_adj_lmax, dim3((unsigned)nb_part), dim3(256), 0, st, lx, m, part);
        HIP_TRY(c, hipGetLastError());
        HIP_TRY(c, hipMemcpyAsync(hp.data(), part, sizeof(double) * (size_t)nb_part, hipMemcpyDeviceToHost, st));
        HIP_TRY(c, mpd::launch_fill_f64((double*)acc, 6 * (int64_t)T, 0.0, st));  // all-zero words
        HIP_TRY(c, hipStreamSynchronize(st));
        for (int64_t b = 0; b < nb_part; b++) lmax_x = std::max(lmax_x, hp[(size_t)b]);
        const double bound = std::fabs(q_over_m) * dt * ((double)steps * dt) * 1.7320508075688772 * lmax_x *
                             (double)m * (double)steps;
        if (bound > 0.0 && std::isfinite(bound)) {
            int e = 0;
            (void)std::frexp(bound, &e);  // bound < 2^e
            scale = std::ldexp(1.0, std::max(-1000, std::min(1000, 62 - e)));
        }
    }
    for (int32_t sgi = nseg - 1; sgi >= 0 && m > 0; sgi--) {
        const int32_t k0 = sgi * C, k1 = std::min(steps, k0 + C);
        HIP_TRY(c, copy_state(cks[sgi], W, m, st));
        for (int32_t k = k0; k < k1; k++) {
            int32_t* pk = path + (int64_t)(k - k0) * m;
            if (wall)
                hipLaunchKernelGGL((mpd::k_adj_step<0, 1>), dim3(gm), dim3(256), 0, st, c->tets, W, m, q_over_m, dt,
                                   cap, pk, states + (int64_t)(k - k0) * 6 * m, (double*)nullptr, nullptr, nullptr);
            else
                hipLaunchKernelGGL((mpd::k_adj_step<0, 0>), dim3(gm), dim3(256), 0, st, c->tets, W, m, q_over_m, dt,
                                   cap, pk, (double*)nullptr, (double*)nullptr, nullptr, nullptr);
            HIP_TRY(c, hipGetLastError());
        }
#define MP_ADJ_SWEEP(W)                                                                                        \
    do {                                                                                                       \
        if (det)                                                                                               \
            hipLaunchKernelGGL((mpd::k_adj_sweep_fixed<W>), dim3(gm), dim3(256), 0, st, path, k0, k1, m, lx, lv, \
                               q_over_m, dt, acc, scale, c->tets, states, cap, over);                          \
        else                                                                                                   \
            hipLaunchKernelGGL((mpd::k_adj_sweep<W>), dim3(gm), dim3(256), 0, st, path, k0, k1, m, lx, lv,       \
                               q_over_m, dt, dE, c->tets, states, cap, over);                                  \
    } while (0)
        if (wall) MP_ADJ_SWEEP(1);
        else MP_ADJ_SWEEP(0);
#undef MP_ADJ_SWEEP
        HIP_TRY(c, hipGetLastError());
    }
    double D = 0.0;
    // deterministic mode: lambda_v . dir in fixed point with a bound of its own,
    // |lambda_v . dir| <= |lambda_v| |dir| <= sqrt(3) K dt sqrt(3) max|lambda_x| |dir| per particle
    double dscale = 1.0;
    std::vector<unsigned long long> hpf;
    if (det && m > 0) {
        const double dn = std::sqrt((drift_dir[0] * drift_dir[0] + drift_dir[1] * drift_dir[1]) +
                                    drift_dir[2] * drift_dir[2]);
        const double bound = 3.0 * ((double)steps * dt) * lmax_x * dn * (double)m;
        if (bound > 0.0 && std::isfinite(bound)) {
            int e = 0;
            (void)std::frexp(bound, &e);
            dscale = std::ldexp(1.0, std::max(-1000, std::min(1000, 62 - e)));
        }
        unsigned long long* pf = buf.get<unsigned long long>(2 * nb_part);
        if (!pf) return set_err(c, MP_ERR_ALLOC, "mp_adjoint: workspace allocation failed");
        hipLaunchKernelGGL(mpd::k_adj_drift_fixed, dim3((unsigned)nb_part), dim3(256), 0, st, lv, m, drift_dir[0],
                           drift_dir[1], drift_dir[2], dscale, pf);
        HIP_TRY(c, hipGetLastError());
        hpf.resize(2 * (size_t)nb_part);
        HIP_TRY(c, hipMemcpyAsync(hpf.data(), pf, sizeof(unsigned long long) * hpf.size(), hipMemcpyDeviceToHost, st));
    } else if (m > 0) {
        hipLaunchKernelGGL(mpd::k_adj_drift, dim3((unsigned)nb_part), dim3(256), 0, st, lv, m, drift_dir[0], drift_dir[1],
                           drift_dir[2], part);
        HIP_TRY(c, hipGetLastError());
        HIP_TRY(c, hipMemcpyAsync(hp.data(), part, sizeof(double) * (size_t)nb_part, hipMemcpyDeviceToHost, st));
    }
    if (det) {
        hipLaunchKernelGGL(mpd::k_adj_unfix, dim3(mpd::g256(3 * (int64_t)T)), dim3(256), 0, st, acc, 3 * (int64_t)T,
                           1.0 / scale, dE);
        HIP_TRY(c, hipGetLastError());
    }
    HIP_TRY(c, hipMemcpyAsync(dL_dE, dE, sizeof(double) * 3 * (size_t)T, hipMemcpyDeviceToHost, st));
    int32_t h_over = 0;
    HIP_TRY(c, hipMemcpyAsync(&h_over, over, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIP_TRY(c, hipStreamSynchronize(st));
    if (h_over)
        return set_err(c, MP_ERR_STATE, "mp_adjoint: a step reflected one particle more than " +
                                             std::to_string(mpd::kMaxRefl) + " times");
    if (det && m > 0) {
        // exact 128-bit sum of the block partials, rounded to double once
        unsigned __int128 sum = 0;
        for (int64_t b = 0; b < nb_part; b++)
            sum += ((unsigned __int128)hpf[2 * (size_t)b] << 64) + (unsigned __int128)hpf[2 * (size_t)b + 1];
        D = (double)(__int128)sum * 5.421010862427522e-20 / dscale;  // x 2^-64 x 2^-S
    } else if (m > 0) {
        for (int64_t b = 0; b < nb_part; b++) D += hp[(size_t)b];
    }
    *loss = L;
    *dL_ddrift = D;
    return MP_OK;
}

extern "C" int mp_adjoint(mp_ctx* c, const double* pos0, const double* vel0, const int32_t* cell0, int64_t n,
                          double q_over_m, double dt, double spwt, int32_t steps, const double* den_target,
                          const double drift_dir[3], int32_t checkpoint_every, double* loss, double* dL_dE,
                          double* dL_ddrift) {
    return mp_adjoint_ex(c, pos0, vel0, cell0, n, q_over_m, dt, spwt, steps, den_target, drift_dir, checkpoint_every,
                         loss, dL_dE, dL_ddrift, MP_SCATTER_ATOMIC);
}
