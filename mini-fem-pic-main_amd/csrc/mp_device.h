// Device-side data layout and the per-particle arithmetic shared by every kernel.
//
// HBM layout of the tet table.  A tet's record is 12 chunks of 16 B:
//   chunks 0-7  alpha_i, beta_i | gamma_i, delta_i of vertex i (two chunks per vertex)
//   chunk  8    vol, ef[0]
//   chunk  9    ef[1], ef[2]
//   chunk 10    nbr[4] (= Tetra::cell_con)
//   chunk 11    con[4]
// Records are grouped in blocks of kTetBlock consecutive tets, chunk-major inside a block:
// chunk k of tet j of the block at k * 16 * kTetBlock + j * 16.  The default, one tet per
// block, is one 192 B record per tet: what the walk, the push (start-cell E) and the scatter
// (node ids) need for one tet in one or two 128 B lines.  Blocks of 8 tets (a cell-sorted
// wave's gather instruction then touches one line instead of one per lane) measured no faster
// for the single-step push at 20M particles on 60^3 and 7% slower for the multi-step kernel,
// which reloads whole records at random as particles move: the push waits on the latency of
// its dependent gathers, not on the lookups.
//
// Arithmetic follows the reference kernel with no contraction (built -ffp-contract=off):
//   push   src/mini-pic.cpp:291-295   v = v + (qm*E)*dt ; x = x + v*dt
//   walk   src/meshes.h:75-100        lc_i = ((1/6)*(((a_i - x b_i) + y g_i) - z d_i)) / V
//                                     inside <=> no lc_i < 0 and no lc_i > 1
//                                     move across the face of the first minimum lc
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpd {

#ifndef MP_TET_BLOCK
#define MP_TET_BLOCK 1
#endif
constexpr int kTetBlock = MP_TET_BLOCK;
static_assert(kTetBlock >= 1 && (kTetBlock & (kTetBlock - 1)) == 0, "tets per block: a power of two");

struct alignas(16) TetChunk {
    union {
        double d[2];
        int32_t i[4];
    };
};

struct TetBlock {
    TetChunk ch[12][kTetBlock];  // ch[k][j]: chunk k of tet j of the block
};
static_assert(sizeof(TetBlock) == 192 * kTetBlock, "192 B per tet");

constexpr int kChVolEx = 8, kChEyEz = 9, kChNbr = 10, kChCon = 11;

__host__ __device__ inline int64_t tet_blocks(int64_t n_tets) { return (n_tets + kTetBlock - 1) / kTetBlock; }

// One tet's record in the table (read side): c(v, f) is coefficient f (alpha, beta, gamma,
// delta) of vertex v.
struct TetView {
    const TetChunk* q;  // chunk 0; chunk k at q[k * kTetBlock]
    __device__ __forceinline__ double c(int v, int f) const { return q[(2 * v + (f >> 1)) * kTetBlock].d[f & 1]; }
    __device__ __forceinline__ double vol() const { return q[kChVolEx * kTetBlock].d[0]; }
    __device__ __forceinline__ double ex() const { return q[kChVolEx * kTetBlock].d[1]; }
    __device__ __forceinline__ double ey() const { return q[kChEyEz * kTetBlock].d[0]; }
    __device__ __forceinline__ double ez() const { return q[kChEyEz * kTetBlock].d[1]; }
    __device__ __forceinline__ int32_t nbr(int v) const { return q[kChNbr * kTetBlock].i[v]; }
    __device__ __forceinline__ int32_t con(int v) const { return q[kChCon * kTetBlock].i[v]; }
};

// Tet e's chunk 0, writable (the table's builder and the field upload); chunk k at [k * kTetBlock].
__host__ __device__ inline TetChunk* tet_chunks(TetBlock* tets, int64_t e) {
    return &tets[e / kTetBlock].ch[0][e % kTetBlock];
}

// The record of tet c >= 0.
__device__ __forceinline__ TetView tet_at(const TetBlock* __restrict__ tets, int32_t c) {
    const uint32_t u = (uint32_t)c;
    return TetView{&tets[u / kTetBlock].ch[0][u % kTetBlock]};
}

// Structure-of-arrays particle state in HBM.
struct Soa {
    double* x;
    double* y;
    double* z;
    double* vx;
    double* vy;
    double* vz;
    int32_t* cell;
    int32_t* pid;
    uint16_t* run;  // batch: run of each particle (nullptr for a single run)
};

// The slab segments a det_algo 2 push places its records into (mp_common.h, place_records).  A
// node owns one slab, or 2^b slabs cut from its pid range when it is hot (sinfo).
struct PlaceArgs {
    const int32_t* base;      // slabs + 1: segment starts (capacity of s: base[s + 1] - base[s])
    int32_t* cur;             // slabs: claim cursors, zero before the push
    int32_t* flags;           // [0]: a record did not fit
    const int2* sinfo;        // nn: (first slab, b) per node; nullptr: no hot node, slab n = node n
    unsigned long long mul;   // pid digit multiplier, floor(2^48 / next pid)
};

// Words (4 B) per placed record (mp_common.h, PlacedRec).
constexpr int kPlacedWords = 4;

enum WalkStatus : int { kInside = 0, kLost = 1, kCapped = 2 };

// Most steps one multi-step launch takes (its 2048-slot LDS hash holds ~40 keys per step and
// workgroup after a cell sort; keys beyond its capacity go to HBM directly).
constexpr int kStepMaxLaunch = 32;

__device__ __forceinline__ double bary(double a, double b, double g, double d, double x, double y,
                                       double z, double vol) {
    return ((1.0 / 6.0) * (((a - x * b) + y * g) - z * d)) / vol;
}

struct WalkStats {
    int moves;
    int refl;
};

// No reflection recorder (walk's default).
struct NoRefl {
    __device__ __forceinline__ void operator()(double, double, double) const {}
};

// XtoLtet (src/meshes.h:75-100) from cell c.  On return c is the last tet visited and
// l0..l3 its barycentric weights.  WALL == 1 reflects at boundary faces instead of
// reporting the particle lost; a reflection counts as one move toward cap, and rec(g) is
// called with each reflection's plane normal g, in order.  The tet's fields are read where
// they are used rather than staged in registers, which keeps the loop's live registers few.
template <int WALL, class Rec = NoRefl>
__device__ __forceinline__ int walk(const TetBlock* __restrict__ tets, int32_t& c, double& x,
                                    double& y, double& z, double& vx, double& vy, double& vz,
                                    double& l0, double& l1, double& l2, double& l3, int cap,
                                    WalkStats& st, Rec&& rec = Rec{}) {
    while (true) {
        const TetView t = tet_at(tets, c);
        const double vol = t.vol();
        l0 = bary(t.c(0, 0), t.c(0, 1), t.c(0, 2), t.c(0, 3), x, y, z, vol);
        l1 = bary(t.c(1, 0), t.c(1, 1), t.c(1, 2), t.c(1, 3), x, y, z, vol);
        l2 = bary(t.c(2, 0), t.c(2, 1), t.c(2, 2), t.c(2, 3), x, y, z, vol);
        l3 = bary(t.c(3, 0), t.c(3, 1), t.c(3, 2), t.c(3, 3), x, y, z, vol);
        const bool out = (l0 < 0 || l0 > 1.0) || (l1 < 0 || l1 > 1.0) || (l2 < 0 || l2 > 1.0) ||
                         (l3 < 0 || l3 > 1.0);
        if (!out) return kInside;
        int mi = 0;
        double ml = l0;
        if (l1 < ml) { ml = l1; mi = 1; }
        if (l2 < ml) { ml = l2; mi = 2; }
        if (l3 < ml) { ml = l3; mi = 3; }
        const int32_t nb = t.nbr(mi);
        if (st.moves >= cap) return kCapped;
        st.moves++;
        if (nb >= 0) {
            c = nb;
            continue;
        }
        if (WALL == 0) return kLost;
        // Specular reflection about the plane lc_mi = 0.  grad lc_mi is proportional to
        // g = (-beta, gamma, -delta) / (6 V); x' = x - 2 lc g (6V)/|g|^2, v' = v - 2 (v.g) g/|g|^2.
        const double gx = -t.c(mi, 1), gy = t.c(mi, 2), gz = -t.c(mi, 3);
        rec(gx, gy, gz);
        const double gg = (gx * gx + gy * gy) + gz * gz;
        const double s = ((2.0 * ml) * (6.0 * vol)) / gg;
        x = x - s * gx;
        y = y - s * gy;
        z = z - s * gz;
        const double w = (2.0 * ((vx * gx + vy * gy) + vz * gz)) / gg;
        vx = vx - w * gx;
        vy = vy - w * gy;
        vz = vz - w * gz;
        st.refl++;
    }
}

// The start cell's fields, loaded in one batch with the particle state (they share the tet
// record's cache lines with the field the push needs), so the first walk iteration does not
// wait for a second round trip.  Plain scalars: no dynamically indexed register arrays.
struct StartTet {
    double a0, b0, g0, d0, a1, b1, g1, d1, a2, b2, g2, d2, a3, b3, g3, d3, vol;
    int32_t n0, n1, n2, n3;
};

__device__ __forceinline__ void load_start(const TetView t, StartTet& s) {
    s.a0 = t.c(0, 0); s.b0 = t.c(0, 1); s.g0 = t.c(0, 2); s.d0 = t.c(0, 3);
    s.a1 = t.c(1, 0); s.b1 = t.c(1, 1); s.g1 = t.c(1, 2); s.d1 = t.c(1, 3);
    s.a2 = t.c(2, 0); s.b2 = t.c(2, 1); s.g2 = t.c(2, 2); s.d2 = t.c(2, 3);
    s.a3 = t.c(3, 0); s.b3 = t.c(3, 1); s.g3 = t.c(3, 2); s.d3 = t.c(3, 3);
    s.vol = t.vol();
    s.n0 = t.nbr(0); s.n1 = t.nbr(1); s.n2 = t.nbr(2); s.n3 = t.nbr(3);
}

// The first iteration of XtoLtet on the start tet's fields s (already in registers).  Returns
// kInside, kLost, kCapped (cap <= 0 and outside: a caller's way to evaluate the start tet
// only), or kContinue: the walk goes on in walk() from c with st.moves counted so far (the
// neighbour with one move, or the same tet for a reflecting boundary face, which walk()
// re-evaluates and reflects).
constexpr int kContinue = 3;

template <int WALL>
__device__ __forceinline__ int first_step(const StartTet& s, int32_t& c, double x, double y, double z,
                                          double& l0, double& l1, double& l2, double& l3, int cap,
                                          WalkStats& st) {
    l0 = bary(s.a0, s.b0, s.g0, s.d0, x, y, z, s.vol);
    l1 = bary(s.a1, s.b1, s.g1, s.d1, x, y, z, s.vol);
    l2 = bary(s.a2, s.b2, s.g2, s.d2, x, y, z, s.vol);
    l3 = bary(s.a3, s.b3, s.g3, s.d3, x, y, z, s.vol);
    const bool out = (l0 < 0 || l0 > 1.0) || (l1 < 0 || l1 > 1.0) || (l2 < 0 || l2 > 1.0) ||
                     (l3 < 0 || l3 > 1.0);
    if (!out) return kInside;
    if (cap <= 0) return kCapped;
    int mi = 0;
    double ml = l0;
    if (l1 < ml) { ml = l1; mi = 1; }
    if (l2 < ml) { ml = l2; mi = 2; }
    if (l3 < ml) { ml = l3; mi = 3; }
    const int32_t nb = mi == 0 ? s.n0 : mi == 1 ? s.n1 : mi == 2 ? s.n2 : s.n3;
    if (nb >= 0) {  // cap >= 1 here, so the first move is never capped
        st.moves = 1;
        c = nb;
        return kContinue;
    }
    if (WALL == 0) {
        st.moves = 1;
        return kLost;
    }
    return kContinue;
}

// XtoLtet from cell c whose start-tet fields s are already in registers.
template <int WALL>
__device__ __forceinline__ int walk_start(const TetBlock* __restrict__ tets, const StartTet& s, int32_t& c,
                                          double& x, double& y, double& z, double& vx, double& vy,
                                          double& vz, double& l0, double& l1, double& l2, double& l3,
                                          int cap, WalkStats& st) {
    const int r = first_step<WALL>(s, c, x, y, z, l0, l1, l2, l3, cap, st);
    if (r != kContinue) return r;
    return walk<WALL>(tets, c, x, y, z, vx, vy, vz, l0, l1, l2, l3, cap, st);
}

// A tet's whole record in registers: what the next step's push (E), walk (coefficients,
// neighbours) and scatter (node ids) read, for kernels that keep a particle across steps.
struct TetRegs {
    StartTet g;
    double ex, ey, ez;
    int32_t m0, m1, m2, m3;
};

__device__ __forceinline__ void load_tet_regs(const TetView t, TetRegs& r) {
    load_start(t, r.g);
    r.ex = t.ex();
    r.ey = t.ey();
    r.ez = t.ez();
    r.m0 = t.con(0);
    r.m1 = t.con(1);
    r.m2 = t.con(2);
    r.m3 = t.con(3);
}

}  // namespace mpd
