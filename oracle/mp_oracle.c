/*
 * mp_oracle.c - CPU oracle of MoveParticles (test infrastructure only; see mp_oracle.h).
 * Build: -O3 -ffp-contract=off, never -ffast-math.
 */
#include "mp_oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { OR_INSIDE = 0, OR_LOST = 1, OR_CAPPED = 2 };

/* XtoLtet, src/meshes.h:75-100, plus this build's cap and reflecting-wall extension:
 * the cap counts face moves (and reflections); at the cap the particle is lost.  A
 * reflection mirrors position and velocity about the plane lc_min = 0, whose inward
 * normal is g = (-beta, gamma, -delta) of that vertex. */
static int or_walk_rec(const or_tet* tets, int32_t* cell, double* x, double* v, double lc[4], int wall,
                       int cap, int* moves, int* refl, double* rec, int rec_max, int* n_rec) {
    for (;;) {
        const or_tet* t = &tets[*cell];
        int inside = 1;
        for (int i = 0; i < 4; i++) {
            lc[i] = (1.0 / 6.0) * (t->alpha[i] - x[0] * t->beta[i] + x[1] * t->gamma[i] - x[2] * t->delta[i]) /
                    t->volume;
            if (lc[i] < 0 || lc[i] > 1.0) inside = 0;
        }
        if (inside) return OR_INSIDE;
        int min_i = 0;
        double min_lc = lc[0];
        for (int i = 1; i < 4; i++)
            if (lc[i] < min_lc) {
                min_lc = lc[i];
                min_i = i;
            }
        if (*moves >= cap) return OR_CAPPED;
        (*moves)++;
        if (t->cell_con[min_i] >= 0) {
            *cell = t->cell_con[min_i];
            continue;
        }
        if (wall == 0) return OR_LOST;
        const double gx = -t->beta[min_i], gy = t->gamma[min_i], gz = -t->delta[min_i];
        const double gg = (gx * gx + gy * gy) + gz * gz;
        const double s = ((2.0 * min_lc) * (6.0 * t->volume)) / gg;
        x[0] = x[0] - s * gx;
        x[1] = x[1] - s * gy;
        x[2] = x[2] - s * gz;
        const double w = (2.0 * ((v[0] * gx + v[1] * gy) + v[2] * gz)) / gg;
        v[0] = v[0] - w * gx;
        v[1] = v[1] - w * gy;
        v[2] = v[2] - w * gz;
        if (rec) { /* the reflection's plane normal, in order (the adjoint replays them backwards) */
            if (*n_rec >= rec_max) return -1;
            rec[3 * *n_rec + 0] = gx;
            rec[3 * *n_rec + 1] = gy;
            rec[3 * *n_rec + 2] = gz;
            (*n_rec)++;
        }
        (*refl)++;
    }
}

static int or_walk(const or_tet* tets, int32_t* cell, double* x, double* v, double lc[4], int wall,
                   int cap, int* moves, int* refl) {
    return or_walk_rec(tets, cell, x, v, lc, wall, cap, moves, refl, NULL, 0, NULL);
}

static void add_stats(int64_t* stats, int lost, int crossed, int moves, int capped, int refl) {
    if (!stats) return;
    stats[0] += 1;
    stats[1] += lost;
    stats[2] += crossed;
    stats[3] += moves;
    stats[4] += capped;
    stats[5] += refl;
}

int64_t or_move_aos(const or_tet* tets, const double* ef, or_particle* parts, int64_t n, double qm,
                    double dt, double* den, int32_t walk_cap, int64_t* stats) {
    for (int64_t p = 0; p < n; p++) {
        or_particle* part = &parts[p];
        /* src/mini-pic.cpp:291-295 */
        for (int i = 0; i < 3; i++) part->vel[i] += qm * ef[3 * (int64_t)part->cell_index + i] * dt;
        for (int i = 0; i < 3; i++) part->pos[i] += part->vel[i] * dt;
        int32_t c = part->cell_index;
        int moves = 0, refl = 0;
        const int r = or_walk(tets, &c, part->pos, part->vel, part->lc, 0, walk_cap, &moves, &refl);
        if (r == OR_INSIDE) {
            part->cell_index = c;
            const or_tet* t = &tets[c];
            for (int v = 0; v < 4; v++) den[t->con[v]] += part->lc[v];
        } else {
            part->cell_index = -1;
        }
        add_stats(stats, r != OR_INSIDE, moves > 0, moves, r == OR_CAPPED, refl);
    }
    return n;
}

int64_t or_move(const or_tet* tets, const double* ef, double* pos, double* vel, int32_t* cell,
                double* lc_out, int64_t n, double qm, double dt, double* den, int32_t wall_mode,
                int32_t walk_cap, int64_t* stats) {
    int64_t done = 0;
    for (int64_t p = 0; p < n; p++) {
        int32_t c = cell[p];
        if (c < 0) continue;
        double* x = pos + 3 * p;
        double* v = vel + 3 * p;
        for (int i = 0; i < 3; i++) v[i] += qm * ef[3 * (int64_t)c + i] * dt;
        for (int i = 0; i < 3; i++) x[i] += v[i] * dt;
        double lc[4];
        int moves = 0, refl = 0;
        const int r = or_walk(tets, &c, x, v, lc, wall_mode, walk_cap, &moves, &refl);
        if (lc_out) memcpy(lc_out + 4 * p, lc, sizeof(lc));
        if (r == OR_INSIDE) {
            cell[p] = c;
            const or_tet* t = &tets[c];
            for (int k = 0; k < 4; k++) den[t->con[k]] += lc[k];
        } else {
            cell[p] = -1;
        }
        add_stats(stats, r != OR_INSIDE, (moves - refl) > 0, moves, r == OR_CAPPED, refl);
        done++;
    }
    return done;
}

int64_t or_move_omp(const or_tet* tets, const double* ef, double* pos, double* vel, int32_t* cell,
                    int64_t n, double qm, double dt, double* den, int32_t wall_mode, int32_t walk_cap,
                    int32_t* threads) {
    int64_t done = 0;
#ifdef _OPENMP
    if (threads) *threads = omp_get_max_threads();
#else
    if (threads) *threads = 1;
#endif
#pragma omp parallel for schedule(static) reduction(+ : done)
    for (int64_t p = 0; p < n; p++) {
        int32_t c = cell[p];
        if (c < 0) continue;
        double* x = pos + 3 * p;
        double* v = vel + 3 * p;
        for (int i = 0; i < 3; i++) v[i] += qm * ef[3 * (int64_t)c + i] * dt;
        for (int i = 0; i < 3; i++) x[i] += v[i] * dt;
        double lc[4];
        int moves = 0, refl = 0;
        const int r = or_walk(tets, &c, x, v, lc, wall_mode, walk_cap, &moves, &refl);
        if (r == OR_INSIDE) {
            cell[p] = c;
            const or_tet* t = &tets[c];
            for (int k = 0; k < 4; k++) {
#pragma omp atomic
                den[t->con[k]] += lc[k];
            }
        } else {
            cell[p] = -1;
        }
        done++;
    }
    return done;
}

int64_t or_compact(double* pos, double* vel, int32_t* cell, double* lc, int64_t* pid, int64_t n) {
    int64_t k = 0;
    for (int64_t i = 0; i < n; i++) {
        if (cell[i] < 0) continue;
        if (k != i) {
            if (pos) memcpy(pos + 3 * k, pos + 3 * i, 3 * sizeof(double));
            if (vel) memcpy(vel + 3 * k, vel + 3 * i, 3 * sizeof(double));
            if (lc) memcpy(lc + 4 * k, lc + 4 * i, 4 * sizeof(double));
            if (pid) pid[k] = pid[i];
            cell[k] = cell[i];
        }
        k++;
    }
    return k;
}

int64_t or_compact_aos(or_particle* p, int64_t n) {
    int64_t k = 0;
    for (int64_t i = 0; i < n; i++) {
        if (p[i].cell_index < 0) continue;
        if (k != i) p[k] = p[i];
        k++;
    }
    return k;
}

void or_normalise(double* den, const double* vol, int32_t nn, double spwt) {
    for (int32_t n = 0; n < nn; n++) den[n] *= spwt / vol[n];
}

/*
 * Adjoint of K steps under a frozen per-element field E, with every particle's cell
 * path held fixed (the loss is piecewise smooth; derivatives are those of the piece the
 * inputs sit in).  Forward step k+1 for one particle, c_k = cell at the start of step:
 *     v_{k+1} = v_k + (qm E[c_k]) dt,   x_{k+1} = x_k + v_{k+1} dt,   c_{k+1} = walk(x_{k+1})
 * Loss L = sum_n (den[n] - tgt[n])^2, den[n] = (spwt / vol[n]) * sum_p lc_p,v over the
 * (p, v) with con[c_K][v] = n, lc_v = ((1/6)(a_v - x b_v + y g_v - z d_v)) / V.
 * Reverse sweep:
 *     lambda_x(K) = sum_v dL/dlc_v * grad lc_v,  grad lc_v = (-b_v, g_v, -d_v) / (6 V)
 *     lambda_v(K) = 0
 *     for k = K-1 .. 0:  lambda_v(k+1) += lambda_x(k+1) dt ;  lambda_x(k) = lambda_x(k+1)
 *                        dL/dE[c_k] += qm dt lambda_v(k+1) ;   lambda_v(k) = lambda_v(k+1)
 *     dL/d drift = lambda_v(0) . drift_dir   (v_0 = u + drift * drift_dir)
 * Reflecting walls (wall_mode 1): a reflection in step k maps x -> R x + b, v -> R v with
 * R = I - 2 g g^T / |g|^2 (g the wall's plane normal, b fixed by the tet), so before the
 * update above step k's reflections R_m .. R_1 (last one first) are applied to lambda_x and
 * lambda_v (R is symmetric).  The reverse sweep finds them by re-running step k from its
 * stored start state.
 * Particles lost before step K contribute nothing.  This oracle stores the full cell
 * path (K ints per particle; with reflecting walls also the K start states); it is meant
 * for small n.
 */
int or_adjoint(const or_tet* tets, const double* ef, int32_t n_tets, const double* vol, int32_t nn,
               const double* pos0, const double* vel0, const int32_t* cell0, int64_t n, double qm,
               double dt, double spwt, int32_t steps, const double* den_target,
               const double drift_dir[3], int32_t wall_mode, int32_t walk_cap, double* loss, double* dL_dE,
               double* dL_ddrift) {
    (void)n_tets;
    if (steps < 1) return 1;
    double* states = NULL;
    if (wall_mode) {
        states = (double*)malloc(sizeof(double) * 6 * (size_t)n * (size_t)steps);
        if (!states) return 2;
    }
    int32_t* path = (int32_t*)malloc(sizeof(int32_t) * (size_t)n * (size_t)steps);
    double* x = (double*)malloc(sizeof(double) * 3 * (size_t)n);
    double* v = (double*)malloc(sizeof(double) * 3 * (size_t)n);
    int32_t* c = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    double* lc = (double*)malloc(sizeof(double) * 4 * (size_t)n);
    double* den = (double*)calloc((size_t)nn, sizeof(double));
    if (!path || !x || !v || !c || !lc || !den) {
        free(path); free(x); free(v); free(c); free(lc); free(den); free(states);
        return 2;
    }
    memcpy(x, pos0, sizeof(double) * 3 * (size_t)n);
    memcpy(v, vel0, sizeof(double) * 3 * (size_t)n);
    memcpy(c, cell0, sizeof(int32_t) * (size_t)n);
    for (int32_t k = 0; k < steps; k++) {
        for (int64_t p = 0; p < n; p++) path[(size_t)k * n + p] = c[p];
        if (states)
            for (int64_t p = 0; p < n; p++) {
                memcpy(states + 6 * ((size_t)k * n + p), x + 3 * p, 3 * sizeof(double));
                memcpy(states + 6 * ((size_t)k * n + p) + 3, v + 3 * p, 3 * sizeof(double));
            }
        memset(den, 0, sizeof(double) * (size_t)nn);
        or_move(tets, ef, x, v, c, lc, n, qm, dt, den, wall_mode, walk_cap, NULL);
    }
    or_normalise(den, vol, nn, spwt);
    double L = 0.0;
    for (int32_t i = 0; i < nn; i++) {
        const double r = den[i] - den_target[i];
        L += r * r;
    }
    *loss = L;
    double dd = 0.0;
    for (int64_t p = 0; p < n; p++) {
        if (c[p] < 0) continue;
        const or_tet* t = &tets[c[p]];
        double lx[3] = {0.0, 0.0, 0.0};
        for (int q = 0; q < 4; q++) {
            const int32_t node = t->con[q];
            const double g = 2.0 * (den[node] - den_target[node]) * (spwt / vol[node]);
            const double f = g / (6.0 * t->volume);
            lx[0] += f * -t->beta[q];
            lx[1] += f * t->gamma[q];
            lx[2] += f * -t->delta[q];
        }
        double lv[3] = {0.0, 0.0, 0.0};
        for (int32_t k = steps - 1; k >= 0; k--) {
            const int32_t ck = path[(size_t)k * n + p];
            if (states) {
                /* step k again from its start state (or_move's arithmetic), recording the
                 * reflections' plane normals */
                const double* s0 = states + 6 * ((size_t)k * n + p);
                double xs[3] = {s0[0], s0[1], s0[2]}, vs[3] = {s0[3], s0[4], s0[5]}, l4[4], gr[3 * 8];
                for (int i = 0; i < 3; i++) vs[i] += qm * ef[3 * (int64_t)ck + i] * dt;
                for (int i = 0; i < 3; i++) xs[i] += vs[i] * dt;
                int32_t cw = ck;
                int moves = 0, refl = 0, nr = 0;
                if (or_walk_rec(tets, &cw, xs, vs, l4, 1, walk_cap, &moves, &refl, gr, 8, &nr) < 0) {
                    free(path); free(x); free(v); free(c); free(lc); free(den); free(states);
                    return 3;
                }
                for (int j = nr - 1; j >= 0; j--) {
                    const double gx = gr[3 * j], gy = gr[3 * j + 1], gz = gr[3 * j + 2];
                    const double gg = (gx * gx + gy * gy) + gz * gz;
                    const double ax = (2.0 * ((lx[0] * gx + lx[1] * gy) + lx[2] * gz)) / gg;
                    lx[0] = lx[0] - ax * gx;
                    lx[1] = lx[1] - ax * gy;
                    lx[2] = lx[2] - ax * gz;
                    const double av = (2.0 * ((lv[0] * gx + lv[1] * gy) + lv[2] * gz)) / gg;
                    lv[0] = lv[0] - av * gx;
                    lv[1] = lv[1] - av * gy;
                    lv[2] = lv[2] - av * gz;
                }
            }
            for (int d = 0; d < 3; d++) {
                lv[d] += lx[d] * dt;
                dL_dE[3 * (int64_t)ck + d] += (qm * dt) * lv[d];
            }
        }
        dd += (lv[0] * drift_dir[0] + lv[1] * drift_dir[1]) + lv[2] * drift_dir[2];
    }
    *dL_ddrift = dd;
    free(path); free(x); free(v); free(c); free(lc); free(den); free(states);
    return 0;
}
