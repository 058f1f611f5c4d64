/*
 * mp_oracle.h - CPU oracle of mini-pic's MoveParticles hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker: only tests/, the smoke check and the
 * CPU-baseline leg of bench.py may load it.  The product path never calls it.
 *
 * A serial restatement, in plain C built with -ffp-contract=off, of
 *   src/mini-pic.cpp:287-308   kernel body: push, XtoLtet, scatter, lost flag
 *   src/meshes.h:75-100        XtoLtet barycentric walk
 *   src/mini-pic.cpp:316       std::erase_if(cell_index < 0), stable
 *   src/mini-pic.cpp:319       den[n] *= spwt / node volume
 * run in particle (pid) order, so den is the ascending-pid serial fold.  Extensions that
 * the reference lacks and this build defines: a walk cap, reflecting walls, and the
 * adjoint of K frozen-field steps.
 */
#ifndef MP_ORACLE_H
#define MP_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same 168 B layout as the reference Tetra (src/meshes.h:33-43). */
typedef struct or_tet {
    int32_t con[4];
    double volume;
    double alpha[4], beta[4], gamma[4], delta[4];
    int32_t cell_con[4];
} or_tet;

/* Same 88 B layout as the reference Particle (src/particles.h:20-25). */
typedef struct or_particle {
    double pos[3];
    double vel[3];
    double lc[4];
    int32_t cell_index;
    int32_t pad;
} or_particle;

/* stats[6]: particle_steps, lost, crossed, moves, cap_hits, reflections (added to). */

/* One serial kernel pass over AoS particles exactly as the reference lambda
 * (src/mini-pic.cpp:288-307), absorbing walls.  den (raw, caller-zeroed) gets the
 * serial fold.  Returns particles processed. */
int64_t or_move_aos(const or_tet* tets, const double* ef, or_particle* p, int64_t n, double qm,
                    double dt, double* den, int32_t walk_cap, int64_t* stats);

/* One serial pass over [n][3] positions/velocities and cells; cell < 0 entries are
 * skipped.  wall_mode 0 absorbs (cell = -1), 1 reflects.  lc_out [n][4] may be NULL. */
int64_t or_move(const or_tet* tets, const double* ef, double* pos, double* vel, int32_t* cell,
                double* lc_out, int64_t n, double qm, double dt, double* den, int32_t wall_mode,
                int32_t walk_cap, int64_t* stats);

/* The same with OpenMP threads and "#pragma omp atomic" density adds: the CPU analogue of
 * the reference's SYCL kernel on host cores, used as the timed CPU baseline.  Returns the
 * processed count; threads used is written to *threads when not NULL. */
int64_t or_move_omp(const or_tet* tets, const double* ef, double* pos, double* vel, int32_t* cell,
                    int64_t n, double qm, double dt, double* den, int32_t wall_mode, int32_t walk_cap,
                    int32_t* threads);

/* Stable removal of cell < 0 from the parallel arrays (any may be NULL except cell). */
int64_t or_compact(double* pos, double* vel, int32_t* cell, double* lc, int64_t* pid, int64_t n);
int64_t or_compact_aos(or_particle* p, int64_t n);

/* den[n] *= spwt / vol[n] (src/mini-pic.cpp:319). */
void or_normalise(double* den, const double* vol, int32_t nn, double spwt);

/* Adjoint of K frozen-field steps (see mp_oracle.c for the derivation).  Inputs are the
 * initial particles (pos0, vel0 [n][3], cell0 [n]); den_target [nn] and the loss
 * L = sum_n (den_K[n] - den_target[n])^2 with den_K normalised by spwt / vol.  wall_mode 0
 * absorbs at boundary faces, 1 reflects.
 * Outputs: loss, dL/dE [T][3] (added to, caller zeroes) and dL/d(drift along drift_dir):
 * sum over particles of lambda_v0 . drift_dir.  Returns 3 if a step reflects a particle more
 * than 8 times. */
int or_adjoint(const or_tet* tets, const double* ef, int32_t n_tets, const double* vol, int32_t nn,
               const double* pos0, const double* vel0, const int32_t* cell0, int64_t n, double qm,
               double dt, double spwt, int32_t steps, const double* den_target,
               const double drift_dir[3], int32_t wall_mode, int32_t walk_cap, double* loss, double* dL_dE,
               double* dL_ddrift);

#ifdef __cplusplus
}
#endif
#endif
