/*
 * mp_fem.h - host finite-element Poisson-Boltzmann solver of mini-pic's coupled loop.
 *
 * The field solve stays on the host, as in the reference; this is the caller on the far side
 * of the mover: each step mini-pic moves the ions, solves for phi from their nodal density
 * and refreshes the per-element E the next push reads.
 *
 * Reference interfaces replaced (AD1641/mini-fem-pic-main):
 *   src/FESolver.h:24-90          class FESolver
 *   src/FESolver.cpp:47-127       FESolver::FESolver (ID, LM, quadrature, computeNX)
 *   src/FESolver.cpp:155-185      startAssembly / addKe / addFe
 *   src/FESolver.cpp:207-287      computeNX / inverse
 *   src/FESolver.cpp:290-321      solve (method dispatch)
 *   src/FESolver.cpp:324-364      solveNonLinear (Newton-Raphson, <= 10 iterations)
 *   src/FESolver.cpp:367-431      buildJmatrix
 *   src/FESolver.cpp:434-485      preAssembly (K, F0)
 *   src/FESolver.cpp:488-526      buildF1Vector
 *   src/FESolver.cpp:528-558      solveLinearLapack (replaced by a dense Cholesky)
 *   src/FESolver.cpp:561-601      solveLinear (Gauss-Seidel)
 *   src/FESolver.cpp:604-619      computePhi
 *   src/FESolver.cpp:622-636      updateEf
 *   src/mini-pic.cpp:89-104       reference values (phi0, n0, kTe) and the g array
 *
 * K and J are held in compressed rows (ascending columns, the pattern of node pairs that
 * share an element); every sum runs over the same nonzero terms in the same order as the
 * reference's dense loops, so d, phi and E are bit-identical to the dense arithmetic.
 * As in the reference, the "ff" force vector F1 accumulates over every buildF1Vector call
 * (it is cleared only at assembly).  With it the Newton iteration of a long coupled run stops
 * converging once the accumulated term dominates (the reference then exits);
 * MP_FE_RESET_F1 clears F1 at each call instead, the textbook Newton residual (an option
 * of this build, off by default).
 */
#ifndef MP_FEM_H
#define MP_FEM_H

#include <stdint.h>
#include "mp_mesh.h"

#ifdef __cplusplus
extern "C" {
#endif

/* FESolver::Method (src/FESolver.h:26) */
#define MP_FE_NONLINEAR    0  /* Newton-Raphson with Gauss-Seidel inner solves            */
#define MP_FE_GAUSS_SEIDEL 1  /* one linearised step, Gauss-Seidel                        */
#define MP_FE_DIRECT       2  /* one linearised step, dense Cholesky of the definite J
                                 (negated when negative definite); stands in for the
                                 Lapack method's dsysv_ (LAPACK is not part of this build),
                                 so it is not bit-identical to it                         */

/* mp_fe_create flags */
#define MP_FE_RESET_F1     1  /* clear F1 at each buildF1Vector (the reference accumulates) */

typedef struct mp_fe mp_fe;

typedef struct mp_fe_stats {
    int32_t nr_iterations;   /* Newton iterations taken (1 for the linearised methods) */
    int32_t gs_sweeps;       /* Gauss-Seidel sweeps summed over the Newton iterations   */
    int32_t gs_unconverged;  /* inner solves that hit 10000 sweeps (reference warns)    */
    int32_t zero_diagonal;   /* rows skipped for |J_uu| < 1e-12 (reference warns)       */
    double  nr_l2;           /* last Newton step norm sqrt(sum y^2) / neq               */
} mp_fe_stats;

/* FESolver(volume) + main's set-up: phi0, n0, kTe; g = 0 on INLET nodes, -wall_potential on
 * FIXED nodes, 0 elsewhere; startAssembly + preAssembly.  NORMAL and OPEN nodes are the
 * unknowns. */
int     mp_fe_create(const mp_mesh* m, double n0, double phi0, double kTe, double wall_potential,
                     int32_t flags, mp_fe** out);
void    mp_fe_free(mp_fe* fe);
int32_t mp_fe_num_unknowns(const mp_fe* fe);
/* computePhi(ion_den, method): ion_den[n_nodes].  A Newton solve that has not converged in 10
 * iterations returns MP_ERR_STATE (the reference exits).  st may be NULL. */
int     mp_fe_compute_phi(mp_fe* fe, const double* ion_den, int32_t method, mp_fe_stats* st);
/* updateEf: ef[n_tets][3] from the current phi. */
int     mp_fe_update_ef(const mp_fe* fe, double* ef);
/* uh[n_nodes]: the current nodal potential. */
int     mp_fe_get_phi(const mp_fe* fe, double* uh);
/* The assembled stiffness matrix in compressed rows over the unknowns:
 * row_ptr[neq+1], col[nnz], val[nnz]; any pointer may be NULL (then only *nnz is set). */
int     mp_fe_get_stiffness(const mp_fe* fe, int64_t* nnz, int64_t* row_ptr, int32_t* col, double* val);
/* The solver's own method timings, the rows FESolver's TRACE_ME calls add to trace.csv
 * (src/FESolver.cpp:48-622, src/trace.cpp:36-76): entry k in [0, MP_FE_PROFILE_ENTRIES) gives
 * the method's name, its call count, its inclusive wall seconds and the seconds spent in its
 * traced callees, accumulated since mp_fe_create.  "solveLinearDirect" is MP_FE_DIRECT's solve
 * (the reference's solveLinearLapack); ~FESolver, startAssembly (part of preAssembly here),
 * addKe/addFe and inverse have no separate rows.  Any output pointer may be NULL. */
#define MP_FE_PROFILE_ENTRIES 11
int     mp_fe_get_profile(const mp_fe* fe, int32_t k, const char** name, int64_t* calls, double* total_s,
                          double* child_s);
const char* mp_fe_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* MP_FEM_H */
