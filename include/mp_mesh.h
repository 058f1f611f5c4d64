/*
 * mp_mesh.h - host-side mesh tables, frozen fields and synthetic particles for the
 * MI355X particle mover.  Pure CPU code in the same shared library as mp_abi.h.
 *
 * The tables are the ones mini-pic's LoadVolumeMesh precomputes once and MoveParticles
 * consumes every step:
 *   src/meshes.cpp:27-265   LoadVolumeMesh: SALOME .dat parse (coordinates / 1000),
 *                           det4 volume with ABCD->ADBC flip, alpha..delta via det3,
 *                           face-neighbour table cell_con, node volumes at centroids,
 *                           OPEN node marking.
 *   src/maths.cpp:22-57     det4 / det3 evaluation order.
 *   src/FESolver.cpp:207-287,622-636  computeNX / inverse / updateEf (E from nodal phi).
 * Box meshes, sinusoidal fields and seeded particle generators are synthetic inputs
 * that the reference does not have; they feed the GPU path and the oracle alike.
 */
#ifndef MP_MESH_H
#define MP_MESH_H

#include <stdint.h>
#include "mp_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Node types, src/meshes.h:22 (enum NodeType {NORMAL,OPEN,INLET,FIXED}). */
#define MP_NODE_NORMAL 0
#define MP_NODE_OPEN   1
#define MP_NODE_INLET  2
#define MP_NODE_FIXED  3

typedef struct mp_mesh mp_mesh;

/* Structured box of nx*ny*nz hexes on [0,lx]x[0,ly]x[0,lz], six tets per hex sharing the
 * hex's main diagonal (Kuhn split).  Node id = i + (nx+1)*(j + (ny+1)*k); tet id =
 * 6*(i + nx*(j + ny*k)) + p, p indexing the axis order of the path corner(0,0,0) ->
 * corner(1,1,1): p = 0:xyz 1:xzy 2:yxz 3:yzx 4:zxy 5:zyx.  Tables are then computed with
 * the LoadVolumeMesh arithmetic. */
int mp_mesh_box(int32_t nx, int32_t ny, int32_t nz, double lx, double ly, double lz,
                mp_mesh** out);
/* SALOME .dat volume mesh as LoadVolumeMesh reads it (src/meshes.cpp:28-265). */
int mp_mesh_load_dat(const char* path, mp_mesh** out);
/* Surface mesh: marks listed nodes with node_type (src/meshes.cpp:268-288); for
 * MP_NODE_INLET also builds the inlet faces (src/meshes.cpp:290-354). */
int mp_mesh_load_surface(mp_mesh* m, const char* path, int32_t node_type, int32_t invert_normals);
void mp_mesh_free(mp_mesh* m);

int32_t          mp_mesh_num_nodes(const mp_mesh* m);
int32_t          mp_mesh_num_tets(const mp_mesh* m);
const mp_tetra*  mp_mesh_tets(const mp_mesh* m);
const double*    mp_mesh_node_pos(const mp_mesh* m);     /* [n_nodes][3] */
const double*    mp_mesh_node_volume(const mp_mesh* m);  /* [n_nodes]    */
const int32_t*   mp_mesh_node_type(const mp_mesh* m);    /* [n_nodes]    */
const char*      mp_mesh_last_error(void);

/* Inlet faces (struct Face, src/meshes.h:46-54). */
typedef struct mp_face {
    int32_t con[3];
    double  area;
    double  u[3];
    double  v[3];
    int32_t cell_con;
    double  normal[3];
} mp_face;
int32_t        mp_mesh_num_inlet_faces(const mp_mesh* m);
const mp_face* mp_mesh_inlet_faces(const mp_mesh* m);

/* Containing tet of each point of a box mesh by the closed-form Kuhn rule (sort the
 * fractional coordinates); -1 outside the box. */
int mp_mesh_locate_box(const mp_mesh* m, const double* pos, int64_t n, int32_t* cell);

/* E per element from nodal potential phi[n_nodes]: computeNX + updateEf
 * (src/FESolver.cpp:207-287, 622-636).  ef: [n_tets][3]. */
int mp_field_from_potential(const mp_mesh* m, const double* phi, double* ef);
/* Frozen field of phi = amp * sin(pi x/Lx) sin(pi y/Ly) sin(pi z/Lz) over the mesh's
 * bounding box; ef: [n_tets][3]. */
int mp_field_sinusoid(const mp_mesh* m, double amp, double* ef);

/* Seeded synthetic particles.  Counter-based SplitMix64 (draw j of particle p is
 * splitmix64_mix(seed + (16*p + j + 1) * 0x9E3779B97F4A7C15)), Box-Muller normals.
 * Positions uniform in [lo, hi); velocities N(0, sigma^2) per component + drift. */
int mp_synth_uniform(uint64_t seed, int64_t n, const double lo[3], const double hi[3],
                     double sigma, const double drift[3], double* pos, double* vel);
/* Positions from a Gaussian mixture (n_comp components, weights w, centres c[k][3],
 * common sigma_pos), resampled until inside [lo, hi). */
int mp_synth_mixture(uint64_t seed, int64_t n, const double lo[3], const double hi[3],
                     int32_t n_comp, const double* w, const double* centres, double sigma_pos,
                     double sigma_vel, const double drift[3], double* pos, double* vel);

#ifdef __cplusplus
}
#endif
#endif /* MP_MESH_H */
