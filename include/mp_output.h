/*
 * mp_output.h - mini-pic's VTK output files, written from host arrays.
 *
 * Reference interfaces replaced (AD1641/mini-fem-pic-main):
 *   src/meshes.cpp:360-449     OutputMesh: mesh_NNNN.vtu (ASCII VTK UnstructuredGrid) with
 *                              node_index, node_type, phi, ion_den per node and ef, cell_volume
 *                              per tet
 *   src/particles.cpp:20-45    OutputParticles: particles.vtp (ASCII VTK PolyData of positions)
 * The files have the reference's layout and number formatting (C++ stream defaults: six
 * significant digits), so ParaView state files and scripts written for it keep working.
 */
#ifndef MP_OUTPUT_H
#define MP_OUTPUT_H

#include <stdint.h>
#include "mp_mesh.h"

#ifdef __cplusplus
extern "C" {
#endif

/* OutputMesh to `path` (the reference names it mesh_%04d.vtu with ts + 1).  phi, ion_den:
 * [n_nodes]; ef: [n_tets][3]. */
int mp_write_mesh_vtu(const mp_mesh* m, const char* path, const double* phi, const double* ef,
                      const double* ion_den);
/* OutputParticles to `path`: n positions read at pos + k * stride_bytes (24 for [n][3] arrays,
 * 88 for the Particle record). */
int mp_write_particles_vtp(const char* path, const double* pos, int64_t n, int64_t stride_bytes);
const char* mp_output_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* MP_OUTPUT_H */
