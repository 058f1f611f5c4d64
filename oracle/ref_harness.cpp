// Harness around the reference's own mesh and walk code, compiled from the reference
// sources where they lie (see oracle/Makefile, target "ref") into oracle/_ref/libref.so.
// TEST INFRASTRUCTURE ONLY: used by tests/ on the build host to pin the oracle and the
// host table code against the reference itself.  It is never built on, or shipped to,
// the GPU side and the product never loads it.
//
// Linked reference units: src/meshes.cpp (LoadVolumeMesh, LoadSurfaceMesh, XtoLtet),
// src/maths.cpp (det3, det4, rnd), src/trace.cpp (TRACE_ME timers).  src/mini-pic.cpp
// needs SYCL and cannot be compiled here, so this file plays the role of its main():
// it defines the two trace globals that src/mini-pic.cpp:31-32 defines and restates the
// kernel lambda body of src/mini-pic.cpp:288-307 serially around the reference's XtoLtet.
#include <cstdint>
#include <cstring>
#include <vector>

#include "maths.h"
#include "meshes.h"
#include "particles.h"
#include "trace.h"

int trace::enabled = 0;
Trace trace::current = Trace("__TRACE_BASE__");

static Volume* g_vol = nullptr;

extern "C" {

int ref_load_volume(const char* path) {
    delete g_vol;
    g_vol = new Volume;
    if (!LoadVolumeMesh(path, *g_vol)) return -1;
    return 0;
}

int ref_load_surface(const char* path, int node_type, int invert) {
    if (!g_vol) return -1;
    return LoadSurfaceMesh(path, *g_vol, (NodeType)node_type, invert != 0) ? 0 : -1;
}

int32_t ref_num_nodes() { return g_vol ? (int32_t)g_vol->nodes.size() : 0; }
int32_t ref_num_tets() { return g_vol ? (int32_t)g_vol->elements.size() : 0; }
int32_t ref_num_inlet_faces() { return g_vol ? (int32_t)g_vol->inlet_faces.size() : 0; }

// Tetra records (168 B each, the reference layout), node positions/volumes/types.
void ref_get_tables(void* tets, double* node_pos, double* node_vol, int32_t* node_type) {
    static_assert(sizeof(Tetra) == 168, "Tetra layout");
    std::memcpy(tets, g_vol->elements.data(), sizeof(Tetra) * g_vol->elements.size());
    for (size_t n = 0; n < g_vol->nodes.size(); n++) {
        for (int d = 0; d < 3; d++) node_pos[3 * n + d] = g_vol->nodes[n].pos[d];
        node_vol[n] = g_vol->nodes[n].volume;
        node_type[n] = (int32_t)g_vol->nodes[n].type;
    }
}

// Inlet faces as (con[3], cell_con, area, u[3], v[3], normal[3]) rows.
void ref_get_inlet(int32_t* con, int32_t* cell_con, double* area, double* u, double* v, double* normal) {
    for (size_t f = 0; f < g_vol->inlet_faces.size(); f++) {
        const Face& F = g_vol->inlet_faces[f];
        for (int i = 0; i < 3; i++) {
            con[3 * f + i] = F.con[i];
            u[3 * f + i] = F.u[i];
            v[3 * f + i] = F.v[i];
            normal[3 * f + i] = F.normal[i];
        }
        cell_con[f] = F.cell_con;
        area[f] = F.area;
    }
}

// Serial MoveParticles kernel body over reference Particle records, with the walk done
// by the reference's own XtoLtet (src/meshes.h:75-100) on the given Tetra table.
void ref_move_serial(const void* tets_v, const double* ef, void* parts_v, int64_t n, double q_over_m,
                     double dt, double* den) {
    const Tetra* elements = (const Tetra*)tets_v;
    Particle* particles = (Particle*)parts_v;
    for (int64_t p = 0; p < n; p++) {
        Particle& part = particles[p];
        for (int i = 0; i < 3; i++) part.vel[i] += q_over_m * ef[3 * (int64_t)part.cell_index + i] * dt;
        for (int i = 0; i < 3; i++) part.pos[i] += part.vel[i] * dt;
        if (XtoLtet(part, elements)) {
            const Tetra& tet = elements[part.cell_index];
            for (int v = 0; v < 4; v++) den[tet.con[v]] += part.lc[v];
        } else {
            part.cell_index = -1;
        }
    }
}

// First n draws of the reference's rnd() (std::mt19937 seeded 0, src/maths.cpp:17-19).
void ref_rnd(int64_t n, double* out) {
    for (int64_t i = 0; i < n; i++) out[i] = rnd();
}

// The reference's det3 / det4 on row-major inputs.
double ref_det3(const double* m) {
    double M[3][3];
    std::memcpy(M, m, sizeof(M));
    return det3(M);
}
double ref_det4(const double* m) {
    double M[4][4];
    std::memcpy(M, m, sizeof(M));
    return det4(M);
}

}  // extern "C"
