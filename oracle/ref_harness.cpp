// Harness around the reference's own mesh and walk code, compiled from the reference
// sources where they lie (see oracle/Makefile, target "ref") into oracle/_ref/libref.so.
// TEST INFRASTRUCTURE ONLY: used by tests/ on the build host to pin the oracle and the
// host table code against the reference itself.  It is never built on, or shipped to,
// the GPU side and the product never loads it.
//
// Linked reference units: src/meshes.cpp (LoadVolumeMesh, LoadSurfaceMesh, XtoLtet,
// OutputMesh), src/maths.cpp (det3, det4, rnd, matVecMultiply), src/trace.cpp (TRACE_ME
// timers), src/FESolver.cpp (the FE Poisson-Boltzmann solver) and src/particles.cpp
// (OutputParticles).  src/mini-pic.cpp needs SYCL and cannot be compiled here, so this file
// plays the role of its main(): it defines the two trace globals that src/mini-pic.cpp:31-32
// defines, sets the solver up as src/mini-pic.cpp:86-104 does, and restates the kernel lambda
// body of src/mini-pic.cpp:288-307 serially around the reference's XtoLtet.
//
// FESolver.cpp's Lapack method calls dsysv_ from an external LAPACK, which is not linked: the
// recipe makes the object's LAPACK references weak, so the library loads and the NonLinear
// and GaussSeidel methods (which never reach LAPACK) run from the reference's own code; the
// harness refuses the Lapack method.
#include <cstdint>
#include <cstring>
#include <vector>

#include "FESolver.h"
#include "maths.h"
#include "meshes.h"
#include "particles.h"
#include "trace.h"

void OutputMesh(int ts, Volume& volume, double* phi, double (*ef)[3], double* ion_den);
void OutputParticles(std::vector<Particle>& particles);

int trace::enabled = 0;
Trace trace::current = Trace("__TRACE_BASE__");

static Volume* g_vol = nullptr;
static FESolver* g_fe = nullptr;

extern "C" {

int ref_load_volume(const char* path) {
    delete g_fe;
    g_fe = nullptr;
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

// FESolver on the loaded volume, set up as src/mini-pic.cpp:86-104: phi0, n0, kTe, g = 0 on
// INLET nodes, -wall_potential on FIXED nodes, 0 elsewhere, then startAssembly + preAssembly.
int ref_fe_create(double n0, double phi0, double kTe, double wall_potential) {
    if (!g_vol) return -1;
    delete g_fe;
    g_fe = new FESolver(*g_vol);
    g_fe->phi0 = phi0;
    g_fe->n0 = n0;
    g_fe->kTe = kTe;
    for (size_t n = 0; n < g_vol->nodes.size(); n++) {
        if (g_vol->nodes[n].type == INLET) g_fe->g[n] = 0;
        else if (g_vol->nodes[n].type == FIXED) g_fe->g[n] = -wall_potential;
        else g_fe->g[n] = 0;
    }
    g_fe->startAssembly();
    g_fe->preAssembly();
    return g_fe->neq;
}

// computePhi(ion_den, method) with method 0 NonLinear or 1 GaussSeidel (the reference exits
// when a Newton solve does not converge).
int ref_fe_compute_phi(const double* ion_den, int method) {
    if (!g_fe || (method != 0 && method != 1)) return -1;
    std::vector<double> den(ion_den, ion_den + g_vol->nodes.size());
    g_fe->computePhi(den.data(), method == 0 ? FESolver::NonLinear : FESolver::GaussSeidel);
    return 0;
}

// The Lapack method's linear system without the solve: buildF1Vector, G = K d - F0 - F1 and
// buildJmatrix(Lapack) as FESolver::solve does before dsysv_ (src/FESolver.cpp:303-313).
// J (row-major neq x neq, only j >= i set, the triangle dsysv_ reads as 'L' in column order)
// and G are copied out; d is left unchanged.
int ref_fe_lapack_system(const double* ion_den, double* J, double* G) {
    if (!g_fe) return -1;
    const int neq = g_fe->neq;
    std::vector<double> den(ion_den, ion_den + g_vol->nodes.size());
    g_fe->buildF1Vector(den.data());
    std::vector<double> g(neq);
    matVecMultiply(g.data(), g_fe->K, g_fe->d, neq);
    vecVecSubtract(g.data(), g.data(), g_fe->F0, neq);
    vecVecSubtract(g.data(), g.data(), g_fe->F1, neq);
    g_fe->buildJmatrix(FESolver::Lapack);
    for (int i = 0; i < neq; i++) {
        G[i] = g[i];
        for (int j = 0; j < neq; j++) J[(int64_t)i * neq + j] = j >= i ? g_fe->J[i][j] : 0.0;
    }
    return 0;
}

// uh[n_nodes] and, after updateEf, ef[n_tets][3].
void ref_fe_get(double* uh, double* ef) {
    const size_t nn = g_vol->nodes.size(), ne = g_vol->elements.size();
    for (size_t n = 0; n < nn; n++) uh[n] = g_fe->uh[n];
    g_fe->updateEf();
    for (size_t e = 0; e < ne; e++)
        for (int d = 0; d < 3; d++) ef[3 * e + d] = g_fe->ef[e][d];
}

// OutputMesh(ts, volume, phi, ef, ion_den): writes mesh_{ts+1:04}.vtu in the working directory.
void ref_output_mesh(int ts, const double* phi, const double* ef, const double* ion_den) {
    const size_t nn = g_vol->nodes.size(), ne = g_vol->elements.size();
    std::vector<double> p(phi, phi + nn), den(ion_den, ion_den + nn);
    std::vector<double> e3(ef, ef + 3 * ne);
    OutputMesh(ts, *g_vol, p.data(), (double(*)[3])e3.data(), den.data());
}

// OutputParticles over n Particle records: writes particles.vtp in the working directory.
void ref_output_particles(const void* parts, int64_t n) {
    std::vector<Particle> v((const Particle*)parts, (const Particle*)parts + n);
    OutputParticles(v);
}

}  // extern "C"
