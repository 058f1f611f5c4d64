// Ion injection and its random stream (declared in mp_inject.h), following
// src/mini-pic.cpp:208-265 and src/maths.cpp:17-19 operation for operation
// (built with -ffp-contract=off).
#include "mp_inject.h"

#include <cmath>
#include <random>
#include <string>

#include "synth_rng.h"

// The reference's stream (mt19937 + uniform_real_distribution<double>) for positions, plus the
// thermal extension's counter-based normals: the k-th particle this stream injects takes
// normals 0..2 of item k under the stream's seed (their own generator, so the position draws
// stay the reference's whatever the temperature).
struct mp_rng {
    std::mt19937 gen;
    std::uniform_real_distribution<double> dist{0, 1.0};
    uint64_t seed;
    int64_t injected = 0;
    explicit mp_rng(uint32_t s) : gen(s), seed(mprng::mix64((uint64_t)s ^ 0x7468726D616C0000ull)) {}
    double next() { return dist(gen); }
    void thermal(double z[3]) {
        double z3;
        mprng::normal2(seed, injected, 0, z[0], z[1]);
        mprng::normal2(seed, injected, 2, z[2], z3);
        injected++;
    }
};

namespace {

thread_local std::string g_inj_err;

int fail(const std::string& m) {
    g_inj_err = m;
    return MP_ERR_ARG;
}

// The per-face macroparticle count and the carried fraction (src/mini-pic.cpp:216-227).
int face_count(const mp_face& f, double plasma_den, double ion_velocity, double dt, double spwt, double& rem) {
    const double num_per_sec = plasma_den * ion_velocity * f.area;
    const double num_real = num_per_sec * dt;
    const double fnum_mp = num_real / spwt + rem;
    const int num_mp = (int)fnum_mp;
    rem = fnum_mp - num_mp;
    return num_mp;
}

}  // namespace

extern "C" {

int mp_rng_create(uint32_t seed, mp_rng** out) {
    if (!out) return fail("mp_rng_create: null argument");
    *out = new mp_rng(seed);
    return MP_OK;
}

void mp_rng_free(mp_rng* r) { delete r; }

int mp_rng_uniform(mp_rng* r, double* out, int64_t n) {
    if (!r || n < 0 || (n > 0 && !out)) return fail("mp_rng_uniform: bad argument");
    for (int64_t i = 0; i < n; i++) out[i] = r->next();
    return MP_OK;
}

int mp_inject_count(const mp_mesh* m, double plasma_den, double ion_velocity, double dt, double spwt, double rem,
                    int64_t* n_new) {
    if (!m || !n_new) return fail("mp_inject_count: null argument");
    const int32_t nf = mp_mesh_num_inlet_faces(m);
    const mp_face* faces = mp_mesh_inlet_faces(m);
    int64_t total = 0;
    for (int32_t k = 0; k < nf; k++) total += face_count(faces[k], plasma_den, ion_velocity, dt, spwt, rem);
    *n_new = total;
    return MP_OK;
}

int mp_inject_ions_thermal(const mp_mesh* m, mp_rng* r, double plasma_den, double ion_velocity,
                           double v_thermal, double dt, double spwt, double q_over_m, const double* ef,
                           double* rem, mp_particle* out, int64_t cap, int64_t* n_new) {
    if (!m || !r || !ef || !rem || !n_new) return fail("mp_inject_ions: null argument");
    if (!(v_thermal >= 0) || std::isinf(v_thermal)) return fail("mp_inject_ions: v_thermal must be finite and >= 0");
    int64_t need = 0;
    int rc = mp_inject_count(m, plasma_den, ion_velocity, dt, spwt, *rem, &need);
    if (rc) return rc;
    if (need > cap || (need > 0 && !out)) return fail("mp_inject_ions: output capacity " + std::to_string(cap) +
                                                      " < " + std::to_string(need) + " new particles");
    const int32_t nf = mp_mesh_num_inlet_faces(m);
    const mp_face* faces = mp_mesh_inlet_faces(m);
    const double* npos = mp_mesh_node_pos(m);
    const int32_t n_tets = mp_mesh_num_tets(m);
    int64_t k = 0;
    for (int32_t fi = 0; fi < nf; fi++) {
        const mp_face& face = faces[fi];
        if (face.cell_con < 0 || face.cell_con >= n_tets) return fail("mp_inject_ions: inlet face without a tet");
        const int num_mp = face_count(face, plasma_den, ion_velocity, dt, spwt, *rem);
        const double* p0 = npos + 3 * (size_t)face.con[0];
        for (int p = 0; p < num_mp; p++) {
            mp_particle part{};
            double a = r->next();
            double b = r->next();
            if ((a + b) > 1) {
                a = 1 - a;
                b = 1 - b;
            }
            for (int i = 0; i < 3; i++) {
                part.pos[i] = a * face.u[i] + b * face.v[i] + p0[i];
                part.vel[i] = face.normal[i] * ion_velocity;  // the reference's cold beam
            }
            if (v_thermal > 0) {
                double z[3];
                r->thermal(z);
                for (int i = 0; i < 3; i++) part.vel[i] = part.vel[i] + v_thermal * z[i];
            }
            part.cell_index = face.cell_con;
            const double* e = ef + 3 * (size_t)face.cell_con;
            for (int i = 0; i < 3; i++) part.vel[i] -= q_over_m * e[i] * (0.5 * dt);
            out[k++] = part;
        }
    }
    *n_new = k;
    return MP_OK;
}

int mp_inject_ions(const mp_mesh* m, mp_rng* r, double plasma_den, double ion_velocity, double dt, double spwt,
                   double q_over_m, const double* ef, double* rem, mp_particle* out, int64_t cap,
                   int64_t* n_new) {
    return mp_inject_ions_thermal(m, r, plasma_den, ion_velocity, 0.0, dt, spwt, q_over_m, ef, rem, out, cap, n_new);
}

}  // extern "C"
