/*
 * mp_inject.h - ion injection of mini-pic's coupled loop, and its random number stream.
 *
 * Reference interfaces replaced (AD1641/mini-fem-pic-main):
 *   src/maths.cpp:17-19        std::mt19937 mt_gen(0); uniform_real_distribution<double>(0, 1);
 *                              double rnd()
 *   src/mini-pic.cpp:208-265   int InjectIons(Species&, Volume&, FESolver&, Parameters)
 *
 * The generator is the C++ standard library's mt19937 with uniform_real_distribution<double>
 * over [0, 1), the reference's own choice, so a stream seeded alike draws the reference's
 * numbers (with the same standard library).
 */
#ifndef MP_INJECT_H
#define MP_INJECT_H

#include <stdint.h>
#include "mp_mesh.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mp_rng mp_rng;

/* mt19937 seeded with `seed` (the reference seeds 0). */
int    mp_rng_create(uint32_t seed, mp_rng** out);
void   mp_rng_free(mp_rng* r);
/* n draws of rnd() into out. */
int    mp_rng_uniform(mp_rng* r, double* out, int64_t n);

/* Macroparticles InjectIons would create this step for the inlet faces of m, given the
 * carried fraction rem (no random draws). */
int    mp_inject_count(const mp_mesh* m, double plasma_den, double ion_velocity, double dt, double spwt,
                       double rem, int64_t* n_new);
/* InjectIons: for each inlet face in order, num = (plasma_den*ion_velocity*area*dt)/spwt + rem
 * macroparticles (rem carries the fraction across faces and steps), each at a uniform point
 * of the face (two draws, folded into the triangle), velocity normal*ion_velocity, cell = the
 * face's tet, then the velocity rewound by half a step in that tet's field:
 * v -= ((charge/mass)*E)*(0.5*dt).  q_over_m is charge/mass; ef[n_tets][3].  Writes *n_new
 * particles to out (cap must hold mp_inject_count's figure; lc is zeroed), updates *rem. */
int    mp_inject_ions(const mp_mesh* m, mp_rng* r, double plasma_den, double ion_velocity, double dt,
                      double spwt, double q_over_m, const double* ef, double* rem, mp_particle* out,
                      int64_t cap, int64_t* n_new);
/* InjectIons with a thermal spread (an extension: the reference injects the cold beam of
 * src/mini-pic.cpp:245-246).  Everything as mp_inject_ions (same counts, same two position
 * draws per particle from the same stream), and before the half-step rewind each velocity
 * component gets v_thermal * z: vel[i] = normal[i]*ion_velocity + v_thermal*z[i].  v_thermal is
 * the Maxwellian's per-component standard deviation sqrt(kT/m) in m/s.  z[0..2] of the k-th
 * particle the stream injects (k = 0, 1, ... over the stream's life) are Box-Muller normals
 * from draws 0..3 of item k of counter-based SplitMix64 (draw j of item k is
 * mix64(s + (16k + j + 1) * 0x9E3779B97F4A7C15) >> 11, times 2^-53) with
 * s = mix64(seed ^ 0x7468726D616C0000), seed as given to mp_rng_create; the normals use a
 * generator of their own, so the positions are the cold beam's for any v_thermal.
 * v_thermal = 0 injects exactly what mp_inject_ions does. */
int    mp_inject_ions_thermal(const mp_mesh* m, mp_rng* r, double plasma_den, double ion_velocity,
                              double v_thermal, double dt, double spwt, double q_over_m, const double* ef,
                              double* rem, mp_particle* out, int64_t cap, int64_t* n_new);

#ifdef __cplusplus
}
#endif
#endif /* MP_INJECT_H */
