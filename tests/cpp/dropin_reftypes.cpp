// Compile check (build container only): the drop-in header binds the reference's own types.
// Compiled with -fsyntax-only against the reference's headers where they lie; the templates
// DeviceData(Volume&) and MoveParticles(Species&, Volume&, FESolver&, const Parameters&,
// DeviceData&) are instantiated with src/meshes.h, src/particles.h, src/FESolver.h and
// src/parameters.h, in the call sequence of src/mini-pic.cpp:119 and :135.
#include "FESolver.h"
#include "meshes.h"
#include "parameters.h"
#include "particles.h"

#include "minipic_amd/move_particles.hpp"

static_assert(sizeof(Particle) == sizeof(mp_particle), "Particle layout");
static_assert(sizeof(Tetra) == sizeof(mp_tetra), "Tetra layout");

void move_one_step(Species& ions, Volume& volume, FESolver& solver, const Parameters& params) {
    DeviceData device(volume);
    device.reserveParticles(ions.particles.size());
    MoveParticles(ions, volume, solver, params, device);
}
