// Drop-in replacement of mini-pic's device offload: DeviceData and MoveParticles.
//
// Replaces src/mini-pic.cpp:36-53 (struct DeviceData), the MoveParticles prototype at
// :59 and the definitions at :267-355.  Include this header after the reference's own
// particles.h, meshes.h, FESolver.h and parameters.h; the templates bind to the
// reference's Species, Volume, FESolver and Parameters types (or any types with the same
// members), and the Particle and Tetra records are passed to the C ABI as they are: their
// byte layouts are checked here against mp_particle (88 B) and mp_tetra (168 B).
//
// Behaviour kept from the reference:
//   * DeviceData(volume) uploads the Tetra table once and exits(-1) with a message if the
//     device lacks what is needed (src/mini-pic.cpp:322-338);
//   * MoveParticles copies the field every step, pushes, walks and scatters on the GPU,
//     removes particles with cell_index < 0 keeping the order of the rest (std::erase_if),
//     writes ions.den = raw * (spwt / node volume), writes lc, and blocks until done
//     (src/mini-pic.cpp:268-320);
//   * any device failure prints a message and exits(-1), like the reference.
// Added: DeviceData::scatter_mode selects the deterministic scatter
// (MP_SCATTER_DETERMINISTIC: density bit-identical to a serial MoveParticles).
#pragma once

#include <cstddef>
#include <cstdlib>
#include <iostream>
#include <vector>

#include "mp_abi.h"

struct DeviceData {
    mp_ctx* ctx = nullptr;
    int n_nodes = 0;
    int n_elements = 0;
    int scatter_mode = MP_SCATTER_ATOMIC;

    template <class VolumeT>
    explicit DeviceData(VolumeT& volume) {
        static_assert(sizeof(volume.elements[0]) == sizeof(mp_tetra), "Tetra must keep the 168 B reference layout");
        n_nodes = (int)volume.nodes.size();
        n_elements = (int)volume.elements.size();
        std::vector<double> node_volume((size_t)n_nodes);
        for (int n = 0; n < n_nodes; n++) node_volume[(size_t)n] = volume.nodes[(size_t)n].volume;
        const int rc = mp_create(reinterpret_cast<const mp_tetra*>(volume.elements.data()), n_elements,
                                 node_volume.data(), n_nodes, &ctx);
        if (rc != MP_OK) {
            std::cerr << "Failed to set up the GPU for MoveParticles: " << mp_last_error(nullptr) << std::endl;
            std::exit(-1);
        }
    }
    ~DeviceData() { mp_destroy(ctx); }
    DeviceData(const DeviceData&) = delete;
    DeviceData& operator=(const DeviceData&) = delete;

    // Capacity is managed by the library (grows to 2n on overflow, as the reference does).
    void reserveParticles(size_t) {}
};

template <class SpeciesT, class VolumeT, class SolverT, class ParamsT>
void MoveParticles(SpeciesT& ions, VolumeT& volume, SolverT& solver, const ParamsT& params,
                   DeviceData& device) {
    static_assert(sizeof(ions.particles[0]) == sizeof(mp_particle), "Particle must keep the 88 B reference layout");
    (void)volume;
    int rc = mp_upload_field(device.ctx, &solver.ef[0][0], device.n_elements);
    size_t n = ions.particles.size();
    if (rc == MP_OK) {
        const double q_over_m = ions.charge / ions.mass;
        rc = mp_step_aos(device.ctx, reinterpret_cast<mp_particle*>(ions.particles.data()), &n, q_over_m,
                         params.dt, ions.spwt, ions.den, device.scatter_mode);
    }
    if (rc != MP_OK) {
        std::cerr << "MoveParticles failed: " << mp_last_error(device.ctx) << std::endl;
        std::exit(-1);
    }
    ions.particles.erase(ions.particles.begin() + (std::ptrdiff_t)n, ions.particles.end());
}
