// MarchingCubes.h — header-only C++ drop-in for the reference class
// `MarchingCubes` (BoltzmannFluids/Source/MarchingCubes/MarchingCubes.h:13-38),
// implemented on the C ABI of bf.h.
//
// The reference computes into GL buffers from a Texture3D; this version takes
// the scalar field as a host array (X*Y*Z floats in cell order) or reads the
// density resident in a solver, extracts the surface on the device and returns
// host arrays: vertices (3 floats each) and triangle indices (3 uint32 each).
// The rendering members (render(...)) are GL programs outside this library.
//
// Link with libbfamd.so.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "bf.h"

namespace bf_compat {

struct IsoMesh {
    std::vector<float> vertices;   // x, y, z per vertex
    std::vector<uint32_t> indices; // three per triangle
};

}  // namespace bf_compat

class MarchingCubes {
public:
    MarchingCubes() = default;

    // worst-case buffer sizes (MarchingCubes.h:17-18)
    template <class IVec3>
    size_t get_index_buffer_needed_size_in_bytes(IVec3 volume_resolution) {
        int64_t v = 0, i = 0;
        capacity(volume_resolution, v, i);
        return (size_t)i * sizeof(uint32_t);
    }
    template <class IVec3>
    size_t get_vertex_buffer_needed_size_in_bytes(IVec3 volume_resolution) {
        int64_t v = 0, i = 0;
        capacity(volume_resolution, v, i);
        return (size_t)v * 3 * sizeof(float);
    }

    // MarchingCubes::compute (MarchingCubes.cpp:3-48): the reference hard-codes
    // threshold 1 and mesh scale 1 (MarchingCubes.cpp:46-47); both are arguments here.
    template <class IVec3>
    bf_compat::IsoMesh compute(const std::vector<float> &volume, IVec3 resolution, float threshold = 1.0f,
                               float scale_x = 1.0f, float scale_y = 1.0f, float scale_z = 1.0f) {
        const int32_t res[3] = {(int32_t)resolution.x, (int32_t)resolution.y, (int32_t)resolution.z};
        if ((int64_t)volume.size() != (int64_t)res[0] * res[1] * res[2])
            throw std::runtime_error("[MarchingCubes Error] volume size does not match the resolution");
        const float sc[3] = {scale_x, scale_y, scale_z};
        return run([&](float *v, int64_t vc, uint32_t *i, int64_t ic, int64_t *nv, int64_t *ni) {
            return bf_marching_cubes(device, volume.data(), res, threshold, sc, v, vc, i, ic, nv, ni);
        });
    }

    // The same for the density a solver saved on its last save tick, read on
    // the device (for LBM::native_handle()).
    bf_compat::IsoMesh compute(bf_solver *solver, float threshold = 1.0f, float scale_x = 1.0f,
                               float scale_y = 1.0f, float scale_z = 1.0f) {
        const float sc[3] = {scale_x, scale_y, scale_z};
        return run([&](float *v, int64_t vc, uint32_t *i, int64_t ic, int64_t *nv, int64_t *ni) {
            return bf_isosurface(solver, threshold, sc, v, vc, i, ic, nv, ni);
        });
    }

    int32_t device = 0;

private:
    template <class IVec3>
    static void capacity(IVec3 r, int64_t &v, int64_t &i) {
        const int32_t res[3] = {(int32_t)r.x, (int32_t)r.y, (int32_t)r.z};
        if (bf_marching_cubes_capacity(res, &v, &i) != BF_OK)
            throw std::runtime_error("[MarchingCubes Error] resolution is not valid");
    }

    template <class F>
    static bf_compat::IsoMesh run(F call) {
        int64_t nv = 0, ni = 0;
        if (call(nullptr, 0, nullptr, 0, &nv, &ni) != BF_OK) throw std::runtime_error(bf_marching_cubes_last_error());
        bf_compat::IsoMesh m;
        m.vertices.resize((size_t)nv * 3);
        m.indices.resize((size_t)ni);
        if (call(m.vertices.data(), nv, m.indices.data(), ni, &nv, &ni) != BF_OK)
            throw std::runtime_error(bf_marching_cubes_last_error());
        return m;
    }
};
