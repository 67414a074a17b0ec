// Scenario code in the style of the reference's 3-D demos
// (BoltzmannFluids/Source/Demos/LBMDemo3D.cpp), compiled against the drop-ins
// include/bf/LBM.h and include/bf/MarchingCubes.h and run on the GPU: a
// Shan-Chen droplet relaxes, then its surface is extracted twice — from the
// density resident on the device and from the host copy of the saved field.
// Writes to argv[1]: the density (float per cell), then for each mesh the
// vertex and index counts followed by the arrays.
#include <cmath>
#include <cstdio>
#include <vector>

#include "bf/LBM.h"
#include "bf/MarchingCubes.h"

static void write_mesh(FILE *f, const bf_compat::IsoMesh &m) {
    const int64_t nv = (int64_t)m.vertices.size() / 3, ni = (int64_t)m.indices.size();
    std::fwrite(&nv, sizeof nv, 1, f);
    std::fwrite(&ni, sizeof ni, 1, f);
    std::fwrite(m.vertices.data(), sizeof(float), m.vertices.size(), f);
    std::fwrite(m.indices.data(), sizeof(uint32_t), m.indices.size(), f);
}

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s out.bin ticks\n", argv[0]);
        return 2;
    }
    const int ticks = std::atoi(argv[2]);
    const int n = 32;
    LBM solver;
    solver.set_intermolecular_interaction_strength(-5.0f);
    solver.initialize_fields(
        [&](glm::ivec3 c, LBM::FluidProperties &p) {
            const float dx = c.x - 15.5f, dy = c.y - 15.5f, dz = c.z - 15.5f;
            const float r = std::sqrt(dx * dx + dy * dy + dz * dz);
            p.density = 0.12f + 1.77f * 0.5f * (1.0f - std::tanh((r - 8.0f) / 1.5f));
        },
        glm::ivec3(n, n, n), 1.0f, true, true, true, VelocitySet::D3Q19, FloatingPointAccuracy::fp32, true);
    solver.iterate_ticks(ticks);
    std::vector<float> vd = solver.get_velocity_density();
    std::vector<float> rho((size_t)n * n * n);
    for (size_t c = 0; c < rho.size(); ++c) rho[c] = vd[4 * c + 3];
    MarchingCubes mc;
    const auto on_device = mc.compute(solver.native_handle(), 1.0f, (float)n, (float)n, (float)n);
    const auto from_host = mc.compute(rho, glm::ivec3(n, n, n), 1.0f, (float)n, (float)n, (float)n);
    FILE *f = std::fopen(argv[1], "wb");
    if (!f) return 3;
    std::fwrite(rho.data(), sizeof(float), rho.size(), f);
    write_mesh(f, on_device);
    write_mesh(f, from_host);
    std::fclose(f);
    std::printf("droplet: %lld vertices, %lld triangles (index buffer bound %zu bytes)\n",
                (long long)on_device.vertices.size() / 3, (long long)on_device.indices.size() / 3,
                mc.get_index_buffer_needed_size_in_bytes(glm::ivec3(n, n, n)));
    return 0;
}
