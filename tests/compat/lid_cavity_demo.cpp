// Scenario code written the way the reference's demos are
// (BoltzmannFluids/Source/Demos/LBMDemo2D.cpp), compiled against the drop-in
// include/bf/LBM.h and run on the GPU: a lid-driven cavity with a moving top
// wall.  Writes the saved velocity/density field (float4 per cell) and the tick
// count to the file named by argv[1].
#include <cstdio>
#include <vector>

#include "bf/LBM.h"

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s out.bin ticks\n", argv[0]);
        return 2;
    }
    const int ticks = std::atoi(argv[2]);
    LBM solver;
    glm::ivec3 resolution(64, 64, 1);
    solver.clear_boundry_properties();
    solver.set_boundry_properties(1, glm::vec3(0, 0, 0));
    solver.set_boundry_properties(2, glm::vec3(1.6f, 0, 0) / 16.0f);
    solver.initialize_fields(
        [&](glm::ivec3 c, LBM::FluidProperties &p) {
            if (c.x == 0 || c.x == resolution.x - 1 || c.y == 0) p.boundry_id = 1;
            if (c.y == resolution.y - 1) p.boundry_id = 2;
        },
        resolution, 0.6f, false, false, true, VelocitySet::D2Q9, FloatingPointAccuracy::fp32, false);
    for (int t = 0; t + 1 < ticks; ++t) solver.iterate_time(0);  // paced loop of the UI/standalone mains
    solver.iterate_ticks(1);                                    // last tick saves the fields
    std::vector<float> vd = solver.get_velocity_density();
    std::vector<double> sums = solver.get_conservation_sums();
    FILE *f = std::fopen(argv[1], "wb");
    if (!f) return 3;
    const int32_t t = solver.get_total_ticks_elapsed();
    std::fwrite(&t, sizeof t, 1, f);
    std::fwrite(vd.data(), sizeof(float), vd.size(), f);
    std::fwrite(sums.data(), sizeof(double), sums.size(), f);
    std::fclose(f);
    std::printf("lid cavity: %d ticks, mass %.12f\n", t, sums[0]);
    return 0;
}
