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

    // the display switch of the standalone loop and the UI viewport
    // (Mains/main.cpp:294-313, Viewport3D.cpp:329-346), unchanged, with a
    // renderer registered once: it receives the field and the caller's camera
    struct Camera {
        int id = 7;
    } camera;
    int renders = 0, bad = 0;
    double rendered_density = 0.0;
    solver.set_field_renderer([&](LBM &s, LBM::RenderedField field, bool three_d, void *cam, int32_t samples) {
        ++renders;
        if (three_d && (static_cast<Camera *>(cam)->id != 7 || samples != 128)) ++bad;
        if (field == LBM::RenderedField::density) {
            const std::vector<float> v = s.get_velocity_density();
            for (size_t c = 0; c < v.size() / 4; ++c) rendered_density += v[4 * c + 3];
        }
    });
    for (int display_mode = 1; display_mode <= 5; ++display_mode) {
        switch (display_mode) {
        case 1:
            if (solver.get_dimentionality() == 3) solver.render3d_density(camera);
            else solver.render2d_density();
            break;
        case 2:
            if (solver.get_dimentionality() == 3) solver.render3d_boundries(camera);
            else solver.render2d_boundries();
            break;
        case 3:
            if (solver.get_dimentionality() == 3) solver.render3d_velocity(camera);
            else solver.render2d_velocity();
            break;
        case 4:
            if (solver.get_dimentionality() == 3) solver.render3d_forces(camera);
            else solver.render2d_forces();
            break;
        case 5:
            if (solver.get_dimentionality() == 3) solver.render3d_temperature(camera);
            else solver.render2d_temperature();
            break;
        }
    }
    solver.render3d_velocity(camera);  // the 3-D form passes the camera through
    std::printf("renders %d bad %d rendered density %.6f\n", renders, bad, rendered_density);
    if (renders != 6 || bad != 0) return 4;

    // checkpoint paths as the reference declares them (std::filesystem::path)
    if (argc > 3) {
        const std::filesystem::path dir(argv[3]);
        solver.save_current_tick_macroscopic(dir);
        solver.load_tick_macroscopic(dir, solver.get_total_ticks_elapsed());
    }
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
