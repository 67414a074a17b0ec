// LBM.h — header-only C++ drop-in for the reference solver class `LBM`
// (BoltzmannFluids/Source/LBM/LBM.h:26-332), implemented on the C ABI of bf.h.
//
// Scenario code written against the reference (demo lambdas, the Qt
// SimulationController, the standalone main loop) keeps its calls:
// initialize_fields(lambda, resolution, tau, periodic..., VelocitySet,
// FloatingPointAccuracy, multiphase), set_boundry_properties(...),
// iterate_time(target_tps), get_total_ticks_elapsed(), the public knobs
// (velocity_limit, velocity_limit_extreme, is_collide_esoteric, ...).
// The GL texture getters are replaced by host readbacks (get_velocity_density
// / get_boundries / get_force_temperature), whose layouts match the reference
// textures (RGBA32F u.xyz+rho, R8 id, RGBA32F F.xyz+T; collide.comp:784-792).
// The render2d_* / render3d_* calls keep their signatures and hand the field
// to a renderer that the application registers once (set_field_renderer), so
// the call sites of the UI viewport and the standalone loop compile unchanged.
//
// Link with libbfamd.so.  Uses GLM when it is on the include path, otherwise a
// minimal stand-in with the few vector types the interface needs.
#pragma once

#include <chrono>
#include <cmath>
#include <cstdint>
#include <filesystem>
#include <functional>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "bf.h"

#if defined(__has_include)
#if __has_include(<glm/glm.hpp>)
#include <glm/glm.hpp>
#define BF_HAVE_GLM 1
#endif
#endif

#ifndef BF_HAVE_GLM
namespace glm {
struct ivec3 {
    int x = 0, y = 0, z = 0;
    ivec3() = default;
    ivec3(int a, int b, int c) : x(a), y(b), z(c) {}
    bool operator==(const ivec3 &o) const { return x == o.x && y == o.y && z == o.z; }
    bool operator!=(const ivec3 &o) const { return !(*this == o); }
};
struct vec3 {
    float x = 0, y = 0, z = 0;
    vec3() = default;
    explicit vec3(float v) : x(v), y(v), z(v) {}
    vec3(float a, float b, float c) : x(a), y(b), z(c) {}
    vec3 operator/(float s) const { return vec3(x / s, y / s, z / s); }
    vec3 operator*(float s) const { return vec3(x * s, y * s, z * s); }
    bool operator==(const vec3 &o) const { return x == o.x && y == o.y && z == o.z; }
    bool operator!=(const vec3 &o) const { return !(*this == o); }
};
}  // namespace glm
#endif

// VelocitySet.h:13-23 and FloatingPointAccuracy.h:6-9 (bf16 is an extension).
enum SimplifiedVelocitySet { D2Q5 = 0, D3Q7 = 1 };
enum VelocitySet { D2Q9 = 2, D3Q15 = 3, D3Q19 = 4, D3Q27 = 5 };
enum FloatingPointAccuracy { fp16 = 0, fp32 = 1, bf16 = 2 };

class LBM {
public:
    constexpr static uint32_t not_a_boundry = 0;
    constexpr static uint32_t max_boundry_count = 255;
    constexpr static float referance_temperature = 1.0f;
    constexpr static float referance_boundry_density = 1.0f;

    struct FluidProperties {
        glm::vec3 velocity = glm::vec3(0);
        glm::vec3 force = glm::vec3(0);
        float density = 1;
        float temperature = referance_temperature;
        uint32_t boundry_id = not_a_boundry;
    };

    LBM() = default;
    LBM(const LBM &) = delete;
    LBM &operator=(const LBM &) = delete;
    ~LBM() { bf_destroy(solver_); }

    // --- simulation controls (LBM.cpp:220-250) ---------------------------
    void iterate_time(float target_tick_per_second = 0) {
        require_solver("iterate_time");
        if (first_iteration) {
            simulation_begin = std::chrono::system_clock::now();
            last_visual_update = std::chrono::system_clock::now();
        }
        const size_t targeted = (size_t)(target_tick_per_second * get_total_time_elapsed().count() / 1000.0f);
        if (target_tick_per_second <= 0 || total_ticks_elapsed < targeted || first_iteration) {
            const auto since = std::chrono::system_clock::now() - last_visual_update;
            const bool visuals = first_iteration ||
                std::chrono::duration_cast<std::chrono::milliseconds>(since).count() > 1000.0 / 60;
            if (visuals) last_visual_update = std::chrono::system_clock::now();
            push_params();
            check(bf_step(solver_, 1, visuals ? BF_STEP_SAVE_MACRO : 0u), "iterate_time");
            total_ticks_elapsed++;
        }
        first_iteration = false;
    }

    // Advance n ticks at once (no pacing); the macroscopic fields are saved on the last.
    void iterate_ticks(int64_t n) {
        require_solver("iterate_ticks");
        push_params();
        check(bf_step(solver_, n, BF_STEP_SAVE_MACRO), "iterate_ticks");
        total_ticks_elapsed += (size_t)n;
        first_iteration = false;
    }

    int32_t get_total_ticks_elapsed() { return (int32_t)total_ticks_elapsed; }
    std::chrono::duration<double, std::milli> get_total_time_elapsed() {
        return std::chrono::system_clock::now() - simulation_begin;
    }
    glm::ivec3 get_resolution() { return resolution; }
    int32_t get_velocity_set_vector_count() { return bf_velocity_count(velocity_set); }

    // --- field initialisation (LBM.cpp:415-506) --------------------------
    void initialize_fields(std::function<void(glm::ivec3, FluidProperties &)> initialization_lambda,
                           glm::ivec3 resolution, float relaxation_time, bool periodic_x = true,
                           bool periodic_y = true, bool periodic_z = true,
                           VelocitySet velocity_set = VelocitySet::D2Q9,
                           FloatingPointAccuracy fp_accuracy = FloatingPointAccuracy::fp32,
                           bool is_flow_multiphase = false) {
        bf_destroy(solver_);
        solver_ = nullptr;
        bf_config cfg{};
        cfg.resolution[0] = resolution.x;
        cfg.resolution[1] = resolution.y;
        cfg.resolution[2] = resolution.z;
        cfg.relaxation_time = relaxation_time;
        cfg.periodic[0] = periodic_x;
        cfg.periodic[1] = periodic_y;
        cfg.periodic[2] = periodic_z;
        cfg.velocity_set = velocity_set;
        cfg.storage = fp_accuracy == fp32 ? BF_STORE_FP32 : (fp_accuracy == fp16 ? BF_STORE_FP16 : BF_STORE_BF16);
        cfg.multiphase = is_flow_multiphase;
        cfg.boundary_mode = compat_fullway_bounce_back ? BF_BB_FULLWAY_COMPAT : BF_BB_HALFWAY;
        cfg.device = device;
        if (bf_create(&cfg, &solver_) != BF_OK) fatal(bf_last_error(nullptr));
        this->resolution = resolution;
        this->relaxation_time = relaxation_time;
        this->velocity_set = velocity_set;
        this->floating_point_accuracy = fp_accuracy;
        this->periodic_x = periodic_x;
        this->periodic_y = periodic_y;
        this->periodic_z = periodic_z;
        this->is_flow_multiphase = is_flow_multiphase;
        this->thermal_relaxation_time = 2;  // LBM.cpp:461

        const size_t n = (size_t)resolution.x * resolution.y * resolution.z;
        std::vector<float> rho(n), u(3 * n), force(3 * n), temperature(n);
        std::vector<uint8_t> ids(n);
        for (int32_t z = 0; z < resolution.z; z++)
            for (int32_t y = 0; y < resolution.y; y++)
                for (int32_t x = 0; x < resolution.x; x++) {
                    FluidProperties p;
                    initialization_lambda(glm::ivec3(x, y, z), p);
                    const size_t c = ((size_t)z * resolution.y + y) * resolution.x + x;
                    if (p.boundry_id > max_boundry_count)
                        fatal("[LBM Error] initialize_fields(): boundry_id greater than 255");
                    rho[c] = p.density;
                    u[3 * c] = p.velocity.x; u[3 * c + 1] = p.velocity.y; u[3 * c + 2] = p.velocity.z;
                    force[3 * c] = p.force.x; force[3 * c + 1] = p.force.y; force[3 * c + 2] = p.force.z;
                    temperature[c] = p.temperature;
                    ids[c] = (uint8_t)p.boundry_id;
                }
        for (size_t id = 1; id < objects_.size(); ++id) {
            const Object &o = objects_[id];
            if (!o.defined) continue;
            const float v[3] = {o.velocity.x, o.velocity.y, o.velocity.z};
            const float w[3] = {o.angular.x, o.angular.y, o.angular.z};
            const float c[3] = {o.center.x, o.center.y, o.center.z};
            check(bf_set_object(solver_, (uint32_t)id, v, w, c, o.temperature, o.effective_density), "set_boundry_properties");
        }
        push_params();
        check(bf_init_fields(solver_, rho.data(), u.data(), force.data(), temperature.data(), ids.data()),
              "initialize_fields");
        bf_info info{};
        bf_get_info(solver_, &info);
        constant_force = glm::vec3(force[0], force[1], force[2]);
        is_forcing_scheme = info.forcing != 0;
        is_force_field_constant = info.forcing != 2;
        is_flow_thermal = info.thermal != 0;
        total_ticks_elapsed = 0;
        first_iteration = true;
    }

    // --- boundary objects (LBM.cpp:1015-1070) ----------------------------
    void set_boundry_properties(uint32_t boundry_id, glm::vec3 velocity_translational, glm::vec3 velocity_angular,
                                glm::vec3 center_of_mass, float temperature = referance_temperature,
                                float effective_density = referance_boundry_density) {
        if (boundry_id > max_boundry_count)
            fatal("[LBM Error] LBM::set_boundry_properties() is called but boundry_id is greater than maximum(255)");
        if (boundry_id == 0)
            fatal("[LBM Error] LBM::set_boundry_properties() is called but boundry_id(0) is defined to be fluid, it "
                  "cannot be treated as an object");
        if (boundry_id >= objects_.size()) objects_.resize(boundry_id + 1);
        objects_[boundry_id] = Object{true, velocity_translational, velocity_angular, center_of_mass, temperature,
                                      effective_density};
        if (solver_) {
            const float v[3] = {velocity_translational.x, velocity_translational.y, velocity_translational.z};
            const float w[3] = {velocity_angular.x, velocity_angular.y, velocity_angular.z};
            const float c[3] = {center_of_mass.x, center_of_mass.y, center_of_mass.z};
            check(bf_set_object(solver_, boundry_id, v, w, c, temperature, effective_density), "set_boundry_properties");
        }
    }
    void set_boundry_properties(uint32_t boundry_id, glm::vec3 velocity_translational,
                                float temperature = referance_temperature,
                                float effective_density = referance_boundry_density) {
        set_boundry_properties(boundry_id, velocity_translational, glm::vec3(0), glm::vec3(0), temperature,
                               effective_density);
    }
    void set_boundry_properties(uint32_t boundry_id, float temperature,
                                float effective_density = referance_boundry_density) {
        set_boundry_properties(boundry_id, glm::vec3(0), glm::vec3(0), glm::vec3(0), temperature, effective_density);
    }
    void clear_boundry_properties() {
        objects_.clear();
        if (solver_) check(bf_clear_objects(solver_), "clear_boundry_properties");
    }

    // --- host readback of the macroscopic fields (replaces get_*_texture) ---
    std::vector<float> get_velocity_density() { return read_floats(BF_FIELD_VELOCITY_DENSITY); }
    std::vector<float> get_force_temperature() { return read_floats(BF_FIELD_FORCE_TEMPERATURE); }
    std::vector<uint8_t> get_boundries() {
        require_solver("get_boundries");
        std::vector<uint8_t> out(bf_field_bytes(solver_, BF_FIELD_BOUNDARY));
        check(bf_read_field(solver_, BF_FIELD_BOUNDARY, out.data(), out.size()), "get_boundries");
        return out;
    }
    // FP64 sums over fluid cells: rho, raw momentum (3), temperature.
    std::vector<double> get_conservation_sums() {
        require_solver("get_conservation_sums");
        std::vector<double> out(5);
        check(bf_conservation(solver_, out.data()), "get_conservation_sums");
        return out;
    }

    // --- rendering (LBM.h:85-95) ------------------------------------------
    // The reference renders its GL textures with GraphicsCortex programs.  Here
    // each render call passes the field to the registered renderer, which reads
    // it with get_velocity_density() / get_boundries() / get_force_temperature()
    // (e.g. into a glTexImage3D upload) and draws it; `camera` is the caller's
    // camera object, passed through untouched (null for 2-D calls).
    enum class RenderedField { density, velocity, boundries, forces, temperature };
    using FieldRenderer =
        std::function<void(LBM &solver, RenderedField field, bool three_d, void *camera, int32_t sample_count)>;
    void set_field_renderer(FieldRenderer renderer) { renderer_ = std::move(renderer); }

    void render2d_density() { render(RenderedField::density, false, nullptr, 0, "render2d_density"); }
    void render2d_velocity() { render(RenderedField::velocity, false, nullptr, 0, "render2d_velocity"); }
    void render2d_boundries() { render(RenderedField::boundries, false, nullptr, 0, "render2d_boundries"); }
    void render2d_forces() { render(RenderedField::forces, false, nullptr, 0, "render2d_forces"); }
    void render2d_temperature() { render(RenderedField::temperature, false, nullptr, 0, "render2d_temperature"); }
    template <class Camera>
    void render3d_density(Camera &camera, int32_t sample_count = 128) {
        render(RenderedField::density, true, &camera, sample_count, "render3d_density");
    }
    template <class Camera>
    void render3d_velocity(Camera &camera, int32_t sample_count = 128) {
        render(RenderedField::velocity, true, &camera, sample_count, "render3d_velocity");
    }
    template <class Camera>
    void render3d_boundries(Camera &camera, int32_t sample_count = 128) {
        render(RenderedField::boundries, true, &camera, sample_count, "render3d_boundries");
    }
    template <class Camera>
    void render3d_forces(Camera &camera, int32_t sample_count = 128) {
        render(RenderedField::forces, true, &camera, sample_count, "render3d_forces");
    }
    template <class Camera>
    void render3d_temperature(Camera &camera, int32_t sample_count = 128) {
        render(RenderedField::temperature, true, &camera, sample_count, "render3d_temperature");
    }

    // --- low level accessors (LBM.h:105-120) -----------------------------
    float get_relaxation_time() { return relaxation_time; }
    VelocitySet get_velocity_set() { return velocity_set; }
    int32_t get_dimentionality() { return velocity_set == D2Q9 ? 2 : 3; }
    FloatingPointAccuracy get_floating_point_accuracy() { return floating_point_accuracy; }
    bool get_periodic_boundry_x() { return periodic_x; }
    bool get_periodic_boundry_y() { return periodic_y; }
    bool get_periodic_boundry_z() { return periodic_z; }
    bool get_is_forcing_scheme() { return is_forcing_scheme; }
    bool get_is_force_field_constant() { return is_force_field_constant; }
    bool get_is_flow_multiphase() { return is_flow_multiphase; }
    glm::vec3 get_constant_force() { return constant_force; }
    float get_intermolecular_interaction_strength() { return intermolecular_interaction_strength; }

    void set_relaxation_time(float value) { relaxation_time = value; }
    void set_constant_force(glm::vec3 value) {
        constant_force = value;
        if (solver_) {
            const float f[3] = {value.x, value.y, value.z};
            check(bf_set_param(solver_, BF_PARAM_CONSTANT_FORCE, f, 3), "set_constant_force");
        }
    }
    void set_intermolecular_interaction_strength(float value) { intermolecular_interaction_strength = value; }

    // checkpoint files (LBM.h:123-126 of the reference, which declares them
    // without a body): <save_path>/tick_<n>.bfmacro / .bfmeso
    void save_current_tick_macroscopic(std::filesystem::path save_path) {
        require_solver("save_current_tick_macroscopic");
        check(bf_save_macroscopic(solver_, save_path.string().c_str()), "save_current_tick_macroscopic");
    }
    void save_current_tick_mesoscropic(std::filesystem::path save_path) {
        require_solver("save_current_tick_mesoscropic");
        check(bf_save_mesoscopic(solver_, save_path.string().c_str()), "save_current_tick_mesoscropic");
    }
    void load_tick_macroscopic(std::filesystem::path save_path, int32_t target_tick) {
        require_solver("load_tick_macroscopic");
        check(bf_load_macroscopic(solver_, save_path.string().c_str(), target_tick), "load_tick_macroscopic");
        total_ticks_elapsed = (size_t)target_tick;
        pull_params();
    }
    void load_tick_mesoscropic(std::filesystem::path save_path, int32_t target_tick) {
        require_solver("load_tick_mesoscropic");
        check(bf_load_mesoscopic(solver_, save_path.string().c_str(), target_tick), "load_tick_mesoscropic");
        total_ticks_elapsed = (size_t)target_tick;
        pull_params();
    }

    // The C handle, for callers that use the C ABI directly (interop, tests).
    bf_solver *native_handle() { return solver_; }

    // --- public knobs of the reference (LBM.h:139-145, 242-257) ----------
    bool is_lattice_texture3d = false;        // storage layout is chosen by the library
    bool is_collide_esoteric = true;          // the update is always in place (AA pattern)
    bool compat_fullway_bounce_back = false;  // reproduce the reference's implicit full-way walls
    bool shan_chen_synchronous = false;       // extension: current-tick neighbour psi (default: the reference's lag)
    float velocity_limit = 0.25f;
    float velocity_limit_extreme = 0.30f;
    float thermal_relaxation_time = 2.0f;
    float thermal_expension_coeficient = 0.5f;
    float intermolecular_interaction_strength = -6.0f;
    float referance_temperature_value = 1.0f;  // T_ref of the Boussinesq term (collide.comp:713)
    int32_t device = 0;

private:
    struct Object {
        bool defined = false;
        glm::vec3 velocity, angular, center;
        float temperature = referance_temperature;
        float effective_density = referance_boundry_density;
    };

    [[noreturn]] static void fatal(const std::string &msg) {
        std::cout << msg << std::endl;
        throw std::runtime_error(msg);
    }
    void check(bf_status st, const char *what) {
        if (st != BF_OK) fatal(std::string("[LBM Error] ") + what + ": " + bf_last_error(solver_));
    }
    void require_solver(const char *what) {
        if (!solver_) fatal(std::string("[LBM Error] ") + what + " is called before initialize_fields()");
    }
    // the parameters a loaded checkpoint carries become the object's
    void pull_params() {
        float v = 0.0f;
        check(bf_get_param(solver_, BF_PARAM_RELAXATION_TIME, &v, 1), "get_param"); relaxation_time = v;
        check(bf_get_param(solver_, BF_PARAM_INTERACTION_STRENGTH, &v, 1), "get_param");
        intermolecular_interaction_strength = v;
        check(bf_get_param(solver_, BF_PARAM_THERMAL_EXPANSION, &v, 1), "get_param"); thermal_expension_coeficient = v;
        check(bf_get_param(solver_, BF_PARAM_THERMAL_RELAXATION_TIME, &v, 1), "get_param"); thermal_relaxation_time = v;
        check(bf_get_param(solver_, BF_PARAM_REFERENCE_TEMPERATURE, &v, 1), "get_param"); referance_temperature_value = v;
        check(bf_get_param(solver_, BF_PARAM_VELOCITY_LIMIT, &v, 1), "get_param"); velocity_limit = v;
        check(bf_get_param(solver_, BF_PARAM_VELOCITY_LIMIT_EXTREME, &v, 1), "get_param"); velocity_limit_extreme = v;
        check(bf_get_param(solver_, BF_PARAM_SHAN_CHEN_SYNCHRONOUS, &v, 1), "get_param"); shan_chen_synchronous = v != 0.0f;
    }
    void push_params() {
        const float values[][1] = {{relaxation_time}, {intermolecular_interaction_strength},
                                   {thermal_expension_coeficient}, {thermal_relaxation_time},
                                   {referance_temperature_value}, {velocity_limit}, {velocity_limit_extreme},
                                   {shan_chen_synchronous ? 1.0f : 0.0f}};
        const int keys[] = {BF_PARAM_RELAXATION_TIME, BF_PARAM_INTERACTION_STRENGTH, BF_PARAM_THERMAL_EXPANSION,
                            BF_PARAM_THERMAL_RELAXATION_TIME, BF_PARAM_REFERENCE_TEMPERATURE,
                            BF_PARAM_VELOCITY_LIMIT, BF_PARAM_VELOCITY_LIMIT_EXTREME,
                            BF_PARAM_SHAN_CHEN_SYNCHRONOUS};
        for (size_t k = 0; k < sizeof(keys) / sizeof(keys[0]); ++k) check(bf_set_param(solver_, keys[k], values[k], 1), "set_param");
    }
    void render(RenderedField field, bool three_d, void *camera, int32_t sample_count, const char *what) {
        require_solver(what);
        if (!renderer_) fatal(std::string("[LBM Error] ") + what + " is called but no renderer is registered (set_field_renderer)");
        renderer_(*this, field, three_d, camera, sample_count);
    }
    std::vector<float> read_floats(int field) {
        require_solver("readback");
        std::vector<float> out(bf_field_bytes(solver_, field) / sizeof(float));
        check(bf_read_field(solver_, field, out.data(), out.size() * sizeof(float)), "readback");
        return out;
    }

    bf_solver *solver_ = nullptr;
    FieldRenderer renderer_;
    std::vector<Object> objects_;
    bool first_iteration = true;
    size_t total_ticks_elapsed = 0;
    std::chrono::time_point<std::chrono::system_clock> simulation_begin = std::chrono::system_clock::now();
    std::chrono::time_point<std::chrono::system_clock> last_visual_update = std::chrono::system_clock::now();
    VelocitySet velocity_set = D2Q9;
    FloatingPointAccuracy floating_point_accuracy = fp32;
    glm::ivec3 resolution = glm::ivec3(0, 0, 0);
    float relaxation_time = 0.53f;
    bool periodic_x = true, periodic_y = true, periodic_z = true;
    bool is_forcing_scheme = false, is_force_field_constant = true, is_flow_thermal = false;
    bool is_flow_multiphase = false;
    glm::vec3 constant_force = glm::vec3(0);
};
