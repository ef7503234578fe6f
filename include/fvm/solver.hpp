// solver.hpp -- C++ drop-in for the reference's CFDSolver (src/include/solver.cuh)
// and SimParams / ObstacleShape / CellType (src/include/common.h), implemented
// over the C ABI of <fvm/fvm.h> (libfvm_amd.so, HIP for gfx950).
//
// Same class, method names, signatures and call semantics as the reference:
// step(params) uses params.dt and advances params.t_current / params.step;
// errors print a message and exit the process, as the reference's CUDA_CHECK
// does (solver.cuh:8-19).  Header-only; link with -lfvm_amd.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>

#include "fvm.h"

// Physical constants (common.h:7-19) in the reference's SI units.
constexpr float GAMMA = 1.4f;
constexpr float R_GAS = 287.05f;
constexpr float CV = R_GAS / (GAMMA - 1.0f);
constexpr float CP = GAMMA * CV;
constexpr float PI = 3.14159265358979323846f;
constexpr float MU_REF = 1.716e-5f;
constexpr float T_REF = 273.15f;
constexpr float S_SUTHERLAND = 110.4f;
constexpr float PRANDTL_NUMBER = 0.71f;

enum class ObstacleShape { CIRCLE = 0, STAR = 1, DIAMOND = 2, CAPSULE = 3, TRIANGLE = 4, STARSHIP = 5 };

enum CellType : uint8_t { CELL_FLUID = 0, CELL_SOLID = 1, CELL_GHOST = 2, CELL_INFLOW = 3, CELL_OUTFLOW = 4 };

// SimParams (common.h:50-102), field for field, followed by the run-time unit
// system and the scheme selection of fvm_params (defaults reproduce the reference).
struct SimParams {
    int nx = 512;
    int ny = 256;
    float dx = 1.0f;
    float dy = 1.0f;
    float domain_width = 10.0f;
    float domain_height = 5.0f;
    float mach = 3.0f;
    float T_inf = 300.0f;
    float p_inf = 101325.0f;
    float rho_inf = 0.0f;
    float u_inf = 0.0f;
    float v_inf = 0.0f;
    float c_inf = 0.0f;
    float obstacle_x = 2.5f;
    float obstacle_y = 2.5f;
    float obstacle_r = 0.5f;
    float obstacle_rotation = 0.0f;
    ObstacleShape obstacle_shape = ObstacleShape::CIRCLE;
    float wing_rotation = 0.0f;
    float cfl = 0.5f;
    float dt = 1e-6f;
    float t_current = 0.0f;
    int step = 0;
    bool enable_viscosity = false;
    float cfl_visc = 0.4f;
    bool adiabatic_wall = true;
    float T_wall = 300.0f;
    bool paused = true;
    // run-time unit system (defaults: src/solver.cu:10-14, 39-41, 52; common.h:7-19)
    float gamma = GAMMA, r_gas = R_GAS, mu_ref = MU_REF, t_ref = T_REF, s_sutherland = S_SUTHERLAND;
    float prandtl = PRANDTL_NUMBER;
    float min_density = 1e-4f, min_pressure = 1e-2f, max_temperature = 5000.0f, min_temperature = 50.0f;
    float low_density_threshold = 0.01f, max_velocity_limit = 5000.0f, max_ghost_velocity = 2500.0f;
    float slip_reference_velocity = 100.0f;
    // additions
    int flux_scheme = FVM_FLUX_HLLC, limiter = FVM_LIMITER_MINMOD, wall_mode = FVM_WALL_REFERENCE;
    int precision = FVM_PRECISION_FP32, dt_mode = FVM_DT_HOST;
    int noise_field = FVM_NOISE_NONE;
    float noise_amplitude = 0.0f;
    uint32_t noise_seed = 0;

    // computeDerived (common.h:93-101)
    void computeDerived() {
        rho_inf = p_inf / (r_gas * T_inf);
        c_inf = std::sqrt(gamma * r_gas * T_inf);
        u_inf = mach * c_inf;
        dx = domain_width / (float)nx;
        dy = domain_height / (float)ny;
        obstacle_y = domain_height / 2.0f;
    }

    fvm_params to_c() const {
        fvm_params p;
        fvm_params_default(&p);
        p.nx = nx; p.ny = ny; p.dx = dx; p.dy = dy;
        p.domain_width = domain_width; p.domain_height = domain_height;
        p.mach = mach; p.T_inf = T_inf; p.p_inf = p_inf; p.rho_inf = rho_inf;
        p.u_inf = u_inf; p.v_inf = v_inf; p.c_inf = c_inf;
        p.obstacle_x = obstacle_x; p.obstacle_y = obstacle_y; p.obstacle_r = obstacle_r;
        p.obstacle_rotation = obstacle_rotation; p.obstacle_shape = (int32_t)obstacle_shape;
        p.wing_rotation = wing_rotation; p.cfl = cfl; p.dt = dt; p.t_current = t_current; p.step = step;
        p.enable_viscosity = enable_viscosity ? 1 : 0; p.cfl_visc = cfl_visc;
        p.adiabatic_wall = adiabatic_wall ? 1 : 0; p.T_wall = T_wall; p.paused = paused ? 1 : 0;
        p.gamma = gamma; p.r_gas = r_gas; p.mu_ref = mu_ref; p.t_ref = t_ref; p.s_sutherland = s_sutherland;
        p.prandtl = prandtl; p.min_density = min_density; p.min_pressure = min_pressure;
        p.max_temperature = max_temperature; p.min_temperature = min_temperature;
        p.low_density_threshold = low_density_threshold; p.max_velocity_limit = max_velocity_limit;
        p.max_ghost_velocity = max_ghost_velocity; p.slip_reference_velocity = slip_reference_velocity;
        p.flux_scheme = flux_scheme; p.limiter = limiter; p.wall_mode = wall_mode; p.precision = precision;
        p.dt_mode = dt_mode; p.noise_field = noise_field; p.noise_amplitude = noise_amplitude;
        p.noise_seed = noise_seed;
        return p;
    }
};

#define FVM_CHECK(call)                                                                       \
    do {                                                                                      \
        int fvm_rc_ = (call);                                                                 \
        if (fvm_rc_ != FVM_OK) {                                                              \
            std::fprintf(stderr, "fvm call failed '%s' in line %d : %s.\n", __FILE__, __LINE__, \
                         fvm_last_error());                                                   \
            std::exit(EXIT_FAILURE);                                                          \
        }                                                                                     \
    } while (0)

class CFDSolver {
public:
    CFDSolver() { FVM_CHECK(fvm_create(0, &h_)); }
    ~CFDSolver() { fvm_destroy(h_); }
    CFDSolver(const CFDSolver&) = delete;
    CFDSolver& operator=(const CFDSolver&) = delete;

    void initialize(const SimParams& params) {
        fvm_params p = params.to_c();
        FVM_CHECK(fvm_initialize(h_, &p));
        nx_ = params.nx;
        ny_ = params.ny;
    }
    void resize(int nx, int ny) {
        FVM_CHECK(fvm_resize(h_, nx, ny));
        if (nx > 0 && ny > 0) { nx_ = nx; ny_ = ny; }
    }
    void step(SimParams& params) {
        fvm_params p = params.to_c();
        FVM_CHECK(fvm_step(h_, &p));
        params.t_current = p.t_current;
        params.step = p.step;
    }
    void reset(const SimParams& params) {
        fvm_params p = params.to_c();
        FVM_CHECK(fvm_reset(h_, &p));
        FVM_CHECK(fvm_synchronize(h_));
    }
    void updateObstacleGeometry(const SimParams& params) {
        fvm_params p = params.to_c();
        FVM_CHECK(fvm_update_obstacle(h_, &p));
    }
    void getCellTypes(uint8_t* host_types) { FVM_CHECK(fvm_get_cell_types(h_, host_types)); }
    void convertCellTypesToDevice(float* devOutput) { FVM_CHECK(fvm_field_to_device(h_, FVM_FIELD_CELL_TYPE, devOutput)); }
    void computeTemperatureToDevice(float* d) { FVM_CHECK(fvm_field_to_device(h_, FVM_FIELD_TEMPERATURE, d)); }
    void computePressureToDevice(float* d) { FVM_CHECK(fvm_field_to_device(h_, FVM_FIELD_PRESSURE, d)); }
    void computeDensityToDevice(float* d) { FVM_CHECK(fvm_field_to_device(h_, FVM_FIELD_DENSITY, d)); }
    void computeVelocityMagToDevice(float* d) { FVM_CHECK(fvm_field_to_device(h_, FVM_FIELD_VELOCITY_MAG, d)); }
    void computeMachToDevice(float* d) { FVM_CHECK(fvm_field_to_device(h_, FVM_FIELD_MACH, d)); }
    // Velocity-arrow vertices for the OpenGL overlay (src/solver.cu:3809-3845)
    // into the caller's mapped VBO; returns the number of vertices written
    // (at most max_vertices).
    int generateVectorArrows(float* dev_vertex_data, int max_vertices, int step, float u_inf, float max_arrow_length,
                             float arrow_head_angle, float arrow_head_length) {
        int n = 0;
        FVM_CHECK(fvm_generate_vector_arrows(h_, dev_vertex_data, max_vertices, step, u_inf, max_arrow_length,
                                             arrow_head_angle, arrow_head_length, &n));
        return n;
    }
    int getNx() const { return nx_; }
    int getNy() const { return ny_; }
    float computeStableTimeStep(const SimParams& params) {
        fvm_params p = params.to_c();
        float dt = 0;
        FVM_CHECK(fvm_compute_stable_dt(h_, &p, &dt));
        return dt;
    }
    void queueTimeStepComputation(const SimParams& params) {
        fvm_params p = params.to_c();
        FVM_CHECK(fvm_queue_dt(h_, &p));
    }
    float readTimeStepResult(const SimParams& params) const {
        fvm_params p = params.to_c();
        float dt = 0;
        FVM_CHECK(fvm_read_dt(h_, &p, &dt));
        return dt;
    }
    float getMaxTemperature() {
        float v = 0;
        FVM_CHECK(fvm_max_temperature(h_, &v));
        return v;
    }
    float getMaxMach() {
        float v = 0;
        FVM_CHECK(fvm_max_mach(h_, &v));
        return v;
    }
    static void getGPUMemoryInfo(size_t& freeMem, size_t& totalMem) { FVM_CHECK(fvm_gpu_memory_info(0, &freeMem, &totalMem)); }
    size_t getSimulationMemoryUsage() {
        size_t b = 0;
        FVM_CHECK(fvm_memory_usage(h_, &b));
        return b;
    }
    // Replaces the callers' cudaDeviceSynchronize() (main.cpp:365, kernel_benchmark.cu:103,113,132).
    void synchronize() { FVM_CHECK(fvm_synchronize(h_)); }
    fvm_solver* handle() { return h_; }

private:
    fvm_solver* h_ = nullptr;
    int nx_ = 1024;
    int ny_ = 512;
};
