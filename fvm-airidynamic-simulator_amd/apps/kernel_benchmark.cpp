// kernel_benchmark.cpp -- headless timing harness with the command line and
// schedule of the reference's KernelBenchmark (src/test/kernel_benchmark.cu:33-145):
// SI Mach-3 viscous case, circle r = 0.5 m in a 10 x 5 m domain, dt computed once,
// 10 steps to develop the flow, 3 warm-up steps, 5 timed steps; prints ms/step.
#include <chrono>
#include <cstdlib>
#include <iomanip>
#include <iostream>

#include "fvm/solver.hpp"

static constexpr int DEFAULT_NX = 1024;
static constexpr int DEFAULT_NY = 512;
static constexpr int WARMUP_ITERATIONS = 3;
static constexpr int BENCHMARK_ITERATIONS = 5;

int main(int argc, char* argv[]) {
    int nx = (argc > 1) ? std::atoi(argv[1]) : DEFAULT_NX;
    int ny = (argc > 2) ? std::atoi(argv[2]) : DEFAULT_NY;
    if (nx <= 0 || ny <= 0) {
        std::cerr << "Usage: " << argv[0] << " [nx=" << DEFAULT_NX << "] [ny=" << DEFAULT_NY << "]" << std::endl;
        return 1;
    }
    std::cout << "Grid size:  " << nx << " x " << ny << " = " << (long long)nx * ny << " cells" << std::endl;
    size_t freeMem = 0, totalMem = 0;
    CFDSolver::getGPUMemoryInfo(freeMem, totalMem);
    std::cout << "VRAM:       " << std::fixed << std::setprecision(0) << totalMem / (1024.0 * 1024.0) << " MB" << std::endl;

    SimParams params;
    params.nx = nx;
    params.ny = ny;
    params.mach = 3.0f;
    params.T_inf = 300.0f;
    params.p_inf = 101325.0f;
    params.enable_viscosity = true;
    params.cfl = 0.5f;
    params.cfl_visc = 0.4f;
    params.obstacle_shape = ObstacleShape::CIRCLE;
    params.obstacle_r = 0.5f;
    params.domain_width = 10.0f;
    params.domain_height = 5.0f;
    params.computeDerived();

    CFDSolver solver;
    solver.initialize(params);
    solver.reset(params);
    params.dt = solver.computeStableTimeStep(params);
    std::cout << "Time step:  dt=" << std::scientific << std::setprecision(2) << params.dt << " s" << std::endl;

    for (int i = 0; i < 10; i++) solver.step(params);
    solver.synchronize();
    for (int i = 0; i < WARMUP_ITERATIONS; i++) solver.step(params);
    solver.synchronize();

    auto t0 = std::chrono::high_resolution_clock::now();
    for (int i = 0; i < BENCHMARK_ITERATIONS; i++) solver.step(params);
    solver.synchronize();
    auto t1 = std::chrono::high_resolution_clock::now();
    double elapsed = std::chrono::duration<double, std::milli>(t1 - t0).count();

    std::cout << "Avg per step:   " << std::fixed << std::setprecision(3) << elapsed / BENCHMARK_ITERATIONS << " ms" << std::endl;
    std::cout << "Total steps:    " << params.step << std::endl;
    std::cout << "Final sim time: " << std::scientific << std::setprecision(4) << params.t_current << " s" << std::endl;
    std::cout << "Cell-updates/s: " << std::scientific << std::setprecision(3)
              << (double)nx * ny * BENCHMARK_ITERATIONS / (elapsed * 1e-3) << std::endl;
    return 0;
}
