// front_end_sequence.cpp -- the reference front end's CFDSolver call order
// (src/main.cpp:203-368), headless: initialize, getCellTypes, reset,
// convertCellTypesToDevice, updateObstacleGeometry, computeStableTimeStep,
// step, the *ToDevice readbacks into a device buffer the program owns (read
// back on the null stream, as the front end's graphics interop does), the
// queued dt, the maxima and the memory figure.  It prints one "key value"
// line per figure; the drop-in test compares them with the Python binding.
// SI Mach-3 viscous circle on 256x128 (the reference defaults).
// usage: front_end_sequence [steps=20]
#include <hip/hip_runtime_api.h>
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "fvm/solver.hpp"

static void host_copy(std::vector<float>& h, const float* d) {
    if (hipMemcpy(h.data(), d, h.size() * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) std::exit(4);
}

int main(int argc, char* argv[]) {
    const int nsteps = (argc > 1) ? std::atoi(argv[1]) : 20;
    if (nsteps <= 0) {
        std::fprintf(stderr, "usage: %s [steps=20]\n", argv[0]);
        return 1;
    }
    SimParams params;
    params.nx = 256;
    params.ny = 128;
    params.enable_viscosity = true;
    params.computeDerived();
    const size_t n = (size_t)params.nx * params.ny;

    CFDSolver solver;
    solver.initialize(params);
    std::vector<uint8_t> ct(n);
    solver.getCellTypes(ct.data());
    solver.reset(params);

    float* dev = nullptr;
    if (hipMalloc(&dev, n * sizeof(float)) != hipSuccess) return 4;
    std::vector<float> h(n);
    solver.convertCellTypesToDevice(dev);
    host_copy(h, dev);
    long mismatch = 0, solid = 0;
    for (size_t i = 0; i < n; i++) { mismatch += h[i] != (float)ct[i]; solid += ct[i] == 1; }
    std::printf("ct_mismatch %ld\nsolid %ld\n", mismatch, solid);

    params.obstacle_x += 0.3f;
    solver.updateObstacleGeometry(params);
    std::vector<uint8_t> ct2(n);
    solver.getCellTypes(ct2.data());
    long moved = 0;
    for (size_t i = 0; i < n; i++) moved += ct2[i] != ct[i];
    std::printf("moved %ld\n", moved);

    params.dt = solver.computeStableTimeStep(params);
    std::printf("dt %.9e\n", params.dt);
    for (int s = 0; s < nsteps; s++) solver.step(params);
    solver.synchronize();
    std::printf("step %d\nt %.9e\n", params.step, params.t_current);
    solver.queueTimeStepComputation(params);
    std::printf("dt_next %.9e\n", solver.readTimeStepResult(params));

    const char* names[5] = {"T", "p", "rho", "vmag", "mach"};
    for (int f = 0; f < 5; f++) {
        // a clear of the buffer still queued on the null stream must land
        // before the readback, not after it
        if (hipMemsetAsync(dev, 0, n * sizeof(float), nullptr) != hipSuccess) return 4;
        switch (f) {
            case 0: solver.computeTemperatureToDevice(dev); break;
            case 1: solver.computePressureToDevice(dev); break;
            case 2: solver.computeDensityToDevice(dev); break;
            case 3: solver.computeVelocityMagToDevice(dev); break;
            default: solver.computeMachToDevice(dev); break;
        }
        host_copy(h, dev);
        double sum = 0; float mx = h[0];
        for (size_t i = 0; i < n; i++) { sum += h[i]; mx = h[i] > mx ? h[i] : mx; }
        std::printf("sum_%s %.12e\nmax_%s %.9e\n", names[f], sum, names[f], mx);
    }
    // velocity arrows into a device buffer sized as the front end sizes its
    // VBO (main.cpp:328-346, 910-934): vector density 20, arrow length 0.8 cell
    // spacings per step, head angle 0.5 rad, head 0.3 of the shaft
    {
        const int step = 20;
        const int max_vertices = ((params.nx + step - 1) / step) * ((params.ny + step - 1) / step) * 8;
        float* vbo = nullptr;
        if (hipMalloc(&vbo, (size_t)max_vertices * 5 * sizeof(float)) != hipSuccess) return 4;
        const float cell = std::min(2.0f / params.nx, 2.0f / params.ny);
        const int nv = solver.generateVectorArrows(vbo, max_vertices, step, params.u_inf, cell * (step * 0.8f), 0.5f, 0.3f);
        std::vector<float> hv((size_t)max_vertices * 5);
        host_copy(hv, vbo);
        double ssum = 0;
        for (int q = 0; q < nv * 5; q++) ssum += hv[q];
        std::printf("arrow_vertices %d\narrow_sum %.12e\n", nv, ssum);
        (void)hipFree(vbo);
    }
    std::printf("maxT %.9e\nmaxMach %.9e\n", solver.getMaxTemperature(), solver.getMaxMach());
    std::printf("memory %zu\n", solver.getSimulationMemoryUsage());
    (void)hipFree(dev);
    return 0;
}
