/*
 * fvm.h -- C ABI of the MI355X-native 2D compressible finite-volume time step.
 *
 * This is the drop-in boundary.  Every entry point is `extern "C"`, takes plain
 * pointers and sizes and returns an fvm_status.  The C++ class `CFDSolver` in
 * <fvm/solver.hpp> wraps these calls with the exact method names of the
 * reference class (src/include/solver.cuh:22-80), so the ImGui front end
 * (src/main.cpp) and KernelBenchmark (src/test/kernel_benchmark.cu) can link
 * against this library instead of src/solver.cu.
 *
 * Threading: one handle per host thread; every handle owns one HIP stream.
 * Errors: every call returns FVM_OK or a negative status; fvm_last_error()
 * gives the message of the last failure on the calling thread.
 */
#ifndef FVM_FVM_H
#define FVM_FVM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FVM_ABI_VERSION 1

typedef enum fvm_status {
    FVM_OK = 0,
    FVM_ERR_INVALID_ARGUMENT = -1,
    FVM_ERR_HIP = -2,
    FVM_ERR_OUT_OF_MEMORY = -3,
    FVM_ERR_NOT_INITIALIZED = -4,
    FVM_ERR_CAPACITY = -5,
    FVM_ERR_NOT_FOUND = -6
} fvm_status;

/* Cell tags; values are ABI (src/include/common.h:105-111). */
enum {
    FVM_CELL_FLUID = 0,
    FVM_CELL_SOLID = 1,
    FVM_CELL_GHOST = 2,
    FVM_CELL_INFLOW = 3,
    FVM_CELL_OUTFLOW = 4
};

/* Obstacle shapes; values are ABI (src/include/common.h:39-47). */
enum {
    FVM_SHAPE_CIRCLE = 0,
    FVM_SHAPE_STAR = 1,
    FVM_SHAPE_DIAMOND = 2,
    FVM_SHAPE_CAPSULE = 3,
    FVM_SHAPE_TRIANGLE = 4,
    FVM_SHAPE_STARSHIP = 5
};

/* Riemann flux (addition: the reference always uses HLLC, solver.cu:1571).
 * AUSM+ (Liou 1996) has no low-Mach pressure diffusion: use it for supersonic
 * flow; past walls at low Mach (e.g. Mach 0.3) it diverges with this scheme. */
enum { FVM_FLUX_HLLC = 0, FVM_FLUX_AUSM_PLUS = 1 };
/* MUSCL limiter (addition: the reference has minmod only, solver.cu:110). */
enum { FVM_LIMITER_MINMOD = 0, FVM_LIMITER_VAN_LEER = 1 };
/* Inviscid wall: reference slip/no-slip blend (solver.cu:2146-2173) or pure slip. */
enum { FVM_WALL_REFERENCE = 0, FVM_WALL_SLIP = 1 };
/* Arithmetic of the state and the update.  FVM_PRECISION_FP32_KAHAN (FP32
 * fluxes, compensated SSP blend, FP64 CFL maxima and dt) is for inviscid cases:
 * fvm_initialize, fvm_reset, fvm_step and fvm_advance reject it with enable_viscosity set
 * (FVM_ERR_INVALID_ARGUMENT, state unchanged). */
enum { FVM_PRECISION_FP32 = 0, FVM_PRECISION_FP32_KAHAN = 1, FVM_PRECISION_FP64 = 2 };
/* Where dt comes from: params.dt (reference step()), or the CFL limit of the
 * current state recomputed on the device before every step. */
enum { FVM_DT_HOST = 0, FVM_DT_DEVICE_CFL = 1 };
/* Seeded initial perturbation (addition; the reference starts from uniform flow). */
enum { FVM_NOISE_NONE = 0, FVM_NOISE_DENSITY = 1, FVM_NOISE_TRANSVERSE_VELOCITY = 2 };

/* Fields for fvm_field_to_device / readback (src/solver.cu:3527-3671, 3444-3459). */
enum {
    FVM_FIELD_TEMPERATURE = 0,
    FVM_FIELD_PRESSURE = 1,
    FVM_FIELD_DENSITY = 2,
    FVM_FIELD_VELOCITY_MAG = 3,
    FVM_FIELD_MACH = 4,
    FVM_FIELD_CELL_TYPE = 5
};

/*
 * Simulation parameters.  The first block mirrors `SimParams`
 * (src/include/common.h:50-91) field for field, with bools and the shape enum
 * widened to int32 so the layout is the same for C, C++ and ctypes.  The
 * second block turns the reference's compile-time physical constants
 * (src/solver.cu:10-14, 39-41, 52; src/include/common.h:7-19) into run-time
 * values so that SI and nondimensional runs share one code path.  The third
 * block selects the additions.  Initialise with fvm_params_default().
 */
typedef struct fvm_params {
    /* ---- SimParams ---- */
    int32_t nx, ny;
    float dx, dy;
    float domain_width, domain_height;
    float mach, T_inf, p_inf, rho_inf, u_inf, v_inf, c_inf;
    float obstacle_x, obstacle_y, obstacle_r, obstacle_rotation;
    int32_t obstacle_shape;
    float wing_rotation;
    float cfl, dt, t_current;
    int32_t step;
    int32_t enable_viscosity;
    float cfl_visc;
    int32_t adiabatic_wall;
    float T_wall;
    int32_t paused;
    /* ---- unit system (defaults: the reference's SI constants) ---- */
    float gamma;                   /* 1.4      common.h:7            */
    float r_gas;                   /* 287.05   common.h:8            */
    float mu_ref;                  /* 1.716e-5 common.h:14           */
    float t_ref;                   /* 273.15   common.h:15           */
    float s_sutherland;            /* 110.4    common.h:16           */
    float prandtl;                 /* 0.71     common.h:19           */
    float min_density;             /* 1e-4     solver.cu:10          */
    float min_pressure;            /* 1e-2     solver.cu:11          */
    float max_temperature;         /* 5000     solver.cu:12          */
    float min_temperature;         /* 50       solver.cu:13          */
    float low_density_threshold;   /* 0.01     solver.cu:39          */
    float max_velocity_limit;      /* 5000     solver.cu:40          */
    float max_ghost_velocity;      /* 2500     solver.cu:41          */
    float slip_reference_velocity; /* 100      solver.cu:52          */
    /* ---- additions ---- */
    int32_t flux_scheme;  /* FVM_FLUX_*      */
    int32_t limiter;      /* FVM_LIMITER_*   */
    int32_t wall_mode;    /* FVM_WALL_*      */
    int32_t precision;    /* FVM_PRECISION_* */
    int32_t dt_mode;      /* FVM_DT_*        */
    int32_t noise_field;  /* FVM_NOISE_*     */
    float noise_amplitude;
    uint32_t noise_seed;
    int32_t reserved[8];
} fvm_params;

/* Reference defaults (SimParams initialisers, common.h:53-91) plus SI units. */
void fvm_params_default(fvm_params* p);
/* SimParams::computeDerived (common.h:93-101) with the run-time gas constant. */
void fvm_params_compute_derived(fvm_params* p);

/* ------------------------------------------------------------------------ */
/* Plain single-case solver: the CFDSolver counterpart.                      */
/* ------------------------------------------------------------------------ */
typedef struct fvm_solver fvm_solver;

/* Handle on HIP device `device`.  Replaces the constructor (solver.cu:2874). */
int fvm_create(int device, fvm_solver** out);
/* Replaces the destructor / freeMemory (solver.cu:2877-2880, 2984-3095). */
int fvm_destroy(fvm_solver* s);
/* initialize(params): allocate for nx*ny and reset (solver.cu:3102-3111).
 * FVM_ERR_INVALID_ARGUMENT when one field plane would exceed 4 GiB
 * (nx*ny*4 bytes in FP32, nx*ny*8 in FP64): the kernels address a plane with
 * a 32-bit byte offset; likewise when nx or ny is 2^24 or more (a cell's
 * plane index is formed with a 24-bit multiply). */
int fvm_initialize(fvm_solver* s, const fvm_params* p);
/* resize(nx, ny): reallocate if the size changed (solver.cu:3119-3129). */
int fvm_resize(fvm_solver* s, int nx, int ny);
/* reset(params): init fields, mark obstacle, BC, CFL maxima (solver.cu:3136-3162). */
int fvm_reset(fvm_solver* s, const fvm_params* p);
/* updateObstacleGeometry(params): re-mark, zero newly exposed fluid, BC
 * (solver.cu:3171-3196). */
int fvm_update_obstacle(fvm_solver* s, const fvm_params* p);
/* step(params): one SSP-RK2 step with params->dt (FVM_DT_HOST) or the device
 * CFL dt (FVM_DT_DEVICE_CFL); advances params->t_current and params->step
 * (solver.cu:3289-3412). */
int fvm_step(fvm_solver* s, fvm_params* p);
/* nsteps steps in one call; with FVM_DT_DEVICE_CFL the step sequence is
 * replayed from a captured HIP graph and dt never leaves the device.
 * params->t_current / step are updated at the end. */
int fvm_advance(fvm_solver* s, fvm_params* p, int nsteps);
/* Captures the HIP graphs fvm_advance replays with FVM_DT_DEVICE_CFL (2, 16
 * and 64 steps, both step parities) for params *p without stepping, so the next
 * fvm_advance only launches.  A no-op with FVM_DT_HOST.  No reference
 * counterpart (the reference launches every kernel of every step). */
int fvm_prepare_graphs(fvm_solver* s, const fvm_params* p);
/* computeStableTimeStep (solver.cu:3273-3278). */
int fvm_compute_stable_dt(fvm_solver* s, const fvm_params* p, float* dt);
/* queueTimeStepComputation / readTimeStepResult (solver.cu:3206-3265). */
int fvm_queue_dt(fvm_solver* s, const fvm_params* p);
int fvm_read_dt(fvm_solver* s, const fvm_params* p, float* dt);
/* getCellTypes: nx*ny bytes into caller host memory (solver.cu:3432-3435). */
int fvm_get_cell_types(fvm_solver* s, uint8_t* host);
/* Signed distance field, nx*ny floats into caller host memory. */
int fvm_get_sdf(fvm_solver* s, float* host);
/* convertCellTypesToDevice / compute{Temperature,Pressure,Density,
 * VelocityMag,Mach}ToDevice: nx*ny floats into caller device memory
 * (solver.cu:3444-3459, 3527-3671).  The write is ordered after the caller's
 * earlier work on the null stream and is complete when the call returns, as
 * the reference's default-stream launch is for its callers. */
int fvm_field_to_device(fvm_solver* s, int field, float* dev);
/* The same fields into caller host memory. */
int fvm_field_to_host(fvm_solver* s, int field, float* host);
/* Conserved state (rho, rho_u, rho_v, E, rho_e), SoA, 5*nx*ny values into / from
 * caller host memory.  The _f64 forms convert; no reference counterpart. */
int fvm_get_state(fvm_solver* s, float* host5);
int fvm_set_state(fvm_solver* s, const float* host5);
int fvm_get_state_f64(fvm_solver* s, double* host5);
/* Domain totals of (rho, rho_u, rho_v, E) summed in FP64 on the device. */
int fvm_conservation_sums(fvm_solver* s, double out4[4]);
/* fvm_conservation_sums without a handle and without allocating: no
 * counterpart in the reference.  For callers that cannot call hipMalloc (code
 * under HIP graph capture, serving loops that must not allocate).  dev_state:
 * a state in caller device memory, SoA planes rho, rho_u, rho_v, E (a fifth
 * plane may follow) of nx*ny values each, row-major, elem_bytes 4 (FP32) or 8
 * (FP64).  dev_out4: 4 doubles of caller device memory, written on `stream`
 * (a hipStream_t; NULL = the null stream); the call returns without waiting.
 * Same kernels and workgroup count as fvm_conservation_sums, so the same bits
 * for the same state.  Its partial sums live in one 32 KiB device array per
 * GPU, kept until the process ends and shared by every call on that GPU:
 * calls that overlap on one GPU race.  Rejects nx*ny > FVM_STATIC_MAX_CELLS. */
#define FVM_STATIC_MAX_CELLS 2147483647
int fvm_conservation_sums_static(const void* dev_state, int elem_bytes, int nx, int ny, double* dev_out4, void* stream);
/* Pressure force on the obstacle per unit depth, (Fx, Fy), summed in FP64 over
 * the wall faces of the current state: the wall flux (0, p, 0, 0, 0) that
 * computeFluxesKernel (solver.cu:1589-1598, 1673-1682) gives a face with a
 * solid or ghost cell on either side, counted where the fluid cell on the
 * other side applies it.  out3[2] = number of such faces.  No reference
 * counterpart (the reference reports only max T and Mach): it is the drag /
 * lift diagnostic for long-run comparisons. */
int fvm_forces(fvm_solver* s, double out3[3]);
/* Device-side simulated time (FP64 sum of the dt of every step since reset)
 * and step count. */
int fvm_get_time(fvm_solver* s, double* t, int64_t* steps);

/*
 * State snapshot (no reference counterpart: the reference keeps its state on
 * the device only).  File layout, little-endian:
 *   fvm_snapshot_header (256 bytes)
 *   fvm_params          (header.params_bytes bytes)
 *   5 planes rho, rho_u, rho_v, E, rho_e of nx*ny values, row-major
 *     idx = j*nx + i, header.elem_bytes each (4: FP32 state, 8: FP64 state)
 *   5 planes of FP32 blend compensation carries when header.kahan != 0
 * The header also keeps the device time, the step count and the CFL maxima
 * the next step's dt is computed from, so a loaded snapshot continues exactly
 * like the run that wrote it.
 */
#define FVM_SNAPSHOT_MAGIC "FVMSNAP"
#define FVM_SNAPSHOT_VERSION 1
typedef struct fvm_snapshot_header {
    char magic[8];              /* "FVMSNAP" + NUL */
    uint32_t version;           /* FVM_SNAPSHOT_VERSION */
    uint32_t header_bytes;      /* sizeof(fvm_snapshot_header) = 256 */
    uint32_t abi_version;       /* FVM_ABI_VERSION of the writer */
    uint32_t params_bytes;      /* sizeof(fvm_params) */
    int32_t nx, ny;
    uint32_t elem_bytes;        /* 4 or 8 */
    uint32_t kahan;             /* 1: carry planes follow the state */
    double t;                   /* device simulated time */
    int64_t steps;              /* device step count */
    uint32_t max_speed_bits;    /* CFL maxima of the state (float bit patterns) */
    uint32_t max_nu_bits;
    double max_speed64;         /* the same in FP64 (FP64 and compensated variants) */
    double max_nu64;
    uint8_t reserved[176];
} fvm_snapshot_header;

/* Writes the current state of the solver to `path`. */
int fvm_save_snapshot(fvm_solver* s, const char* path);
/* Reads `path`: initialises the solver with the stored parameters (as
 * fvm_initialize), then restores state, time, step count, CFL maxima and
 * carries.  params_out (may be NULL) receives the stored parameters. */
int fvm_load_snapshot(fvm_solver* s, const char* path, fvm_params* params_out);
/* generateVectorArrows (solver.cu:3809-3845): velocity arrows for the display
 * overlay into caller device memory (an OpenGL VBO mapped by the caller), as
 * GL_LINES vertices (x, y, r, g, b) in normalised device coordinates, 8 per
 * arrow, one arrow per step-th cell (from step/2) that moves at 1e-6 u_inf or
 * more; *num_vertices receives the vertices written.  Unlike the reference,
 * nothing is written beyond max_vertices (the count is capped), and the two
 * vertices per arrow the reference leaves unwritten form a zero-length line.
 * Arrow order is unspecified.  Ordered after the caller's null-stream work and
 * complete on return, as fvm_field_to_device. */
int fvm_generate_vector_arrows(fvm_solver* s, float* dev_vertex_data, int max_vertices, int step, float u_inf,
                               float max_arrow_length, float arrow_head_angle, float arrow_head_length,
                               int* num_vertices);
/* getMaxTemperature / getMaxMach (solver.cu:3417-3426). */
int fvm_max_temperature(fvm_solver* s, float* out);
int fvm_max_mach(fvm_solver* s, float* out);
/* getSimulationMemoryUsage (solver.cu:3474-3514): bytes this handle holds. */
int fvm_memory_usage(fvm_solver* s, size_t* bytes);
/* getGPUMemoryInfo (solver.cu:3468-3471); argument order is (free, total),
 * as in the definition. */
int fvm_gpu_memory_info(int device, size_t* free_bytes, size_t* total_bytes);
/* Wait for all work of this handle (replaces cudaDeviceSynchronize calls of
 * the callers: main.cpp:365, kernel_benchmark.cu:103,113,132). */
int fvm_synchronize(fvm_solver* s);
/* Grid size (getNx / getNy, solver.cuh:62-63). */
int fvm_get_size(fvm_solver* s, int* nx, int* ny);
/* Tile shape of the stage and viscous kernels, one of the compiled shapes
 * (64x4, 64x6, 64x7, 64x8, 64x12, 64x16, 32x2, 32x3, 32x4, 32x7, 32x8, 32x16, 128x4, 128x8, 96x8;
 * FP64: 64x8, 32x8); 0,0 = the autotuned default for the grid. */
int fvm_set_tile(fvm_solver* s, int tile_x, int tile_y);
/* Schedule of a viscous step (no reference counterpart: the reference always
 * launches its diffusion kernel, solver.cu:3299-3304 and 3343-3354, as its own
 * pass).  FVM_SCHEDULE_SPLIT: the first half step is a launch of its own.
 * FVM_SCHEDULE_FUSED: it is recomputed inside stage 1 on a wider halo (FP32
 * only; tiles whose LDS does not fit, FP64 and the compensated variant keep
 * the split schedule).  FVM_SCHEDULE_AUTO (default): the autotuned choice for
 * the grid, with the tile of fvm_set_tile(0, 0).  All give bitwise the same
 * state. */
#define FVM_SCHEDULE_AUTO 0
#define FVM_SCHEDULE_SPLIT 1
#define FVM_SCHEDULE_FUSED 2
int fvm_set_viscous_schedule(fvm_solver* s, int schedule);
/* Replay captured HIP graphs in fvm_advance with device dt (default on; off:
 * every kernel is launched directly, e.g. for kernel-tracing tools).  No
 * reference counterpart. */
int fvm_use_graph(fvm_solver* s, int enable);
/* Whether FVM_SCHEDULE_FUSED is built for a tile shape and precision. */
int fvm_fused_schedule_supported(int tile_x, int tile_y, int precision, int* supported);
/* Name and average duration of each kernel launched by the last fvm_advance
 * call when timing is enabled (HIP events on the handle's stream). */
int fvm_enable_timing(fvm_solver* s, int enable);
int fvm_kernel_timing(fvm_solver* s, int index, char* name, int name_len, double* avg_ms, int64_t* launches);

/* Message of the last failure on this thread ("" if none). */
const char* fvm_last_error(void);
int fvm_abi_version(void);

/* ------------------------------------------------------------------------ */
/* Batched engine: hundreds of independent cases advanced together.         */
/* No reference counterpart (the reference runs one case per process).      */
/* ------------------------------------------------------------------------ */
typedef struct fvm_batch fvm_batch;

/* mem_budget_bytes: device bytes the engine may hold for case state. */
int fvm_batch_create(int device, size_t mem_budget_bytes, int max_cases, fvm_batch** out);
int fvm_batch_destroy(fvm_batch* b);
/* Admit a case (allocates from the budget, marks, initialises, applies BC).
 * steps_total: number of steps the case runs before it is complete. */
int fvm_batch_add_case(fvm_batch* b, const fvm_params* p, int32_t steps_total, int32_t* case_id);
/* Bytes a case of this size would take (for admission decisions). */
int fvm_batch_case_bytes(const fvm_params* p, size_t* bytes);
int fvm_batch_remove_case(fvm_batch* b, int32_t case_id);
/* Advance every active, unfinished case by up to nsteps steps; each case uses
 * its own CFL dt (FVM_DT_DEVICE_CFL) or its params.dt (FVM_DT_HOST). */
int fvm_batch_advance(fvm_batch* b, int32_t nsteps);
/* Captures the step graphs of the current case set without stepping (see
 * fvm_prepare_graphs); admitting or retiring cases keeps them valid. */
int fvm_batch_prepare_graphs(fvm_batch* b);
/* Case progress: steps done, simulated time, finished flag. */
int fvm_batch_case_status(fvm_batch* b, int32_t case_id, int32_t* steps_done, double* t, int32_t* finished);
int fvm_batch_get_state(fvm_batch* b, int32_t case_id, float* host5);
int fvm_batch_get_cell_types(fvm_batch* b, int32_t case_id, uint8_t* host);
/* fvm_forces for one case of the batch. */
int fvm_batch_forces(fvm_batch* b, int32_t case_id, double out3[3]);
/* Snapshot of one case of the batch (same file format as fvm_save_snapshot). */
int fvm_batch_save_case(fvm_batch* b, int32_t case_id, const char* path);
/* Admits a case from a snapshot (fvm_batch_add_case with the stored
 * parameters, then the stored state); it runs steps_total more steps. */
int fvm_batch_load_case(fvm_batch* b, const char* path, int32_t steps_total, int32_t* case_id);
int fvm_batch_active_cases(fvm_batch* b, int32_t* n);
int fvm_batch_bytes_in_use(fvm_batch* b, size_t* bytes);
int fvm_batch_synchronize(fvm_batch* b);
/* Tile shape of the batched kernels (0,0 = the autotuned default). */
int fvm_batch_set_tile(fvm_batch* b, int tile_x, int tile_y);
/* fvm_set_viscous_schedule for the batched engine (AUTO: the autotuned batch
 * default). */
int fvm_batch_set_viscous_schedule(fvm_batch* b, int schedule);
/* Use a captured HIP graph for fvm_batch_advance (default on). */
int fvm_batch_use_graph(fvm_batch* b, int enable);

#ifdef __cplusplus
}
#endif

#endif /* FVM_FVM_H */
