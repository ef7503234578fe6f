// fvm_api.hip -- the extern "C" boundary (include/fvm/fvm.h).
//
// fvm_solver mirrors CFDSolver (src/include/solver.cuh:22-80, src/solver.cu:
// 2874-3514): one case, owned buffers, reset / step / dt / readback.
// fvm_batch advances many independent cases in one launch per kernel.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>

#include "../../include/fvm/fvm.h"
#include "fvm_engine.hpp"

#pragma clang fp contract(off)

namespace fvm {
const char* last_error();
}

using fvm::Engine;
using fvm::set_error;

struct fvm_solver {
    int device = 0;
    std::unique_ptr<Engine> eng;
    int case_id = -1;
    fvm_params p{};
    bool have_params = false;
    int tile_x = 0, tile_y = 0;
    int schedule = FVM_SCHEDULE_AUTO;
    bool timing = false;
    bool use_graph = true;
};

struct fvm_batch {
    int device = 0;
    std::unique_ptr<Engine> eng;
    size_t budget = 0;
    int max_cases = 0;
    bool use_graph = true;
    int schedule = FVM_SCHEDULE_AUTO;
    bool default_tile = true;
};

namespace {

// Fields that change the constants of a case (everything but time keeping).
bool consts_differ(const fvm_params& a, const fvm_params& b) {
    fvm_params x = a, y = b;
    x.dt = y.dt = 0;
    x.t_current = y.t_current = 0;
    x.step = y.step = 0;
    x.paused = y.paused = 0;
    return std::memcmp(&x, &y, sizeof(x)) != 0;
}

#include "tile_table.inc"

struct TileEntry {
    int viscous, bucket, tx, ty, fused;
};
const TileEntry kTileTable[] = {FVM_TILE_ENTRIES};

// Tile shape and viscous schedule for a case from the autotuned table: same
// viscosity, nearest log4 cell-count bucket (fvm_amd/autotune.py measures and
// writes the table).
int pick_tile(const fvm_params& p, int* tx, int* ty, bool* fused = nullptr) {
    double cells = (double)p.nx * (double)p.ny;
    int bucket = (int)std::lround(std::log(std::max(cells, 1.0) / 32768.0) / std::log(4.0));
    bucket = std::max(0, std::min(4, bucket));
    int best = -1, best_d = 1 << 30;
    for (int i = 0; i < (int)(sizeof(kTileTable) / sizeof(kTileTable[0])); i++) {
        if (kTileTable[i].viscous != (p.enable_viscosity ? 1 : 0)) continue;
        int d = std::abs(kTileTable[i].bucket - bucket);
        if (d < best_d) { best_d = d; best = i; }
    }
    *tx = best >= 0 ? kTileTable[best].tx : 64;
    *ty = best >= 0 ? kTileTable[best].ty : 8;
    if (fused) *fused = best >= 0 && kTileTable[best].fused != 0;
    if (!fvm::tile_supported(*tx, *ty, p.precision == FVM_PRECISION_FP64)) { *tx = 64; *ty = 8; }
    return 0;
}

// The schedule an engine runs: the explicit choice, or under AUTO the table's
// (for an explicit tile: split).
bool schedule_fused(const fvm_solver* s, const fvm_params& p) {
    if (s->schedule != FVM_SCHEDULE_AUTO) return s->schedule == FVM_SCHEDULE_FUSED;
    if (s->tile_x != 0) return false;
    int tx, ty;
    bool fused = false;
    pick_tile(p, &tx, &ty, &fused);
    return fused;
}

// The batched engine's schedule: explicit, or under AUTO the autotuned batch
// default (for an explicit tile: split).
bool batch_fused(const fvm_batch* b) {
    if (b->schedule != FVM_SCHEDULE_AUTO) return b->schedule == FVM_SCHEDULE_FUSED;
    return b->default_tile && FVM_BATCH_FUSED;
}

int ensure_engine(fvm_solver* s, const fvm_params& p) {
    bool fp64 = p.precision == FVM_PRECISION_FP64;
    bool kahan = p.precision == FVM_PRECISION_FP32_KAHAN;
    bool need_new = !s->eng || s->eng->fp64() != fp64 || s->eng->kahan() != kahan ||
                    !s->have_params || s->p.nx != p.nx || s->p.ny != p.ny;
    if (!need_new) return FVM_OK;
    s->eng.reset(new Engine(s->device, fp64, kahan));
    int rc = s->eng->init_stream();
    if (rc) return rc;
    int tx = s->tile_x, ty = s->tile_y;
    if (tx == 0) pick_tile(p, &tx, &ty);
    rc = s->eng->set_tile(tx, ty);
    if (rc) return rc;
    s->eng->set_fused_d1(schedule_fused(s, p));
    s->eng->enable_timing(s->timing);
    int id = -1;
    rc = s->eng->add_case(p, 0, true, &id);
    if (rc) {
        s->eng.reset();
        s->have_params = false;
        return rc;
    }
    s->case_id = id;
    s->p = p;
    s->have_params = true;
    return FVM_OK;
}

int apply_params(fvm_solver* s, const fvm_params& p) {
    if (!s->eng || !s->have_params) return set_error(FVM_ERR_NOT_INITIALIZED, "solver not initialized");
    if (p.nx != s->p.nx || p.ny != s->p.ny) return set_error(FVM_ERR_INVALID_ARGUMENT, "grid size differs; call fvm_resize or fvm_initialize");
    if ((p.precision == FVM_PRECISION_FP64) != s->eng->fp64() || (p.precision == FVM_PRECISION_FP32_KAHAN) != s->eng->kahan())
        return set_error(FVM_ERR_INVALID_ARGUMENT, "precision differs; call fvm_initialize");
    if (consts_differ(p, s->p)) {
        int rc = s->eng->update_params(s->case_id, p);
        if (rc) return rc;
    }
    if (p.dt != s->p.dt) {
        int rc = s->eng->set_host_dt(s->case_id, p.dt);
        if (rc) return rc;
    }
    s->p = p;
    return FVM_OK;
}

// readTimeStepResult, src/solver.cu:3250-3265: FP32 as in the reference, or
// FP64 from the FP64 maxima (FP64 state and compensated variant), rounded to
// the float the interface returns.
float dt_from_maxima(const fvm_params& p, double max_speed, double max_nu, bool dt64) {
    if (dt64) {
        const double min_dx = std::fmin((double)p.dx, (double)p.dy);
        double dt = (double)p.cfl * min_dx / (max_speed + (double)1e-10f);
        if (p.enable_viscosity) dt = std::fmin(dt, (double)p.cfl_visc * min_dx * min_dx / (max_nu + (double)1e-10f));
        return (float)dt;
    }
    float min_dx = std::fmin(p.dx, p.dy);
    float dt = p.cfl * min_dx / ((float)max_speed + 1e-10f);
    if (p.enable_viscosity) {
        float dtv = p.cfl_visc * min_dx * min_dx / ((float)max_nu + 1e-10f);
        dt = std::fmin(dt, dtv);
    }
    return dt;
}

}  // namespace

#define CHECK_SOLVER(s)                                                              \
    do {                                                                             \
        if (!(s)) return set_error(FVM_ERR_INVALID_ARGUMENT, "null solver handle");   \
        if (!(s)->eng || !(s)->have_params)                                          \
            return set_error(FVM_ERR_NOT_INITIALIZED, "solver not initialized");     \
    } while (0)

extern "C" {

int fvm_abi_version(void) { return FVM_ABI_VERSION; }
const char* fvm_last_error(void) { return fvm::last_error(); }

void fvm_params_default(fvm_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->nx = 512; p->ny = 256; p->dx = 1.0f; p->dy = 1.0f;
    p->domain_width = 10.0f; p->domain_height = 5.0f;
    p->mach = 3.0f; p->T_inf = 300.0f; p->p_inf = 101325.0f;
    p->obstacle_x = 2.5f; p->obstacle_y = 2.5f; p->obstacle_r = 0.5f;
    p->obstacle_shape = FVM_SHAPE_CIRCLE;
    p->cfl = 0.5f; p->dt = 1e-6f; p->cfl_visc = 0.4f;
    p->adiabatic_wall = 1; p->T_wall = 300.0f; p->paused = 1;
    p->gamma = 1.4f; p->r_gas = 287.05f; p->mu_ref = 1.716e-5f; p->t_ref = 273.15f;
    p->s_sutherland = 110.4f; p->prandtl = 0.71f;
    p->min_density = 1e-4f; p->min_pressure = 1e-2f;
    p->max_temperature = 5000.0f; p->min_temperature = 50.0f;
    p->low_density_threshold = 0.01f; p->max_velocity_limit = 5000.0f;
    p->max_ghost_velocity = 2500.0f; p->slip_reference_velocity = 100.0f;
}

void fvm_params_compute_derived(fvm_params* p) {
    if (!p) return;
    p->rho_inf = p->p_inf / (p->r_gas * p->T_inf);
    p->c_inf = std::sqrt(p->gamma * p->r_gas * p->T_inf);
    p->u_inf = p->mach * p->c_inf;
    p->dx = p->domain_width / (float)p->nx;
    p->dy = p->domain_height / (float)p->ny;
    p->obstacle_y = p->domain_height / 2.0f;
}

int fvm_create(int device, fvm_solver** out) {
    if (!out) return set_error(FVM_ERR_INVALID_ARGUMENT, "null output");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) return set_error(FVM_ERR_HIP, "no HIP device available");
    if (device < 0 || device >= n) return set_error(FVM_ERR_INVALID_ARGUMENT, "device index out of range");
    fvm_solver* s = new fvm_solver();
    s->device = device;
    *out = s;
    return FVM_OK;
}

int fvm_destroy(fvm_solver* s) {
    if (!s) return FVM_OK;
    s->eng.reset();
    delete s;
    return FVM_OK;
}

int fvm_initialize(fvm_solver* s, const fvm_params* p) {
    if (!s || !p) return set_error(FVM_ERR_INVALID_ARGUMENT, "null argument");
    s->have_params = false;
    s->eng.reset();
    return ensure_engine(s, *p);
}

int fvm_resize(fvm_solver* s, int nx, int ny) {
    if (!s) return set_error(FVM_ERR_INVALID_ARGUMENT, "null solver handle");
    if (nx <= 0 || ny <= 0) return FVM_OK;  // allocateMemory returns silently (solver.cu:2888-2889)
    if (s->have_params && nx == s->p.nx && ny == s->p.ny) return FVM_OK;
    if (!s->have_params) return set_error(FVM_ERR_NOT_INITIALIZED, "resize before initialize");
    fvm_params p = s->p;
    p.nx = nx;
    p.ny = ny;
    s->have_params = false;
    return ensure_engine(s, p);
}

int fvm_reset(fvm_solver* s, const fvm_params* p) {
    if (!s || !p) return set_error(FVM_ERR_INVALID_ARGUMENT, "null argument");
    if (!s->eng || !s->have_params || p->nx != s->p.nx || p->ny != s->p.ny ||
        (p->precision == FVM_PRECISION_FP64) != s->eng->fp64() ||
        (p->precision == FVM_PRECISION_FP32_KAHAN) != s->eng->kahan())
        return fvm_initialize(s, p);
    int rc = s->eng->reset_case(s->case_id, *p);
    if (rc) return rc;
    s->p = *p;
    return FVM_OK;
}

int fvm_update_obstacle(fvm_solver* s, const fvm_params* p) {
    CHECK_SOLVER(s);
    if (!p) return set_error(FVM_ERR_INVALID_ARGUMENT, "null params");
    int rc = s->eng->update_obstacle(s->case_id, *p);
    if (rc) return rc;
    s->p = *p;
    return FVM_OK;
}

int fvm_step(fvm_solver* s, fvm_params* p) {
    CHECK_SOLVER(s);
    if (!p) return set_error(FVM_ERR_INVALID_ARGUMENT, "null params");
    int rc = apply_params(s, *p);
    if (rc) return rc;
    rc = s->eng->advance(1, false);
    if (rc) return rc;
    if (p->dt_mode == FVM_DT_DEVICE_CFL) {
        double t = 0;
        long long n = 0;
        rc = s->eng->read_time(s->case_id, &t, &n);
        if (rc) return rc;
        p->t_current = (float)t;
    } else {
        p->t_current += p->dt;
    }
    p->step++;
    s->p.t_current = p->t_current;
    s->p.step = p->step;
    return FVM_OK;
}

int fvm_advance(fvm_solver* s, fvm_params* p, int nsteps) {
    CHECK_SOLVER(s);
    if (!p) return set_error(FVM_ERR_INVALID_ARGUMENT, "null params");
    if (nsteps < 0) return set_error(FVM_ERR_INVALID_ARGUMENT, "negative step count");
    int rc = apply_params(s, *p);
    if (rc) return rc;
    rc = s->eng->advance(nsteps, s->use_graph && p->dt_mode == FVM_DT_DEVICE_CFL);
    if (rc) return rc;
    if (p->dt_mode == FVM_DT_DEVICE_CFL) {
        double t = 0;
        long long n = 0;
        rc = s->eng->read_time(s->case_id, &t, &n);
        if (rc) return rc;
        p->t_current = (float)t;
    } else {
        for (int i = 0; i < nsteps; i++) p->t_current += p->dt;
    }
    p->step += nsteps;
    s->p.t_current = p->t_current;
    s->p.step = p->step;
    return FVM_OK;
}

int fvm_prepare_graphs(fvm_solver* s, const fvm_params* p) {
    CHECK_SOLVER(s);
    if (!p) return set_error(FVM_ERR_INVALID_ARGUMENT, "null params");
    int rc = apply_params(s, *p);
    if (rc) return rc;
    if (p->dt_mode != FVM_DT_DEVICE_CFL) return FVM_OK;   // host dt: fvm_advance launches directly
    return s->eng->prepare_graphs();
}

int fvm_compute_stable_dt(fvm_solver* s, const fvm_params* p, float* dt) {
    CHECK_SOLVER(s);
    if (!p || !dt) return set_error(FVM_ERR_INVALID_ARGUMENT, "null argument");
    double ms = 0, mn = 0;
    int rc = s->eng->read_maxima(s->case_id, &ms, &mn);
    if (rc) return rc;
    *dt = dt_from_maxima(*p, ms, mn, s->eng->dt64());
    return FVM_OK;
}

int fvm_queue_dt(fvm_solver* s, const fvm_params* p) {
    CHECK_SOLVER(s);
    (void)p;
    return s->eng->queue_maxima(s->case_id);
}

int fvm_read_dt(fvm_solver* s, const fvm_params* p, float* dt) {
    return fvm_compute_stable_dt(s, p, dt);
}

int fvm_get_cell_types(fvm_solver* s, uint8_t* host) {
    CHECK_SOLVER(s);
    if (!host) return set_error(FVM_ERR_INVALID_ARGUMENT, "null buffer");
    return s->eng->read_cell_types(s->case_id, host);
}

int fvm_get_sdf(fvm_solver* s, float* host) {
    CHECK_SOLVER(s);
    if (!host) return set_error(FVM_ERR_INVALID_ARGUMENT, "null buffer");
    return s->eng->read_sdf(s->case_id, host);
}

int fvm_field_to_device(fvm_solver* s, int field, float* dev) {
    CHECK_SOLVER(s);
    if (!dev) return set_error(FVM_ERR_INVALID_ARGUMENT, "null buffer");
    return s->eng->field(s->case_id, field, dev);
}

int fvm_field_to_host(fvm_solver* s, int field, float* host) {
    CHECK_SOLVER(s);
    if (!host) return set_error(FVM_ERR_INVALID_ARGUMENT, "null buffer");
    size_t n = (size_t)s->p.nx * s->p.ny;
    float* dev = nullptr;
    hipError_t e = hipMalloc((void**)&dev, n * sizeof(float));
    if (e != hipSuccess) return fvm::hip_fail(e, "hipMalloc(field)");
    int rc = s->eng->field(s->case_id, field, dev);
    if (!rc) rc = s->eng->sync();
    if (!rc) {
        e = hipMemcpy(host, dev, n * sizeof(float), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = fvm::hip_fail(e, "hipMemcpy(field)");
    }
    (void)hipFree(dev);
    return rc;
}

int fvm_get_state(fvm_solver* s, float* host5) {
    CHECK_SOLVER(s);
    if (!host5) return set_error(FVM_ERR_INVALID_ARGUMENT, "null buffer");
    return s->eng->read_state(s->case_id, host5, false);
}

int fvm_get_state_f64(fvm_solver* s, double* host5) {
    CHECK_SOLVER(s);
    if (!host5) return set_error(FVM_ERR_INVALID_ARGUMENT, "null buffer");
    return s->eng->read_state(s->case_id, host5, true);
}

int fvm_set_state(fvm_solver* s, const float* host5) {
    CHECK_SOLVER(s);
    if (!host5) return set_error(FVM_ERR_INVALID_ARGUMENT, "null buffer");
    return s->eng->write_state(s->case_id, host5);
}

int fvm_conservation_sums(fvm_solver* s, double out4[4]) {
    CHECK_SOLVER(s);
    if (!out4) return set_error(FVM_ERR_INVALID_ARGUMENT, "null buffer");
    return s->eng->conservation(s->case_id, out4);
}

int fvm_forces(fvm_solver* s, double out3[3]) {
    CHECK_SOLVER(s);
    if (!out3) return set_error(FVM_ERR_INVALID_ARGUMENT, "null buffer");
    return s->eng->forces(s->case_id, out3);
}

int fvm_save_snapshot(fvm_solver* s, const char* path) {
    CHECK_SOLVER(s);
    return s->eng->save_case(s->case_id, path);
}

int fvm_load_snapshot(fvm_solver* s, const char* path, fvm_params* params_out) {
    if (!s) return set_error(FVM_ERR_INVALID_ARGUMENT, "null solver handle");
    fvm_snapshot_header h;
    fvm_params p;
    int rc = fvm::read_snapshot_header(path, &h, &p);
    if (rc) return rc;
    rc = fvm_initialize(s, &p);
    if (rc) return rc;
    rc = s->eng->load_case(s->case_id, path);
    if (rc) return rc;
    if (params_out) *params_out = p;
    return FVM_OK;
}

int fvm_get_time(fvm_solver* s, double* t, int64_t* steps) {
    CHECK_SOLVER(s);
    long long n = 0;
    int rc = s->eng->read_time(s->case_id, t, &n);
    if (steps) *steps = n;
    return rc;
}

int fvm_max_temperature(fvm_solver* s, float* out) {
    CHECK_SOLVER(s);
    if (!out) return set_error(FVM_ERR_INVALID_ARGUMENT, "null output");
    return s->eng->stat(s->case_id, 0, out);
}

int fvm_max_mach(fvm_solver* s, float* out) {
    CHECK_SOLVER(s);
    if (!out) return set_error(FVM_ERR_INVALID_ARGUMENT, "null output");
    return s->eng->stat(s->case_id, 1, out);
}

int fvm_generate_vector_arrows(fvm_solver* s, float* dev_vertex_data, int max_vertices, int step, float u_inf,
                               float max_arrow_length, float arrow_head_angle, float arrow_head_length,
                               int* num_vertices) {
    CHECK_SOLVER(s);
    if (!num_vertices) return set_error(FVM_ERR_INVALID_ARGUMENT, "null output");
    return s->eng->arrows(s->case_id, dev_vertex_data, max_vertices, step, u_inf, max_arrow_length, arrow_head_angle,
                          arrow_head_length, num_vertices);
}

int fvm_memory_usage(fvm_solver* s, size_t* bytes) {
    if (!s || !bytes) return set_error(FVM_ERR_INVALID_ARGUMENT, "null argument");
    *bytes = s->eng ? s->eng->bytes_in_use() : 0;
    return FVM_OK;
}

int fvm_gpu_memory_info(int device, size_t* free_bytes, size_t* total_bytes) {
    if (!free_bytes || !total_bytes) return set_error(FVM_ERR_INVALID_ARGUMENT, "null argument");
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return fvm::hip_fail(e, "hipSetDevice");
    e = hipMemGetInfo(free_bytes, total_bytes);
    if (e != hipSuccess) return fvm::hip_fail(e, "hipMemGetInfo");
    return FVM_OK;
}

int fvm_synchronize(fvm_solver* s) {
    CHECK_SOLVER(s);
    return s->eng->sync();
}

int fvm_get_size(fvm_solver* s, int* nx, int* ny) {
    if (!s || !nx || !ny) return set_error(FVM_ERR_INVALID_ARGUMENT, "null argument");
    *nx = s->have_params ? s->p.nx : 0;
    *ny = s->have_params ? s->p.ny : 0;
    return FVM_OK;
}

int fvm_set_tile(fvm_solver* s, int tile_x, int tile_y) {
    if (!s) return set_error(FVM_ERR_INVALID_ARGUMENT, "null solver handle");
    if (tile_x == 0 && tile_y == 0) {
        s->tile_x = s->tile_y = 0;
        if (s->eng) {
            int tx, ty;
            pick_tile(s->p, &tx, &ty);
            s->eng->set_fused_d1(schedule_fused(s, s->p));
            return s->eng->set_tile(tx, ty);
        }
        return FVM_OK;
    }
    if (!fvm::tile_supported(tile_x, tile_y, false)) return set_error(FVM_ERR_INVALID_ARGUMENT, "unsupported tile shape");
    s->tile_x = tile_x;
    s->tile_y = tile_y;
    if (s->eng) {
        s->eng->set_fused_d1(schedule_fused(s, s->p));
        return s->eng->set_tile(tile_x, tile_y);
    }
    return FVM_OK;
}

int fvm_set_viscous_schedule(fvm_solver* s, int schedule) {
    if (!s) return set_error(FVM_ERR_INVALID_ARGUMENT, "null solver handle");
    if (schedule != FVM_SCHEDULE_AUTO && schedule != FVM_SCHEDULE_SPLIT && schedule != FVM_SCHEDULE_FUSED)
        return set_error(FVM_ERR_INVALID_ARGUMENT, "unknown viscous schedule");
    s->schedule = schedule;
    if (s->eng) s->eng->set_fused_d1(schedule_fused(s, s->p));
    return FVM_OK;
}

int fvm_use_graph(fvm_solver* s, int enable) {
    if (!s) return set_error(FVM_ERR_INVALID_ARGUMENT, "null solver handle");
    s->use_graph = enable != 0;
    return FVM_OK;
}

int fvm_fused_schedule_supported(int tile_x, int tile_y, int precision, int* supported) {
    if (!supported) return set_error(FVM_ERR_INVALID_ARGUMENT, "null argument");
    *supported = precision == FVM_PRECISION_FP32 && fvm::fused_d1_supported(tile_x, tile_y, false) ? 1 : 0;
    return FVM_OK;
}

int fvm_enable_timing(fvm_solver* s, int enable) {
    if (!s) return set_error(FVM_ERR_INVALID_ARGUMENT, "null solver handle");
    s->timing = enable != 0;
    if (s->eng) {
        s->eng->enable_timing(s->timing);
        s->eng->clear_timings();
    }
    return FVM_OK;
}

int fvm_kernel_timing(fvm_solver* s, int index, char* name, int name_len, double* avg_ms, int64_t* launches) {
    CHECK_SOLVER(s);
    int rc = s->eng->flush_timings();
    if (rc) return rc;
    const auto& t = s->eng->timings();
    if (index < 0 || index >= (int)t.size()) return set_error(FVM_ERR_NOT_FOUND, "no such timing entry");
    if (name && name_len > 0) {
        std::strncpy(name, t[index].first.c_str(), name_len - 1);
        name[name_len - 1] = 0;
    }
    if (avg_ms) *avg_ms = t[index].second.launches ? t[index].second.total_ms / t[index].second.launches : 0.0;
    if (launches) *launches = t[index].second.launches;
    return FVM_OK;
}

// ----------------------------------------------------------------- batch
int fvm_batch_create(int device, size_t mem_budget_bytes, int max_cases, fvm_batch** out) {
    if (!out) return set_error(FVM_ERR_INVALID_ARGUMENT, "null output");
    if (max_cases <= 0 || mem_budget_bytes < 4096) return set_error(FVM_ERR_INVALID_ARGUMENT, "bad budget or case limit");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) return set_error(FVM_ERR_HIP, "no HIP device available");
    if (device < 0 || device >= n) return set_error(FVM_ERR_INVALID_ARGUMENT, "device index out of range");
    std::unique_ptr<fvm_batch> b(new fvm_batch());
    b->device = device;
    b->budget = mem_budget_bytes;
    b->max_cases = max_cases;
    b->eng.reset(new Engine(device, false, false));
    int rc = b->eng->init_stream();
    if (rc) return rc;
    rc = b->eng->set_tile(FVM_BATCH_TILE_X, FVM_BATCH_TILE_Y);
    if (rc) return rc;
    b->eng->set_fused_d1(batch_fused(b.get()));
    b->eng->set_side_admission(true);
    rc = b->eng->reserve_cases(max_cases);
    if (rc) return rc;
    rc = b->eng->set_arena(mem_budget_bytes);
    if (rc) return rc;
    *out = b.release();
    return FVM_OK;
}

int fvm_batch_destroy(fvm_batch* b) {
    if (!b) return FVM_OK;
    b->eng.reset();
    delete b;
    return FVM_OK;
}

int fvm_batch_case_bytes(const fvm_params* p, size_t* bytes) {
    if (!p || !bytes) return set_error(FVM_ERR_INVALID_ARGUMENT, "null argument");
    *bytes = Engine::case_bytes(*p, false, false, false);
    return FVM_OK;
}

int fvm_batch_add_case(fvm_batch* b, const fvm_params* p, int32_t steps_total, int32_t* case_id) {
    if (!b || !p || !case_id) return set_error(FVM_ERR_INVALID_ARGUMENT, "null argument");
    if (p->precision != FVM_PRECISION_FP32) return set_error(FVM_ERR_INVALID_ARGUMENT, "the batched engine runs FP32 cases");
    if (b->eng->active_count() >= b->max_cases) return set_error(FVM_ERR_CAPACITY, "case limit reached");
    int id = -1;
    int rc = b->eng->add_case(*p, steps_total, false, &id);
    if (rc) return rc;
    *case_id = id;
    return FVM_OK;
}

int fvm_batch_remove_case(fvm_batch* b, int32_t case_id) {
    if (!b) return set_error(FVM_ERR_INVALID_ARGUMENT, "null batch");
    return b->eng->remove_case(case_id);
}

int fvm_batch_advance(fvm_batch* b, int32_t nsteps) {
    if (!b) return set_error(FVM_ERR_INVALID_ARGUMENT, "null batch");
    if (nsteps < 0) return set_error(FVM_ERR_INVALID_ARGUMENT, "negative step count");
    Engine& e = *b->eng;
    int rc0 = e.throttle_begin();
    if (rc0) return rc0;
    int left = nsteps;
    while (left > 0 && e.active_count() > 0) {
        // run until the next case completes (or nsteps are done)
        long long chunk = left;
        for (int id = 0; id < e.case_slots(); id++) {
            fvm::HostCase* c = e.get(id);
            if (c && c->active && !c->finished && c->steps_total > 0)
                chunk = std::min<long long>(chunk, c->steps_total - c->steps_done);
        }
        if (chunk <= 0) chunk = 1;
        int rc = e.advance((int)chunk, b->use_graph);
        if (rc) return rc;
        left -= (int)chunk;
        e.retire_finished();
    }
    return e.throttle_end();
}

int fvm_batch_prepare_graphs(fvm_batch* b) {
    if (!b) return set_error(FVM_ERR_INVALID_ARGUMENT, "null batch");
    if (!b->use_graph || b->eng->active_count() == 0) return FVM_OK;
    return b->eng->prepare_graphs();
}

int fvm_batch_case_status(fvm_batch* b, int32_t case_id, int32_t* steps_done, double* t, int32_t* finished) {
    if (!b) return set_error(FVM_ERR_INVALID_ARGUMENT, "null batch");
    fvm::HostCase* c = b->eng->get(case_id);
    if (!c) return set_error(FVM_ERR_NOT_FOUND, "no such case");
    if (steps_done) *steps_done = (int32_t)c->steps_done;
    if (finished) *finished = c->finished ? 1 : 0;
    if (t) {
        long long n = 0;
        int rc = b->eng->read_time(case_id, t, &n);
        if (rc) return rc;
    }
    return FVM_OK;
}

int fvm_batch_get_state(fvm_batch* b, int32_t case_id, float* host5) {
    if (!b || !host5) return set_error(FVM_ERR_INVALID_ARGUMENT, "null argument");
    return b->eng->read_state(case_id, host5, false);
}

int fvm_batch_get_cell_types(fvm_batch* b, int32_t case_id, uint8_t* host) {
    if (!b || !host) return set_error(FVM_ERR_INVALID_ARGUMENT, "null argument");
    return b->eng->read_cell_types(case_id, host);
}

int fvm_batch_forces(fvm_batch* b, int32_t case_id, double out3[3]) {
    if (!b || !out3) return set_error(FVM_ERR_INVALID_ARGUMENT, "null argument");
    return b->eng->forces(case_id, out3);
}

int fvm_batch_save_case(fvm_batch* b, int32_t case_id, const char* path) {
    if (!b) return set_error(FVM_ERR_INVALID_ARGUMENT, "null batch");
    return b->eng->save_case(case_id, path);
}

int fvm_batch_load_case(fvm_batch* b, const char* path, int32_t steps_total, int32_t* case_id) {
    if (!b || !case_id) return set_error(FVM_ERR_INVALID_ARGUMENT, "null argument");
    fvm_snapshot_header h;
    fvm_params p;
    int rc = fvm::read_snapshot_header(path, &h, &p);
    if (rc) return rc;
    int32_t id = -1;
    rc = fvm_batch_add_case(b, &p, steps_total, &id);
    if (rc) return rc;
    rc = b->eng->load_case(id, path);
    if (rc) {
        (void)b->eng->remove_case(id);
        return rc;
    }
    *case_id = id;
    return FVM_OK;
}

int fvm_batch_active_cases(fvm_batch* b, int32_t* n) {
    if (!b || !n) return set_error(FVM_ERR_INVALID_ARGUMENT, "null argument");
    *n = b->eng->active_count();
    return FVM_OK;
}

int fvm_batch_bytes_in_use(fvm_batch* b, size_t* bytes) {
    if (!b || !bytes) return set_error(FVM_ERR_INVALID_ARGUMENT, "null argument");
    *bytes = b->eng->bytes_in_use();
    return FVM_OK;
}

int fvm_batch_synchronize(fvm_batch* b) {
    if (!b) return set_error(FVM_ERR_INVALID_ARGUMENT, "null batch");
    return b->eng->sync();
}

int fvm_batch_set_viscous_schedule(fvm_batch* b, int schedule) {
    if (!b) return set_error(FVM_ERR_INVALID_ARGUMENT, "null batch");
    if (schedule != FVM_SCHEDULE_AUTO && schedule != FVM_SCHEDULE_SPLIT && schedule != FVM_SCHEDULE_FUSED)
        return set_error(FVM_ERR_INVALID_ARGUMENT, "unknown viscous schedule");
    b->schedule = schedule;
    b->eng->set_fused_d1(batch_fused(b));
    return FVM_OK;
}

int fvm_batch_set_tile(fvm_batch* b, int tile_x, int tile_y) {
    if (!b) return set_error(FVM_ERR_INVALID_ARGUMENT, "null batch");
    b->default_tile = tile_x == 0 && tile_y == 0;
    if (b->default_tile) {
        tile_x = FVM_BATCH_TILE_X;
        tile_y = FVM_BATCH_TILE_Y;
    }
    b->eng->set_fused_d1(batch_fused(b));
    return b->eng->set_tile(tile_x, tile_y);
}

int fvm_batch_use_graph(fvm_batch* b, int enable) {
    if (!b) return set_error(FVM_ERR_INVALID_ARGUMENT, "null batch");
    b->use_graph = enable != 0;
    return FVM_OK;
}

}  // extern "C"
