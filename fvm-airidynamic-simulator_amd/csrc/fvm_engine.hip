// fvm_engine.hip -- host runtime of the FVM step (see fvm_engine.hpp).
#include "fvm_engine.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>

#pragma clang fp contract(off)

namespace fvm {

static thread_local std::string g_last_error;

int set_error(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? FVM_ERR_OUT_OF_MEMORY : FVM_ERR_HIP;
}

const char* last_error() { return g_last_error.c_str(); }

// ------------------------------------------------------------------ arena
static size_t round256(size_t x) { return (x + 255) & ~size_t(255); }

Arena::~Arena() {
    if (base_) (void)hipFree(base_);
}

int Arena::init(size_t bytes) {
    bytes = round256(bytes);
    FVM_HIP(hipMalloc((void**)&base_, bytes));
    cap_ = bytes;
    used_ = 0;
    free_.clear();
    live_.clear();
    free_[0] = bytes;
    return FVM_OK;
}

void* Arena::alloc(size_t bytes) {
    bytes = round256(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second >= bytes) {
            size_t off = it->first, sz = it->second;
            free_.erase(it);
            if (sz > bytes) free_[off + bytes] = sz - bytes;
            live_[off] = bytes;
            used_ += bytes;
            return base_ + off;
        }
    }
    return nullptr;
}

void Arena::release(void* p) {
    if (!p) return;
    size_t off = (size_t)((char*)p - base_);
    auto it = live_.find(off);
    if (it == live_.end()) return;
    size_t sz = it->second;
    live_.erase(it);
    used_ -= sz;
    auto nx = free_.lower_bound(off);
    if (nx != free_.end() && off + sz == nx->first) {  // merge with next
        sz += nx->second;
        nx = free_.erase(nx);
    }
    if (nx != free_.begin()) {  // merge with previous
        auto pv = std::prev(nx);
        if (pv->first + pv->second == off) {
            pv->second += sz;
            return;
        }
    }
    free_[off] = sz;
}

// ------------------------------------------------------------------ engine
Engine::Engine(int device, bool fp64, bool kahan) : device_(device), fp64_(fp64), kahan_(kahan && !fp64) {
    desc_size_ = fp64_ ? sizeof(CaseDesc<double>) : sizeof(CaseDesc<float>);
}

Engine::~Engine() {
    (void)hipSetDevice(device_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (side_) (void)hipStreamSynchronize(side_);
    drop_graphs();
    for (HostCase* c : cases_) {
        if (!c) continue;
        if (c->ev_frozen) (void)hipEventDestroy(c->ev_frozen);
        if (use_arena_) {
            arena_.release(c->mem);
            arena_.release(c->ghost_mem);
        } else {
            if (c->mem) (void)hipFree(c->mem);
            if (c->ghost_mem) (void)hipFree(c->ghost_mem);
        }
        delete c;
    }
    if (d_cases_) (void)hipFree(d_cases_);
    if (!use_arena_) {   // arena-held lists go with the arena
        for (WorkItem* p : {d_tiles_all_, d_tiles_visc_, d_ghosts_, d_case_list_, scr_tiles_, scr_ghosts_})
            if (p) (void)hipFree(p);
    }
    if (scr_cnt_) (void)hipFree(scr_cnt_);
    if (d_counts_) (void)hipFree(d_counts_);
    if (d_stat_) (void)hipFree(d_stat_);
    if (d_cons_) (void)hipFree(d_cons_);
    if (h_pinned_) (void)hipHostFree(h_pinned_);
    for (auto& ts : tset_)
        for (auto& ev : ts.ev)
            if (ev) (void)hipEventDestroy(ev);
    if (ev_null_) (void)hipEventDestroy(ev_null_);
    for (hipEvent_t e : ev_ring_)
        if (e) (void)hipEventDestroy(e);
    for (auto& kv : staging_) {
        if (kv.second.host) (void)hipHostFree(kv.second.host);
        if (kv.second.ev) (void)hipEventDestroy(kv.second.ev);
    }
    if (side_) (void)hipStreamDestroy(side_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

int Engine::init_stream() {
    FVM_HIP(hipSetDevice(device_));
    if (!stream_) FVM_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    if (!main_) main_ = stream_;
    if (!side_) FVM_HIP(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking));
    if (!ev_null_) FVM_HIP(hipEventCreateWithFlags(&ev_null_, hipEventDisableTiming));
    for (hipEvent_t& e : ev_ring_)
        if (!e) FVM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    if (!d_counts_) {
        FVM_HIP(hipMalloc((void**)&d_counts_, 8 * sizeof(int)));
        FVM_HIP(hipMemset(d_counts_, 0, 8 * sizeof(int)));
    }
    if (!d_stat_) FVM_HIP(hipMalloc((void**)&d_stat_, 16 * sizeof(unsigned)));
    if (!scr_cnt_) FVM_HIP(hipMalloc((void**)&scr_cnt_, 4 * sizeof(int)));
    if (!d_cons_) FVM_HIP(hipMalloc((void**)&d_cons_, (kConsBlocks * 4 + 8) * sizeof(double)));
    if (!h_pinned_) FVM_HIP(hipHostMalloc((void**)&h_pinned_, 16 * sizeof(float)));
    for (auto& ts : tset_)
        for (auto& ev : ts.ev)
            if (!ev) FVM_HIP(hipEventCreate(&ev));
    return FVM_OK;
}

// Descriptor table for n cases, allocated up front (the batch does this at
// creation, so admission never grows it).
int Engine::reserve_cases(int n) {
    if (n <= max_cases_) return FVM_OK;
    void* nd = nullptr;
    FVM_HIP(hipStreamSynchronize(stream_));
    FVM_HIP(hipMalloc(&nd, desc_size_ * n));
    if (d_cases_) {
        FVM_HIP(hipMemcpy(nd, d_cases_, desc_size_ * max_cases_, hipMemcpyDeviceToDevice));
        FVM_HIP(hipFree(d_cases_));
    }
    d_cases_ = nd;
    h_descs_.resize(desc_size_ * n);
    max_cases_ = n;
    drop_graphs();
    return FVM_OK;
}

int Engine::set_arena(size_t bytes) {
    int rc = arena_.init(bytes);
    if (rc == FVM_OK) {
        use_arena_ = true;
        persistent_ = true;
    }
    return rc;
}

int Engine::set_tile(int tx, int ty) {
    if (!tile_supported(tx, ty, fp64_)) return set_error(FVM_ERR_INVALID_ARGUMENT, "unsupported tile shape");
    if (tx == tx_ && ty == ty_) return FVM_OK;
    tx_ = tx;
    ty_ = ty;
    for (HostCase* c : cases_) {
        if (!c) continue;
        c->tiles_x = (c->p.nx + tx_ - 1) / tx_;
        c->tiles_y = (c->p.ny + ty_ - 1) / ty_;
        int rc = upload_desc(*c);
        if (rc) return rc;
    }
    lists_dirty_ = true;
    drop_graphs();
    return FVM_OK;
}

void Engine::set_fused_d1(bool on) {
    if (on == fuse_d1_) return;
    fuse_d1_ = on;
    lists_dirty_ = true;   // the persistent grids follow the kernels the schedule launches
    drop_graphs();
}

bool Engine::fused_d1() const {
    return fuse_d1_ && !fp64_ && !kahan_ && fused_d1_supported(tx_, ty_, fp64_);
}

HostCase* Engine::get(int id) {
    if (id < 0 || id >= (int)cases_.size()) return nullptr;
    return cases_[id];
}

int Engine::current_set(const HostCase& c) const {
    if (c.frozen_set >= 0) return c.frozen_set;
    if (c.p.enable_viscosity) return 0;
    return phase_ ? 2 : 0;
}

int Engine::active_count() const {
    int n = 0;
    for (const HostCase* c : cases_)
        if (c && c->active && !c->finished) n++;
    return n;
}

int Engine::retire_finished() {
    int n = 0;
    for (HostCase* c : cases_)
        if (c && c->active && !c->finished && c->steps_total > 0 && c->steps_done >= c->steps_total) {
            c->frozen_set = current_set(*c);
            c->finished = true;
            // marks the end of the case's last step: its read-back and release
            // wait for this, not for the steps queued after it
            if (!c->ev_frozen && hipEventCreateWithFlags(&c->ev_frozen, hipEventDisableTiming) != hipSuccess)
                c->ev_frozen = nullptr;
            if (c->ev_frozen && hipEventRecord(c->ev_frozen, stream_) != hipSuccess) {
                (void)hipEventDestroy(c->ev_frozen);
                c->ev_frozen = nullptr;
            }
            n++;
        }
    if (n) lists_dirty_ = true;
    return n;
}

int Engine::throttle_begin() {
    if (ring_used_[ring_]) FVM_HIP(hipEventSynchronize(ev_ring_[ring_]));
    return FVM_OK;
}

int Engine::throttle_end() {
    FVM_HIP(hipEventRecord(ev_ring_[ring_], stream_));
    ring_used_[ring_] = true;
    ring_ ^= 1;
    return FVM_OK;
}

size_t Engine::bytes_in_use() const {
    size_t b = 0;
    for (const HostCase* c : cases_)
        if (c) b += c->bytes;
    return b;
}

// Planes of the first `nsets` buffer sets: five each, except that set D (the
// fourth) has no density plane of its own, since the first viscous half step
// leaves rho unchanged and D reads A's (upload_desc).
static size_t set_planes(int nsets) { return (size_t)5 * nsets - (nsets > 3 ? 1 : 0); }

// Offset, in planes, of field f of buffer set s (set D's density is A's).
static size_t plane_index(int s, int f) {
    if (s < 3) return (size_t)s * 5 + f;
    return f == 0 ? 0 : (size_t)15 + f - 1;
}

// Plane layout of one case inside its allocation.
struct CaseLayout {
    size_t n, plane, sets_off, k_off, ct_off, sdf_off, slots_off, t_off, steps_off, tmax_off, tmax64_off, total;
};

// Upper bound of a case's tile count over the compiled tile shapes: the
// per-tile CFL maxima arrays are sized by it.
static size_t max_tiles(const fvm_params& p) {
    int mx = 0, my = 0;
    tile_min_dims(&mx, &my);
    return (size_t)((p.nx + mx - 1) / mx) * (size_t)((p.ny + my - 1) / my);
}

static CaseLayout layout_for(const fvm_params& p, bool fp64, bool kahan, int nsets) {
    CaseLayout L;
    L.n = (size_t)p.nx * p.ny;
    L.plane = round256(L.n * (fp64 ? 8 : 4));
    size_t off = 0;
    L.sets_off = off;
    off += L.plane * set_planes(nsets);
    L.k_off = off;
    if (kahan) off += round256(L.n * 4) * 5;
    L.ct_off = off;
    off += round256(L.n);
    L.sdf_off = off;
    off += round256(L.n * 4);
    L.slots_off = off;
    off += 256;
    L.t_off = off;
    off += 256;
    L.steps_off = off;
    off += 256;
    // per-tile CFL maxima, room for the tile count of any compiled shape
    L.tmax_off = off;
    off += round256(max_tiles(p) * 2 * sizeof(float));
    L.tmax64_off = off;
    off += round256(max_tiles(p) * 2 * sizeof(double));
    L.total = off;
    return L;
}

size_t Engine::case_bytes(const fvm_params& p, bool fp64, bool kahan, bool all_sets) {
    int nsets = (all_sets || p.enable_viscosity) ? 4 : 3;
    return layout_for(p, fp64, kahan, nsets).total;
}

template <class T>
static void fill_consts(CaseConsts<T>& k, const fvm_params& p) {
    std::memset(&k, 0, sizeof(k));
    k.nx = p.nx;
    k.ny = p.ny;
    k.visc = p.enable_viscosity ? 1 : 0;
    k.adiabatic = p.adiabatic_wall ? 1 : 0;
    k.flux = p.flux_scheme;
    k.limiter = p.limiter;
    k.wall_mode = p.wall_mode;
    k.dt_mode = p.dt_mode;
    k.noise_field = p.noise_field;
    k.noise_seed = p.noise_seed;
    k.noise_amp = p.noise_amplitude;
    k.cfl = p.cfl;
    k.cfl_visc = p.cfl_visc;
    k.min_dx = std::fmin(p.dx, p.dy);
    k.host_dt = p.dt;
    const T one = T(1);
    k.dx = (T)p.dx;
    k.dy = (T)p.dy;
    k.inv_dx = one / k.dx;
    k.inv_dy = one / k.dy;
    k.inv_2dx = T(0.5) / k.dx;
    k.inv_2dy = T(0.5) / k.dy;
    k.gamma = (T)p.gamma;
    k.gm1 = k.gamma - one;
    k.inv_gm1 = one / k.gm1;
    k.R = (T)p.r_gas;
    k.inv_R = one / k.R;
    T cv = k.R / k.gm1;
    k.cp = k.gamma * cv;
    k.prandtl = (T)p.prandtl;
    k.two_over_gm1 = T(2) / k.gm1;
    k.gm1_over_4 = k.gm1 * T(0.25);
    k.inv_gamma = one / k.gamma;
    k.cs_fac = T(2) * k.gm1 / (k.gamma + one);
    k.min_rho = (T)p.min_density;
    k.min_p = (T)p.min_pressure;
    k.max_T = (T)p.max_temperature;
    k.min_T = (T)p.min_temperature;
    k.low_rho = (T)p.low_density_threshold;
    k.vmax = (T)p.max_velocity_limit;
    k.vghost = (T)p.max_ghost_velocity;
    k.slip_ref = (T)p.slip_reference_velocity;
    k.mu_ref = (T)p.mu_ref;
    k.t_ref = (T)p.t_ref;
    k.s_su = (T)p.s_sutherland;
    k.rho_inf = (T)p.rho_inf;
    k.u_inf = (T)p.u_inf;
    k.v_inf = (T)p.v_inf;
    k.p_inf = (T)p.p_inf;
    k.e_inf = k.p_inf / k.gm1;
    T ke_inf = T(0.5) * k.rho_inf * (k.u_inf * k.u_inf + k.v_inf * k.v_inf);
    k.E_inf = k.e_inf + ke_inf;
    k.T_wall = (T)p.T_wall;
    // FP64 keeps the reference's flux form; FP32 fluxes are taken relative to p_inf
    k.p_gauge = (sizeof(T) == 4) ? k.p_inf : T(0);
    // far-field Riemann invariants (computeFarfieldRiemannInvarX/Y, solver.cu:799-841)
    T c_inf = std::sqrt(k.gamma * k.p_inf / k.rho_inf);
    k.c_far = c_inf;
    k.s_far = std::exp(-k.gamma * std::log(k.rho_inf)) * k.p_inf;
    // geometry: the per-cell cosf/sinf of the reference (solver.cu:190-210,
    // 259-260, 330-331, 421-422, 451-452) evaluated once here
    Geom& g = k.geo;
    g.shape = (p.obstacle_shape >= 0 && p.obstacle_shape <= 5) ? p.obstacle_shape : 0;
    g.obs_x = p.obstacle_x;
    g.obs_y = p.obstacle_y;
    g.obs_r = p.obstacle_r;
    g.cosR = std::cos(-p.obstacle_rotation);
    g.sinR = std::sin(-p.obstacle_rotation);
    g.cosW = std::cos(-p.wing_rotation);
    g.sinW = std::sin(-p.wing_rotation);
    g.band = 1.5f * std::fmax(p.dx, p.dy);
    g.dx = p.dx;
    g.dy = p.dy;
    const float kPI = 3.14159265358979323846f;
    const int N = 5;
    const float outerR = p.obstacle_r, innerR = p.obstacle_r * 0.38f;
    for (int i = 0; i < N; i++) {
        float outerAngle = 2.0f * kPI * (float)i / (float)N - kPI / 2.0f;
        float innerAngle = outerAngle + kPI / (float)N;
        g.star[i * 4 + 0] = outerR * std::cos(outerAngle);
        g.star[i * 4 + 1] = outerR * std::sin(outerAngle);
        g.star[i * 4 + 2] = innerR * std::cos(innerAngle);
        g.star[i * 4 + 3] = innerR * std::sin(innerAngle);
    }
}

int Engine::build_consts(HostCase& hc, const fvm_params& p, void* desc_out) {
    if (fp64_) fill_consts(((CaseDesc<double>*)desc_out)->c, p);
    else fill_consts(((CaseDesc<float>*)desc_out)->c, p);
    (void)hc;
    return FVM_OK;
}

int Engine::upload_desc(HostCase& hc) {
    // the kernels store one maximum pair per tile into arrays of max_tiles()
    if ((size_t)hc.tiles_x * hc.tiles_y > max_tiles(hc.p))
        return set_error(FVM_ERR_INVALID_ARGUMENT, "tile count exceeds the per-tile maxima arrays");
    char* hd = h_descs_.data() + (size_t)hc.id * desc_size_;
    char* base = (char*)hc.mem;
    CaseLayout L = layout_for(hc.p, fp64_, kahan_, hc.nsets);
    CaseBufs b;
    std::memset(&b, 0, sizeof(b));
    // The first viscous half step (A -> D) does not change rho, so D's density
    // plane is A's (plane_index): that half step writes four fields, and the
    // stage kernels and the blend read U^n's density in place.  A's density is
    // next written by the final half step, after the blend, the last reader of D.
    for (int s = 0; s < 4; s++)
        for (int f = 0; f < 5; f++)
            b.U[s][f] = (s < hc.nsets) ? (void*)(base + L.sets_off + L.plane * plane_index(s, f)) : nullptr;
    if (kahan_)
        for (int f = 0; f < 5; f++) b.K[f] = (float*)(base + L.k_off + round256(L.n * 4) * f);
    b.ct = (uint8_t*)(base + L.ct_off);
    b.ghosts = hc.ghost_mem;
    b.gmap = hc.ghost_mem ? (int*)((char*)hc.ghost_mem + hc.gmap_off) : nullptr;
    for (int q = 0; q < 4; q++) b.gbox[q] = hc.gbox[q];
    b.sdf = (float*)(base + L.sdf_off);
    b.slots = (unsigned*)(base + L.slots_off);
    b.t = (double*)(base + L.t_off);
    b.steps = (long long*)(base + L.steps_off);
    b.tile_max = (float*)(base + L.tmax_off);
    if (fp64_) {
        CaseDesc<double>* d = (CaseDesc<double>*)hd;
        fill_consts(d->c, hc.p);
        d->b = b;
        d->tiles_x = hc.tiles_x;
        d->tiles_y = hc.tiles_y;
        d->dt64 = dt64() ? 1 : 0;
        d->slots64 = (double*)(base + L.slots_off + kSlots64Off);
        d->tile_max64 = (double*)(base + L.tmax64_off);
    } else {
        CaseDesc<float>* d = (CaseDesc<float>*)hd;
        fill_consts(d->c, hc.p);
        d->b = b;
        d->tiles_x = hc.tiles_x;
        d->tiles_y = hc.tiles_y;
        d->dt64 = dt64() ? 1 : 0;
        d->slots64 = (double*)(base + L.slots_off + kSlots64Off);
        d->tile_max64 = (double*)(base + L.tmax64_off);
    }
    FVM_HIP(hipMemcpyAsync((char*)d_cases_ + (size_t)hc.id * desc_size_, hd, desc_size_, hipMemcpyHostToDevice, stream_));
    FVM_HIP(hipStreamSynchronize(stream_));
    return FVM_OK;
}

static bool params_valid(const fvm_params& p, std::string& why) {
    if (p.nx < 3 || p.ny < 3) { why = "nx and ny must be >= 3"; return false; }
    // the kernels form a cell's plane index with a 24-bit multiply
    if (p.nx >= (1 << 24) || p.ny >= (1 << 24)) { why = "nx and ny must be below 16777216"; return false; }
    if ((long long)p.nx * p.ny > (1ll << 31) - 1) { why = "grid too large (nx*ny must fit in int32)"; return false; }
    // the kernels address a field plane with a 32-bit byte offset
    if ((long long)p.nx * p.ny * (p.precision == FVM_PRECISION_FP64 ? 8 : 4) > (1ll << 32)) {
        why = "grid too large (one field plane, nx*ny*element bytes, must fit in 4 GiB)";
        return false;
    }
    if (!(p.dx > 0) || !(p.dy > 0)) { why = "dx and dy must be positive"; return false; }
    if (!(p.gamma > 1.0f)) { why = "gamma must exceed 1"; return false; }
    if (!(p.rho_inf > 0) || !(p.p_inf > 0)) { why = "rho_inf and p_inf must be positive (call fvm_params_compute_derived)"; return false; }
    if (p.flux_scheme < 0 || p.flux_scheme > 1) { why = "unknown flux scheme"; return false; }
    if (p.limiter < 0 || p.limiter > 1) { why = "unknown limiter"; return false; }
    if (p.dt_mode < 0 || p.dt_mode > 1) { why = "unknown dt mode"; return false; }
    if (p.precision < 0 || p.precision > 2) { why = "unknown precision"; return false; }
    if (p.wall_mode < 0 || p.wall_mode > 1) { why = "unknown wall mode"; return false; }
    if (p.noise_field < 0 || p.noise_field > 2) { why = "unknown noise field"; return false; }
    if (p.obstacle_shape < 0 || p.obstacle_shape > 5) { why = "unknown obstacle shape"; return false; }
    if (!(p.cfl > 0) || (p.enable_viscosity && !(p.cfl_visc > 0))) { why = "CFL numbers must be positive"; return false; }
    if (p.dt_mode == FVM_DT_HOST && !(p.dt > 0)) { why = "dt must be positive with host dt"; return false; }
    if (!(p.min_density > 0) || !(p.min_pressure > 0)) { why = "density and pressure floors must be positive"; return false; }
    if (p.precision == FVM_PRECISION_FP32_KAHAN && p.enable_viscosity) {
        why = "the compensated (Kahan) update is implemented for inviscid cases only";
        return false;
    }
    return true;
}

// Runs the admission of a batched case on the side stream: every enqueue and
// wait of add_case / reset_case / mark_and_lists goes to stream_, which is the
// side stream for the duration.  The admission only touches the new case's
// memory, its descriptor slot and the engine's admission scratch lists, none
// of which the steps queued on the main stream read.
struct SideStream {
    hipStream_t& s;
    hipStream_t main;
    SideStream(hipStream_t& stream, hipStream_t side, bool on) : s(stream), main(stream) { if (on) s = side; }
    ~SideStream() { s = main; }
};

int Engine::add_case(const fvm_params& p, int steps_total, bool all_sets, int* id_out) {
    std::string why;
    if (!params_valid(p, why)) return set_error(FVM_ERR_INVALID_ARGUMENT, why);
    int rc = init_stream();
    if (rc) return rc;
    // a descriptor-table growth below waits for and copies the table the main
    // stream's steps read: never on the side stream
    const bool side = side_admission_ && (int)cases_.size() < max_cases_;
    SideStream on_side(stream_, side_, side);
    int id = -1;
    for (int i = 0; i < (int)cases_.size(); i++)
        if (!cases_[i]) { id = i; break; }
    if (id < 0) {
        id = (int)cases_.size();
        if (id >= max_cases_) {
            // grow the descriptor table (never while a graph references it)
            int nmax = std::max(4, max_cases_ * 2);
            while (nmax <= id) nmax *= 2;
            int rg = reserve_cases(nmax);
            if (rg) return rg;
        }
        cases_.push_back(nullptr);
    }
    HostCase* hc = new HostCase();
    hc->p = p;
    hc->id = id;
    hc->nsets = (all_sets || p.enable_viscosity) ? 4 : 3;
    hc->steps_total = steps_total;
    hc->tiles_x = (p.nx + tx_ - 1) / tx_;
    hc->tiles_y = (p.ny + ty_ - 1) / ty_;
    CaseLayout L = layout_for(p, fp64_, kahan_, hc->nsets);
    hc->bytes = L.total;
    if (use_arena_) {
        hc->mem = arena_.alloc(L.total);
        if (!hc->mem) {
            delete hc;
            cases_[id] = nullptr;
            return set_error(FVM_ERR_CAPACITY, "device memory budget exhausted");
        }
    } else {
        hipError_t e = hipMalloc(&hc->mem, L.total);
        if (e != hipSuccess) {
            delete hc;
            cases_[id] = nullptr;
            return hip_fail(e, "hipMalloc(case state)");
        }
    }
    FVM_HIP(hipMemsetAsync(hc->mem, 0, L.total, stream_));
    cases_[id] = hc;
    rc = upload_desc(*hc);
    if (rc) return rc;
    hc->active = true;
    rc = reset_case(id, p);
    if (rc) return rc;
    *id_out = id;
    return FVM_OK;
}

int Engine::remove_case(int id) {
    HostCase* hc = get(id);
    if (!hc) return set_error(FVM_ERR_NOT_FOUND, "no such case");
    // A finished case is read by no step queued after its last one (the lists
    // are rebuilt without it): it is released once that step has run.
    // Anything else waits for every queued step.
    if (hc->finished && hc->ev_frozen) FVM_HIP(hipEventSynchronize(hc->ev_frozen));
    else FVM_HIP(hipStreamSynchronize(stream_));
    FVM_HIP(hipStreamSynchronize(side_));
    if (hc->ev_frozen) (void)hipEventDestroy(hc->ev_frozen);
    hc->ev_frozen = nullptr;
    if (use_arena_) {
        arena_.release(hc->mem);
        arena_.release(hc->ghost_mem);
    } else {
        FVM_HIP(hipFree(hc->mem));
        if (hc->ghost_mem) FVM_HIP(hipFree(hc->ghost_mem));
    }
    delete hc;
    cases_[id] = nullptr;
    lists_dirty_ = true;
    if (!persistent_) drop_graphs();
    return FVM_OK;
}

// Device memory of the engine's own lists: from the arena when the engine has
// one (admission and release then never call hipMalloc / hipFree), else
// hipMalloc.
void* Engine::dev_alloc(size_t bytes) {
    if (use_arena_) return arena_.alloc(bytes);
    void* p = nullptr;
    return hipMalloc(&p, bytes) == hipSuccess ? p : nullptr;
}

void Engine::dev_free(void* p) {
    if (!p) return;
    if (use_arena_) arena_.release(p);
    else (void)hipFree(p);
}

// Uploads a work list into *buf (grown geometrically) and its length into
// d_count[count_slot], in stream order after the work queued before it (which
// still reads the old list), without waiting for that work: the copy is made
// from a pinned staging buffer of this list.  A grown device buffer replaces
// the old one only after everything queued has run.
int Engine::upload_items(const std::vector<WorkItem>& v, WorkItem** buf, size_t* cap, int* d_count, int count_slot) {
    if (v.size() > *cap) {
        FVM_HIP(hipStreamSynchronize(stream_));
        dev_free(*buf);
        *buf = nullptr;
        *cap = 0;
        size_t nc = std::max<size_t>(v.size(), 64) * 2;
        *buf = (WorkItem*)dev_alloc(nc * sizeof(WorkItem));
        if (!*buf) return set_error(FVM_ERR_CAPACITY, "device memory exhausted (work list)");
        *cap = nc;
    }
    UploadStaging& st = staging_[buf];
    if (!st.ev) FVM_HIP(hipEventCreateWithFlags(&st.ev, hipEventDisableTiming));
    if (st.used) FVM_HIP(hipEventSynchronize(st.ev));
    const size_t bytes = v.size() * sizeof(WorkItem), need = bytes + 64;
    if (need > st.cap) {
        if (st.host) FVM_HIP(hipHostFree(st.host));
        st.host = nullptr;
        st.cap = 0;
        const size_t nc = std::max<size_t>(need * 2, 4096);
        FVM_HIP(hipHostMalloc((void**)&st.host, nc));
        st.cap = nc;
    }
    if (bytes) std::memcpy(st.host, v.data(), bytes);
    int* const count = (int*)(st.host + st.cap - 64);
    *count = (int)v.size();
    if (bytes) FVM_HIP(hipMemcpyAsync(*buf, st.host, bytes, hipMemcpyHostToDevice, stream_));
    FVM_HIP(hipMemcpyAsync(d_count + count_slot, count, sizeof(int), hipMemcpyHostToDevice, stream_));
    FVM_HIP(hipEventRecord(st.ev, stream_));
    st.used = true;
    return FVM_OK;
}

// Tile order: consecutive dispatch slots go round-robin over the 8 XCDs, so a
// contiguous block of tiles is handed to each XCD (halo rows stay in that
// XCD's L2).  Holes are marked with index -1.
static void xcd_order(const std::vector<WorkItem>& in, std::vector<WorkItem>& out) {
    const int X = 8;
    size_t n = in.size();
    size_t chunk = (n + X - 1) / X;
    out.assign(chunk * X, WorkItem{0, -1});
    for (size_t t = 0; t < n; t++) {
        size_t x = t / chunk, q = t % chunk;
        out[q * X + x] = in[t];
    }
    while (!out.empty() && out.back().index < 0) out.pop_back();
}

int Engine::reset_case(int id, const fvm_params& p) {
    HostCase* hc = get(id);
    if (!hc) return set_error(FVM_ERR_NOT_FOUND, "no such case");
    std::string why;
    if (!params_valid(p, why)) return set_error(FVM_ERR_INVALID_ARGUMENT, why);
    if (p.nx != hc->p.nx || p.ny != hc->p.ny) return set_error(FVM_ERR_INVALID_ARGUMENT, "reset cannot change the grid size");
    if (p.enable_viscosity && hc->nsets < 4) return set_error(FVM_ERR_INVALID_ARGUMENT, "case was admitted without the viscous buffer set");
    hc->p = p;
    hc->finished = false;
    hc->frozen_set = -1;
    int rc = upload_desc(*hc);
    if (rc) return rc;
    size_t n = (size_t)p.nx * p.ny;
    int cur = current_set(*hc);
    launch_init(fp64_, stream_, d_cases_, id, n, cur, 1);
    launch_mark(fp64_, stream_, d_cases_, id, n, 0, 0);
    FVM_HIP(hipGetLastError());
    // time, steps, dt slots of this case
    {
        char* base = (char*)hc->mem;
        CaseLayout L = layout_for(hc->p, fp64_, kahan_, hc->nsets);
        FVM_HIP(hipMemsetAsync(base + L.slots_off, 0, 256 * 3, stream_));
        if (kahan_) FVM_HIP(hipMemsetAsync(base + L.k_off, 0, round256(L.n * 4) * 5, stream_));
    }
    rc = mark_and_lists(*hc, 0);
    if (rc) return rc;
    hc->steps_done = 0;
    FVM_HIP(hipStreamSynchronize(stream_));
    return FVM_OK;
}

// After a (re)mark: build the ghost list on the host, apply the BC of the
// reset path (point-wise in place on the current set, ghost fill with the
// pre-BC copy in set B) and the CFL maxima into the slot the next step reads.
int Engine::mark_and_lists(HostCase& hc, int fixup) {
    (void)fixup;
    size_t n = (size_t)hc.p.nx * hc.p.ny;
    std::vector<uint8_t> ct(n);
    CaseLayout L = layout_for(hc.p, fp64_, kahan_, hc.nsets);
    FVM_HIP(hipMemcpyAsync(ct.data(), (char*)hc.mem + L.ct_off, n, hipMemcpyDeviceToHost, stream_));
    FVM_HIP(hipStreamSynchronize(stream_));
    hc.ghosts.clear();
    int bi0 = hc.p.nx, bj0 = hc.p.ny, bi1 = -1, bj1 = -1;
    for (size_t i = 0; i < n; i++)
        if (ct[i] == CT_GHOST) {
            hc.ghosts.push_back((int)i);
            const int ci = (int)(i % hc.p.nx), cj = (int)(i / hc.p.nx);
            bi0 = std::min(bi0, ci); bi1 = std::max(bi1, ci);
            bj0 = std::min(bj0, cj); bj1 = std::max(bj1, cj);
        }
    // device block: ghost table (cell index now; normal and ranked neighbours
    // from k_ghost_prep), then the ghost-table index of every cell of the
    // ghost cells' bounding box (-1 for the others)
    const size_t gsz = fp64_ ? sizeof(GhostInfo<double>) : sizeof(GhostInfo<float>);
    const int bw = hc.ghosts.empty() ? 0 : bi1 - bi0 + 1, bh = hc.ghosts.empty() ? 0 : bj1 - bj0 + 1;
    const size_t table_bytes = round256(hc.ghosts.size() * gsz);
    const size_t need = table_bytes + (size_t)bw * bh * sizeof(int);
    if (need > hc.ghost_cap) {
        if (hc.ghost_mem) {
            FVM_HIP(hipStreamSynchronize(stream_));
            if (use_arena_) arena_.release(hc.ghost_mem);
            else FVM_HIP(hipFree(hc.ghost_mem));
            hc.ghost_mem = nullptr;
            hc.ghost_cap = 0;
        }
        size_t cap = round256(need + need / 4 + 256);
        if (use_arena_) {
            hc.ghost_mem = arena_.alloc(cap);
            if (!hc.ghost_mem) return set_error(FVM_ERR_CAPACITY, "device memory budget exhausted (ghost table)");
        } else {
            FVM_HIP(hipMalloc(&hc.ghost_mem, cap));
        }
        hc.ghost_cap = cap;
        hc.bytes = L.total + cap;
    }
    hc.gbox[0] = bw ? bi0 : 0;
    hc.gbox[1] = bh ? bj0 : 0;
    hc.gbox[2] = bw;
    hc.gbox[3] = bh;
    hc.gmap_off = table_bytes;
    if (!hc.ghosts.empty()) {
        std::vector<char> blk(need, 0);
        for (size_t g = 0; g < hc.ghosts.size(); g++) std::memcpy(blk.data() + g * gsz, &hc.ghosts[g], sizeof(int));
        int* map = (int*)(blk.data() + table_bytes);
        std::fill(map, map + (size_t)bw * bh, -1);
        for (size_t g = 0; g < hc.ghosts.size(); g++) {
            const int ci = hc.ghosts[g] % hc.p.nx, cj = hc.ghosts[g] / hc.p.nx;
            map[(size_t)(cj - bj0) * bw + (ci - bi0)] = (int)g;
        }
        FVM_HIP(hipMemcpyAsync(hc.ghost_mem, blk.data(), blk.size(), hipMemcpyHostToDevice, stream_));
        FVM_HIP(hipStreamSynchronize(stream_));
    }
    {
        int rc0 = upload_desc(hc);
        if (rc0) return rc0;
    }

    int cur = current_set(hc);
    std::vector<WorkItem> tiles, ordered, gl;
    for (int t = 0; t < hc.tiles_x * hc.tiles_y; t++) tiles.push_back(WorkItem{hc.id, t});
    xcd_order(tiles, ordered);
    for (int g = 0; g < (int)hc.ghosts.size(); g++) gl.push_back(WorkItem{hc.id, g});
    // per-case lists for the reset-path launches (engine scratch, reused)
    int rc = upload_items(ordered, &scr_tiles_, &scr_tiles_cap_, scr_cnt_, 0);
    if (!rc) rc = upload_items(gl, &scr_ghosts_, &scr_ghosts_cap_, scr_cnt_, 1);
    if (rc) return rc;
    WorkItem* const d_t = scr_tiles_;
    WorkItem* const d_g = scr_ghosts_;
    int* const d_cnt = scr_cnt_;

    Roles r{};
    r.in = cur; r.out = cur; r.blend = -1; r.pre0 = 1; r.pre1 = -1;
    r.slot_r = phase_; r.slot_w = phase_; r.zero_slot = 0; r.count_step = 0;
    KernelCall kc{};
    kc.stream = stream_;
    kc.cases = d_cases_;
    kc.work = d_t;
    kc.count = d_cnt;
    kc.ri = r;
    kc.rv = r;
    kc.grid = dim3((unsigned)std::max<size_t>(1, ordered.size()));
    kc.op = OP_BC_INPLACE;
    launch_tiled(fp64_, tx_, ty_, kc);
    launch_ghost_prep(fp64_, stream_, d_cases_, d_g, (int)gl.size());
    if (!gl.empty())
        launch_ghost(fp64_, dt64(), false, dim3((unsigned)((gl.size() + 255) / 256)), stream_, d_cases_, d_g, d_cnt + 1, r, r);
    // CFL maxima of the reset state into the slot the next step reads
    launch_zero_u32(stream_, (unsigned*)((char*)hc.mem + L.slots_off) + phase_ * 2, 2);
    launch_zero_u32(stream_, (unsigned*)((char*)hc.mem + L.slots_off + kSlots64Off) + phase_ * 4, 4);
    kc.op = OP_MAXIMA;
    launch_tiled(fp64_, tx_, ty_, kc);
    FVM_HIP(hipGetLastError());
    FVM_HIP(hipStreamSynchronize(stream_));
    lists_dirty_ = true;
    if (!persistent_) drop_graphs();
    return FVM_OK;
}

int Engine::update_params(int id, const fvm_params& p) {
    HostCase* hc = get(id);
    if (!hc) return set_error(FVM_ERR_NOT_FOUND, "no such case");
    std::string why;
    if (!params_valid(p, why)) return set_error(FVM_ERR_INVALID_ARGUMENT, why);
    if (p.nx != hc->p.nx || p.ny != hc->p.ny) return set_error(FVM_ERR_INVALID_ARGUMENT, "grid size changed; call resize/initialize");
    if (p.enable_viscosity && hc->nsets < 4) return set_error(FVM_ERR_INVALID_ARGUMENT, "case has no viscous buffer set");
    bool visc_changed = (p.enable_viscosity != 0) != (hc->p.enable_viscosity != 0);
    if (visc_changed && !hc->p.enable_viscosity && phase_ == 1) {
        // an inviscid case at odd parity holds U^n in set C; viscous cases keep it in A
        size_t n = (size_t)p.nx * p.ny;
        launch_copy_set(fp64_, stream_, d_cases_, id, n, 2, 0);
    } else if (visc_changed && hc->p.enable_viscosity && phase_ == 1) {
        size_t n = (size_t)p.nx * p.ny;
        launch_copy_set(fp64_, stream_, d_cases_, id, n, 0, 2);
    }
    hc->p = p;
    int rc = upload_desc(*hc);
    if (rc) return rc;
    if (visc_changed) {
        lists_dirty_ = true;
        if (!persistent_) drop_graphs();
    }
    return FVM_OK;
}

int Engine::update_obstacle(int id, const fvm_params& p) {
    HostCase* hc = get(id);
    if (!hc) return set_error(FVM_ERR_NOT_FOUND, "no such case");
    int rc = update_params(id, p);
    if (rc) return rc;
    size_t n = (size_t)p.nx * p.ny;
    int cur = current_set(*hc);
    launch_mark(fp64_, stream_, d_cases_, id, n, 1, cur);
    launch_copy_set(fp64_, stream_, d_cases_, id, n, cur, 1);
    FVM_HIP(hipGetLastError());
    return mark_and_lists(*hc, 1);
}

int Engine::set_host_dt(int id, float dt) {
    HostCase* hc = get(id);
    if (!hc) return set_error(FVM_ERR_NOT_FOUND, "no such case");
    hc->p.dt = dt;
    char* hd = h_descs_.data() + (size_t)id * desc_size_;
    size_t off = fp64_ ? offsetof(CaseDesc<double>, c) + offsetof(CaseConsts<double>, host_dt)
                       : offsetof(CaseDesc<float>, c) + offsetof(CaseConsts<float>, host_dt);
    std::memcpy(hd + off, &dt, sizeof(float));
    FVM_HIP(hipMemcpyAsync((char*)d_cases_ + (size_t)id * desc_size_ + off, hd + off, sizeof(float),
                           hipMemcpyHostToDevice, stream_));
    FVM_HIP(hipStreamSynchronize(stream_));
    return FVM_OK;
}

int Engine::rebuild_lists() {
    std::vector<WorkItem> all, visc, ghosts, oall, ovisc, clist;
    const bool had_visc = any_visc_;
    WorkItem* const old_ptrs[4] = {d_tiles_all_, d_tiles_visc_, d_ghosts_, d_case_list_};
    any_visc_ = false;
    std::vector<WorkItem> heavy;
    for (HostCase* c : cases_) {
        if (!c || !c->active || c->finished) continue;
        int nt = c->tiles_x * c->tiles_y;
        // tiles whose staged region (tile and 2-cell halo) holds a ghost cell:
        // stage 2 computes those ghost values, a chain of dependent loads
        std::vector<char> h((size_t)nt, 0);
        for (int g : c->ghosts) {
            const int ci = g % c->p.nx, cj = g / c->p.nx;
            const int tx0 = std::max(0, (ci - 2) / tx_), tx1 = std::min(c->tiles_x - 1, (ci + 2) / tx_);
            const int ty0 = std::max(0, (cj - 2) / ty_), ty1 = std::min(c->tiles_y - 1, (cj + 2) / ty_);
            for (int b = ty0; b <= ty1; b++)
                for (int a = tx0; a <= tx1; a++) h[(size_t)b * c->tiles_x + a] = 1;
        }
        for (int t = 0; t < nt; t++) {
            if (h[t]) heavy.push_back(WorkItem{c->id, t});
            else all.push_back(WorkItem{c->id, t});
            if (c->p.enable_viscosity) visc.push_back(WorkItem{c->id, t});
        }
        if (c->p.enable_viscosity) any_visc_ = true;
        for (int g = 0; g < (int)c->ghosts.size(); g++) ghosts.push_back(WorkItem{c->id, g});
        for (int q = 0; q * REDUCE_CHUNK < nt; q++) clist.push_back(WorkItem{c->id, q});
    }
    // heavy tiles first: dispatched before the rest and spread over the XCDs
    // (dispatch slots go round-robin over them); then the XCD-contiguous order
    xcd_order(all, oall);
    oall.insert(oall.begin(), heavy.begin(), heavy.end());
    xcd_order(visc, ovisc);
    int rc = upload_items(oall, &d_tiles_all_, &cap_tiles_, d_counts_, 0);
    if (!rc) rc = upload_items(ovisc, &d_tiles_visc_, &cap_visc_, d_counts_, 1);
    if (!rc) rc = upload_items(ghosts, &d_ghosts_, &cap_ghosts_, d_counts_, 2);
    if (!rc) rc = upload_items(clist, &d_case_list_, &cap_case_list_, d_counts_, 3);
    if (rc) return rc;
    n_tiles_ = (int)oall.size();
    n_visc_ = (int)ovisc.size();
    n_ghosts_ = (int)ghosts.size();
    // one active case and graphs rebuilt with the lists: the ghost list is the
    // identity over that case's ghost table, which the final ghost launch can
    // take without reading it
    int n_active = 0, last_id = -1;
    for (HostCase* c : cases_)
        if (c && c->active && !c->finished) { n_active++; last_id = c->id; }
    single_case_ = (!persistent_ && n_active == 1) ? last_id : -1;
    if (persistent_) {
        // fixed grids: captured graphs read the lists from device memory and
        // stay valid unless a list buffer moved, the viscous kernels appear or
        // the grids change.  Each grid is what the CUs hold at once: with
        // more, the surplus workgroups would start their strided share of the
        // items only when resident ones had finished theirs.
        const int old_tiles = grid_tiles_, old_visc = grid_visc_;
        int st = 0, df = 0;
        if (cu_count_ <= 0 && hipDeviceGetAttribute(&cu_count_, hipDeviceAttributeMultiprocessorCount, device_) != hipSuccess)
            cu_count_ = 0;
        tiled_resident_per_cu(fp64_, tx_, ty_, kahan_, any_visc_ && fused_d1(), &st, &df);
        grid_tiles_ = (st > 0 && cu_count_ > 0) ? st * cu_count_ : persistent_grid_;
        grid_visc_ = (df > 0 && cu_count_ > 0) ? df * cu_count_ : persistent_grid_;
        grid_ghosts_ = 512;
        grid_reduce_ = 256;
        if (had_visc != any_visc_ || old_ptrs[0] != d_tiles_all_ || old_ptrs[1] != d_tiles_visc_ ||
            old_ptrs[2] != d_ghosts_ || old_ptrs[3] != d_case_list_ || old_tiles != grid_tiles_ || old_visc != grid_visc_)
            drop_graphs();
    } else {
        grid_tiles_ = std::max(1, n_tiles_);
        grid_visc_ = std::max(1, n_visc_);
        grid_ghosts_ = std::max(1, (n_ghosts_ + 255) / 256);
        grid_reduce_ = std::max(1, std::min((int)clist.size(), 256));
        drop_graphs();
    }
    lists_dirty_ = false;
    return FVM_OK;
}

void Engine::drop_graphs() {
    // an executable graph is destroyed only once no replay of it is queued
    if ((graph_[0] || graph_[1] || graph_long_[0] || graph_long_[1] || graph_xl_[0] || graph_xl_[1]) && main_)
        (void)hipStreamSynchronize(main_);
    for (int i = 0; i < 2; i++) {
        if (graph_[i]) {
            (void)hipGraphExecDestroy(graph_[i]);
            graph_[i] = nullptr;
        }
        if (graph_long_[i]) {
            (void)hipGraphExecDestroy(graph_long_[i]);
            graph_long_[i] = nullptr;
        }
        if (graph_xl_[i]) {
            (void)hipGraphExecDestroy(graph_xl_[i]);
            graph_xl_[i] = nullptr;
        }
    }
    graph_valid_ = false;
    long_valid_ = false;
    xl_valid_ = false;
}

int Engine::flush_timed_set(int set) {
    TimedSet& ts = tset_[set];
    if (!ts.pending) return FVM_OK;
    ts.pending = false;
    if (ts.n == 0) return FVM_OK;
    FVM_HIP(hipEventSynchronize(ts.ev[2 * ts.n - 1]));
    for (int i = 0; i < ts.n; i++) {
        float ms = 0;
        FVM_HIP(hipEventElapsedTime(&ms, ts.ev[2 * i], ts.ev[2 * i + 1]));
        bool found = false;
        for (auto& kv : timing_list_)
            if (kv.first == ts.name[i]) {
                kv.second.total_ms += ms;
                kv.second.launches++;
                found = true;
                break;
            }
        if (!found) {
            KernelStat st;
            st.total_ms = ms;
            st.launches = 1;
            timing_list_.push_back({ts.name[i], st});
        }
    }
    return FVM_OK;
}

int Engine::flush_timings() {
    int rc = flush_timed_set(tcur_ ^ 1);
    if (!rc) rc = flush_timed_set(tcur_);
    return rc;
}

void Engine::clear_timings() {
    tset_[0].pending = tset_[1].pending = false;
    timing_list_.clear();
}

// Opens the event set of one step: the set used two steps ago is read first.
// Each timed kernel is launched with a start / stop event pair on its own
// dispatch, so the interval is the kernel's execution, as in a kernel trace.
int Engine::timed_begin() {
    if (!timing_) return FVM_OK;
    int rc = flush_timed_set(tcur_);
    if (rc) return rc;
    tset_[tcur_].n = 0;
    return FVM_OK;
}

void Engine::timed_events(hipEvent_t* ev0, hipEvent_t* ev1) {
    *ev0 = *ev1 = nullptr;
    if (!timing_) return;
    TimedSet& ts = tset_[tcur_];
    if (ts.n >= kTimedMax) return;
    *ev0 = ts.ev[2 * ts.n];
    *ev1 = ts.ev[2 * ts.n + 1];
}

// Names the interval of the kernel just launched with timed_events().
int Engine::timed(const char* name, int rc_before) {
    if (!timing_ || rc_before) return rc_before;
    TimedSet& ts = tset_[tcur_];
    if (ts.n >= kTimedMax) return set_error(FVM_ERR_INVALID_ARGUMENT, "too many timed launches in one step");
    ts.name[ts.n] = name;
    ts.n++;
    return FVM_OK;
}

// One SSP-RK2 step of every active case, at step parity `ph`.
int Engine::enqueue_step(int ph) {
    const int cur = ph ? 2 : 0, nxt = ph ? 0 : 2;
    const int sr = ph, sw = 1 - ph;
    Roles i_s1{cur, 1, -1, -1, -1, sr, sw, 1, 1};
    Roles i_s2{1, nxt, cur, -1, -1, sr, sw, 0, 0};
    Roles i_g2{-1, nxt, -1, cur, 1, sr, sw, 0, 0};
    Roles v_k0{0, 3, -1, -1, -1, sr, sw, 1, 1};
    // fused: the half step's result goes to set 3 (blend), stage 1's to set 1
    Roles v_f1{0, 1, 3, -1, -1, sr, sw, 1, 1};
    Roles v_s1{3, 1, -1, -1, -1, sr, sw, 0, 0};
    Roles v_s2{1, 2, 3, -1, -1, sr, sw, 0, 0};
    Roles v_k3{2, 0, -1, -1, -1, sr, sw, 0, 0};
    Roles v_g2{-1, 0, -1, 2, -1, sr, sw, 0, 0};

    KernelCall kc{};
    kc.stream = stream_;
    kc.cases = d_cases_;
    kc.kahan = kahan_;
    int rc = timed_begin();
    if (rc) return rc;
    const bool fused = any_visc_ && fused_d1();
    if (any_visc_ && !fused) {
        kc.op = OP_DIFFUSE;
        kc.grid = dim3(grid_visc_);
        kc.work = d_tiles_visc_;
        kc.count = d_counts_ + 1;
        kc.ri = v_k0;
        kc.rv = v_k0;
        timed_events(&kc.ev0, &kc.ev1);
        launch_tiled(fp64_, tx_, ty_, kc);
        rc = timed("diffuse_half1", rc);
    }
    kc.op = fused ? OP_STAGE1_FUSED : OP_STAGE1;
    kc.grid = dim3(grid_tiles_);
    kc.work = d_tiles_all_;
    kc.count = d_counts_ + 0;
    kc.ri = i_s1;
    kc.rv = fused ? v_f1 : v_s1;
    timed_events(&kc.ev0, &kc.ev1);
    launch_tiled(fp64_, tx_, ty_, kc);
    rc = timed(fused ? "stage1_fused" : "stage1", rc);
    kc.op = OP_STAGE2;
    kc.ri = i_s2;
    kc.rv = v_s2;
    timed_events(&kc.ev0, &kc.ev1);
    launch_tiled(fp64_, tx_, ty_, kc);
    rc = timed("stage2", rc);
    if (any_visc_) {
        kc.op = OP_DIFFUSE_FINAL;
        kc.grid = dim3(grid_visc_);
        kc.work = d_tiles_visc_;
        kc.count = d_counts_ + 1;
        kc.ri = v_k3;
        kc.rv = v_k3;
        timed_events(&kc.ev0, &kc.ev1);
        launch_tiled(fp64_, tx_, ty_, kc);
        rc = timed("diffuse_half2", rc);
    }
    hipEvent_t g0, g1;
    timed_events(&g0, &g1);
    launch_ghost(fp64_, dt64(), true, dim3(grid_ghosts_ + grid_reduce_), stream_, d_cases_, d_ghosts_, d_counts_ + 2, i_g2,
                 v_g2, grid_ghosts_, d_case_list_, d_counts_ + 3, single_case_, g0, g1);
    rc = timed("ghost2", rc);
    if (rc) return rc;
    if (timing_) {
        tset_[tcur_].pending = true;
        tcur_ ^= 1;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "kernel launch");
    return FVM_OK;
}

// nsteps steps starting at parity first_phase, captured and instantiated.
int Engine::capture_sequence(int first_phase, int nsteps, hipGraphExec_t* out) {
    hipGraph_t graph = nullptr;
    FVM_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    int rc = FVM_OK;
    for (int s = 0; s < nsteps && !rc; s++) rc = enqueue_step(first_phase ^ (s & 1));
    hipError_t e = hipStreamEndCapture(stream_, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
    return FVM_OK;
}

int Engine::capture_graphs() {
    drop_graphs();
    for (int g = 0; g < 2; g++) {
        int rc = capture_sequence(g, 2, &graph_[g]);
        if (rc) return rc;
    }
    graph_valid_ = true;
    return FVM_OK;
}

int Engine::capture_long_graphs() {
    for (int g = 0; g < 2; g++) {
        int rc = capture_sequence(g, LONG_STEPS, &graph_long_[g]);
        if (rc) return rc;
    }
    long_valid_ = true;
    return FVM_OK;
}

int Engine::capture_xl_graphs() {
    for (int g = 0; g < 2; g++) {
        int rc = capture_sequence(g, XL_STEPS, &graph_xl_[g]);
        if (rc) return rc;
    }
    xl_valid_ = true;
    return FVM_OK;
}

void advance_counters(std::vector<HostCase*>& cases, int nsteps) {
    for (HostCase* c : cases)
        if (c && c->active && !c->finished) c->steps_done += nsteps;
}

int Engine::step_direct(int nsteps) {
    if (lists_dirty_) {
        int rc = rebuild_lists();
        if (rc) return rc;
    }
    for (int s = 0; s < nsteps; s++) {
        int rc = enqueue_step(phase_);
        if (rc) return rc;
        phase_ ^= 1;
    }
    advance_counters(cases_, nsteps);
    return FVM_OK;
}

int Engine::step_graph(int nsteps) {
    if (lists_dirty_) {
        int rc = rebuild_lists();
        if (rc) return rc;
    }
    if (timing_) return step_direct(nsteps);
    int done = 0;
    if (nsteps >= 2 && !graph_valid_) {
        int rc = capture_graphs();
        if (rc) return rc;
    }
    if (nsteps >= LONG_STEPS && !long_valid_) {
        int rc = capture_long_graphs();
        if (rc) return rc;
    }
    if (nsteps >= XL_STEPS && !xl_valid_) {
        int rc = capture_xl_graphs();
        if (rc) return rc;
    }
    while (nsteps - done >= XL_STEPS) {
        FVM_HIP(hipGraphLaunch(graph_xl_[phase_], stream_));
        done += XL_STEPS;
    }
    while (nsteps - done >= LONG_STEPS) {
        FVM_HIP(hipGraphLaunch(graph_long_[phase_], stream_));
        done += LONG_STEPS;
    }
    while (nsteps - done >= 2) {
        FVM_HIP(hipGraphLaunch(graph_[phase_], stream_));
        done += 2;
    }
    advance_counters(cases_, done);
    if (done < nsteps) return step_direct(nsteps - done);
    return FVM_OK;
}

int Engine::prepare_graphs() {
    if (lists_dirty_) {
        int rc = rebuild_lists();
        if (rc) return rc;
    }
    if (timing_) return FVM_OK;   // timed runs launch directly
    if (!graph_valid_) {
        int rc = capture_graphs();
        if (rc) return rc;
    }
    if (!long_valid_) {
        int rc = capture_long_graphs();
        if (rc) return rc;
    }
    if (!xl_valid_) {
        int rc = capture_xl_graphs();
        if (rc) return rc;
    }
    return FVM_OK;
}

int Engine::advance(int nsteps, bool use_graph) {
    if (nsteps <= 0) return FVM_OK;
    return use_graph ? step_graph(nsteps) : step_direct(nsteps);
}

int Engine::sync() {
    FVM_HIP(hipStreamSynchronize(stream_));
    if (side_) FVM_HIP(hipStreamSynchronize(side_));
    return flush_timings();
}

int Engine::read_state(int id, void* host, bool as_f64) {
    HostCase* hc = get(id);
    if (!hc) return set_error(FVM_ERR_NOT_FOUND, "no such case");
    // a finished case is read on the side stream after its last step, while
    // the steps queued after it run on; anything else after the queued steps
    const bool frozen = hc->finished && hc->ev_frozen;
    hipStream_t s = frozen ? side_ : stream_;
    if (frozen) FVM_HIP(hipStreamWaitEvent(side_, hc->ev_frozen, 0));
    CaseLayout L = layout_for(hc->p, fp64_, kahan_, hc->nsets);
    int cur = current_set(*hc);
    size_t es = fp64_ ? 8 : 4;
    std::vector<char> tmp;
    for (int f = 0; f < 5; f++) {
        const char* src = (const char*)hc->mem + L.sets_off + L.plane * plane_index(cur, f);
        if ((as_f64 && fp64_) || (!as_f64 && !fp64_)) {
            FVM_HIP(hipMemcpyAsync((char*)host + f * L.n * es, src, L.n * es, hipMemcpyDeviceToHost, s));
        } else {
            tmp.resize(L.n * es);
            FVM_HIP(hipMemcpyAsync(tmp.data(), src, L.n * es, hipMemcpyDeviceToHost, s));
            FVM_HIP(hipStreamSynchronize(s));
            if (as_f64) {
                double* o = (double*)host + f * L.n;
                const float* v = (const float*)tmp.data();
                for (size_t i = 0; i < L.n; i++) o[i] = v[i];
            } else {
                float* o = (float*)host + f * L.n;
                const double* v = (const double*)tmp.data();
                for (size_t i = 0; i < L.n; i++) o[i] = (float)v[i];
            }
        }
    }
    FVM_HIP(hipStreamSynchronize(s));
    return FVM_OK;
}

int Engine::write_state(int id, const void* host) {
    HostCase* hc = get(id);
    if (!hc) return set_error(FVM_ERR_NOT_FOUND, "no such case");
    FVM_HIP(hipStreamSynchronize(stream_));
    CaseLayout L = layout_for(hc->p, fp64_, kahan_, hc->nsets);
    int cur = current_set(*hc);
    std::vector<double> tmp;
    for (int f = 0; f < 5; f++) {
        char* dst = (char*)hc->mem + L.sets_off + L.plane * plane_index(cur, f);
        const float* s = (const float*)host + f * L.n;
        if (fp64_) {
            tmp.resize(L.n);
            for (size_t i = 0; i < L.n; i++) tmp[i] = s[i];
            FVM_HIP(hipMemcpy(dst, tmp.data(), L.n * 8, hipMemcpyHostToDevice));
        } else {
            FVM_HIP(hipMemcpy(dst, s, L.n * 4, hipMemcpyHostToDevice));
        }
    }
    // maxima of the new state into the slot the next step reads
    std::vector<WorkItem> tiles, ordered;
    for (int t = 0; t < hc->tiles_x * hc->tiles_y; t++) tiles.push_back(WorkItem{hc->id, t});
    xcd_order(tiles, ordered);
    int rc = upload_items(ordered, &scr_tiles_, &scr_tiles_cap_, scr_cnt_, 0);
    if (rc) return rc;
    WorkItem* const d_t = scr_tiles_;
    int* const d_cnt = scr_cnt_;
    FVM_HIP(hipMemsetAsync((char*)hc->mem + L.slots_off + phase_ * 8, 0, 8, stream_));
    FVM_HIP(hipMemsetAsync((char*)hc->mem + L.slots_off + kSlots64Off + phase_ * 16, 0, 16, stream_));
    if (kahan_) FVM_HIP(hipMemsetAsync((char*)hc->mem + L.k_off, 0, round256(L.n * 4) * 5, stream_));
    Roles r{cur, cur, -1, -1, -1, phase_, phase_, 0, 0};
    KernelCall kc{};
    kc.op = OP_MAXIMA;
    kc.stream = stream_;
    kc.cases = d_cases_;
    kc.work = d_t;
    kc.count = d_cnt;
    kc.ri = r;
    kc.rv = r;
    kc.grid = dim3((unsigned)std::max<size_t>(1, ordered.size()));
    launch_tiled(fp64_, tx_, ty_, kc);
    FVM_HIP(hipGetLastError());
    FVM_HIP(hipStreamSynchronize(stream_));
    return FVM_OK;
}

int Engine::read_cell_types(int id, uint8_t* host) {
    HostCase* hc = get(id);
    if (!hc) return set_error(FVM_ERR_NOT_FOUND, "no such case");
    CaseLayout L = layout_for(hc->p, fp64_, kahan_, hc->nsets);
    FVM_HIP(hipStreamSynchronize(stream_));
    FVM_HIP(hipMemcpy(host, (char*)hc->mem + L.ct_off, L.n, hipMemcpyDeviceToHost));
    return FVM_OK;
}

int Engine::read_sdf(int id, float* host) {
    HostCase* hc = get(id);
    if (!hc) return set_error(FVM_ERR_NOT_FOUND, "no such case");
    CaseLayout L = layout_for(hc->p, fp64_, kahan_, hc->nsets);
    FVM_HIP(hipStreamSynchronize(stream_));
    FVM_HIP(hipMemcpy(host, (char*)hc->mem + L.sdf_off, L.n * 4, hipMemcpyDeviceToHost));
    return FVM_OK;
}

int Engine::read_time(int id, double* t, long long* steps) {
    HostCase* hc = get(id);
    if (!hc) return set_error(FVM_ERR_NOT_FOUND, "no such case");
    CaseLayout L = layout_for(hc->p, fp64_, kahan_, hc->nsets);
    // a finished case: after its last step, on the side stream (as read_state)
    const bool frozen = hc->finished && hc->ev_frozen;
    hipStream_t s = frozen ? side_ : stream_;
    if (frozen) FVM_HIP(hipStreamWaitEvent(side_, hc->ev_frozen, 0));
    if (t) FVM_HIP(hipMemcpyAsync(t, (char*)hc->mem + L.t_off, 8, hipMemcpyDeviceToHost, s));
    if (steps) FVM_HIP(hipMemcpyAsync(steps, (char*)hc->mem + L.steps_off, 8, hipMemcpyDeviceToHost, s));
    FVM_HIP(hipStreamSynchronize(s));
    return FVM_OK;
}

int Engine::queue_maxima(int id) {
    HostCase* hc = get(id);
    if (!hc) return set_error(FVM_ERR_NOT_FOUND, "no such case");
    CaseLayout L = layout_for(hc->p, fp64_, kahan_, hc->nsets);
    if (dt64())
        FVM_HIP(hipMemcpyAsync(h_pinned_, (char*)hc->mem + L.slots_off + kSlots64Off + phase_ * 16, 16,
                               hipMemcpyDeviceToHost, stream_));
    else
        FVM_HIP(hipMemcpyAsync(h_pinned_, (char*)hc->mem + L.slots_off + phase_ * 8, 8, hipMemcpyDeviceToHost, stream_));
    return FVM_OK;
}

int Engine::read_maxima(int id, double* max_speed, double* max_nu) {
    int rc = queue_maxima(id);
    if (rc) return rc;
    FVM_HIP(hipStreamSynchronize(stream_));
    double a, b;
    if (dt64()) {
        std::memcpy(&a, h_pinned_, 8);
        std::memcpy(&b, (char*)h_pinned_ + 8, 8);
    } else {
        float fa, fb;
        std::memcpy(&fa, h_pinned_, 4);
        std::memcpy(&fb, (char*)h_pinned_ + 4, 4);
        a = fa;
        b = fb;
    }
    if (max_speed) *max_speed = a;
    if (max_nu) *max_nu = b;
    return FVM_OK;
}

int Engine::field(int id, int fieldno, float* dev_out) {
    HostCase* hc = get(id);
    if (!hc) return set_error(FVM_ERR_NOT_FOUND, "no such case");
    if (fieldno < 0 || fieldno > 5) return set_error(FVM_ERR_INVALID_ARGUMENT, "unknown field");
    // The reference writes these buffers on the legacy default stream, so its
    // callers order their own default-stream work (graphics interop, copies)
    // around the call without synchronising.  The solver's stream does not
    // synchronise with the null stream, so wait here for the caller's earlier
    // null-stream work and return only once the buffer is written.
    FVM_HIP(hipEventRecord(ev_null_, nullptr));
    FVM_HIP(hipStreamWaitEvent(stream_, ev_null_, 0));
    launch_field(fp64_, stream_, d_cases_, id, (size_t)hc->p.nx * hc->p.ny, current_set(*hc), fieldno, dev_out);
    FVM_HIP(hipGetLastError());
    FVM_HIP(hipStreamSynchronize(stream_));
    return FVM_OK;
}

int Engine::stat(int id, int mode, float* out) {
    HostCase* hc = get(id);
    if (!hc) return set_error(FVM_ERR_NOT_FOUND, "no such case");
    launch_zero_u32(stream_, d_stat_, 1);
    launch_stat(fp64_, stream_, d_cases_, id, (size_t)hc->p.nx * hc->p.ny, current_set(*hc), mode, d_stat_);
    FVM_HIP(hipGetLastError());
    unsigned bits = 0;
    FVM_HIP(hipMemcpyAsync(h_pinned_ + 4, d_stat_, 4, hipMemcpyDeviceToHost, stream_));
    FVM_HIP(hipStreamSynchronize(stream_));
    std::memcpy(&bits, h_pinned_ + 4, 4);
    std::memcpy(out, &bits, 4);
    return FVM_OK;
}

int Engine::arrows(int id, float* dev_out, int max_vertices, int step, float u_inf, float max_len, float head_angle,
                   float head_ratio, int* num_vertices) {
    HostCase* hc = get(id);
    if (!hc) return set_error(FVM_ERR_NOT_FOUND, "no such case");
    if (!dev_out || step < 1 || max_vertices < 0 || !(u_inf > 0.0f))
        return set_error(FVM_ERR_INVALID_ARGUMENT, "arrows: need a buffer, step >= 1, max_vertices >= 0 and u_inf > 0");
    const int nax = (hc->p.nx + step - 1) / step, nay = (hc->p.ny + step - 1) / step;
    if ((long long)nax * nay * 8 > (long long)INT32_MAX)
        return set_error(FVM_ERR_INVALID_ARGUMENT, "arrows: more than 2^31 vertices at this step; use a larger step");
    // like field(): ordered after the caller's null-stream work (graphics interop)
    FVM_HIP(hipEventRecord(ev_null_, nullptr));
    FVM_HIP(hipStreamWaitEvent(stream_, ev_null_, 0));
    unsigned* counter = d_stat_ + 1;
    launch_zero_u32(stream_, counter, 1);
    launch_arrows(fp64_, stream_, d_cases_, id, current_set(*hc), step, nax, nay, u_inf, max_len, std::cos(head_angle),
                  std::sin(head_angle), head_ratio, max_vertices, dev_out, counter);
    FVM_HIP(hipGetLastError());
    FVM_HIP(hipMemcpyAsync(h_pinned_ + 8, counter, 4, hipMemcpyDeviceToHost, stream_));
    FVM_HIP(hipStreamSynchronize(stream_));
    unsigned n = 0;
    std::memcpy(&n, h_pinned_ + 8, 4);
    const unsigned cap = (unsigned)(max_vertices / 8) * 8u;
    *num_vertices = (int)std::min(n, cap);
    return FVM_OK;
}

int Engine::conservation(int id, double out[4]) {
    HostCase* hc = get(id);
    if (!hc) return set_error(FVM_ERR_NOT_FOUND, "no such case");
    CaseLayout L = layout_for(hc->p, fp64_, kahan_, hc->nsets);
    const int cur = current_set(*hc);
    const void* planes[4];
    for (int f = 0; f < 4; f++) planes[f] = (const char*)hc->mem + L.sets_off + L.plane * plane_index(cur, f);
    launch_cons(fp64_, stream_, planes, L.n, d_cons_ + 8, kConsBlocks, d_cons_);
    FVM_HIP(hipGetLastError());
    FVM_HIP(hipStreamSynchronize(stream_));
    FVM_HIP(hipMemcpy(out, d_cons_, 4 * sizeof(double), hipMemcpyDeviceToHost));
    return FVM_OK;
}

int Engine::forces(int id, double out[3]) {
    HostCase* hc = get(id);
    if (!hc) return set_error(FVM_ERR_NOT_FOUND, "no such case");
    launch_forces(fp64_, stream_, d_cases_, id, current_set(*hc), d_cons_ + 8, kConsBlocks, d_cons_);
    FVM_HIP(hipGetLastError());
    FVM_HIP(hipStreamSynchronize(stream_));
    double r[4];
    FVM_HIP(hipMemcpy(r, d_cons_, 4 * sizeof(double), hipMemcpyDeviceToHost));
    out[0] = r[0]; out[1] = r[1]; out[2] = r[2];
    return FVM_OK;
}

// ------------------------------------------------------------- snapshots
namespace {
struct File {
    FILE* f = nullptr;
    explicit File(FILE* p) : f(p) {}
    ~File() { if (f) std::fclose(f); }
};
}  // namespace

static_assert(sizeof(fvm_snapshot_header) == 256, "snapshot header is 256 bytes");

int read_snapshot_header(const char* path, fvm_snapshot_header* h, fvm_params* p) {
    if (!path) return set_error(FVM_ERR_INVALID_ARGUMENT, "null path");
    File in(std::fopen(path, "rb"));
    if (!in.f) return set_error(FVM_ERR_INVALID_ARGUMENT, std::string("cannot open snapshot ") + path);
    if (std::fread(h, sizeof(*h), 1, in.f) != 1) return set_error(FVM_ERR_INVALID_ARGUMENT, "snapshot header truncated");
    if (std::memcmp(h->magic, FVM_SNAPSHOT_MAGIC, 8) != 0 || h->version != FVM_SNAPSHOT_VERSION ||
        h->header_bytes != sizeof(fvm_snapshot_header))
        return set_error(FVM_ERR_INVALID_ARGUMENT, "not an fvm snapshot of this version");
    if (h->abi_version != FVM_ABI_VERSION || h->params_bytes != sizeof(fvm_params))
        return set_error(FVM_ERR_INVALID_ARGUMENT, "snapshot written by another ABI version");
    if (h->nx <= 0 || h->ny <= 0 || (h->elem_bytes != 4 && h->elem_bytes != 8))
        return set_error(FVM_ERR_INVALID_ARGUMENT, "snapshot header has bad sizes");
    if (p && std::fread(p, sizeof(*p), 1, in.f) != 1) return set_error(FVM_ERR_INVALID_ARGUMENT, "snapshot params truncated");
    if (p && (p->nx != h->nx || p->ny != h->ny)) return set_error(FVM_ERR_INVALID_ARGUMENT, "snapshot sizes disagree");
    return FVM_OK;
}

int Engine::save_case(int id, const char* path) {
    HostCase* hc = get(id);
    if (!hc) return set_error(FVM_ERR_NOT_FOUND, "no such case");
    if (!path) return set_error(FVM_ERR_INVALID_ARGUMENT, "null path");
    FVM_HIP(hipStreamSynchronize(stream_));
    CaseLayout L = layout_for(hc->p, fp64_, kahan_, hc->nsets);
    char* base = (char*)hc->mem;
    fvm_snapshot_header h;
    std::memset(&h, 0, sizeof(h));
    std::memcpy(h.magic, FVM_SNAPSHOT_MAGIC, 8);
    h.version = FVM_SNAPSHOT_VERSION;
    h.header_bytes = sizeof(h);
    h.abi_version = FVM_ABI_VERSION;
    h.params_bytes = sizeof(fvm_params);
    h.nx = hc->p.nx;
    h.ny = hc->p.ny;
    h.elem_bytes = fp64_ ? 8 : 4;
    h.kahan = kahan_ ? 1 : 0;
    long long steps = 0;
    FVM_HIP(hipMemcpy(&h.t, base + L.t_off, 8, hipMemcpyDeviceToHost));
    FVM_HIP(hipMemcpy(&steps, base + L.steps_off, 8, hipMemcpyDeviceToHost));
    h.steps = steps;
    unsigned bits[2];
    FVM_HIP(hipMemcpy(bits, base + L.slots_off + phase_ * 8, 8, hipMemcpyDeviceToHost));
    h.max_speed_bits = bits[0];
    h.max_nu_bits = bits[1];
    double m64[2];
    FVM_HIP(hipMemcpy(m64, base + L.slots_off + kSlots64Off + phase_ * 16, 16, hipMemcpyDeviceToHost));
    h.max_speed64 = m64[0];
    h.max_nu64 = m64[1];
    File out(std::fopen(path, "wb"));
    if (!out.f) return set_error(FVM_ERR_INVALID_ARGUMENT, std::string("cannot create snapshot ") + path);
    bool ok = std::fwrite(&h, sizeof(h), 1, out.f) == 1 && std::fwrite(&hc->p, sizeof(fvm_params), 1, out.f) == 1;
    const int cur = current_set(*hc);
    std::vector<char> buf(L.n * h.elem_bytes);
    for (int f = 0; ok && f < 5; f++) {
        FVM_HIP(hipMemcpy(buf.data(), base + L.sets_off + L.plane * plane_index(cur, f), buf.size(), hipMemcpyDeviceToHost));
        ok = std::fwrite(buf.data(), 1, buf.size(), out.f) == buf.size();
    }
    for (int f = 0; ok && kahan_ && f < 5; f++) {
        FVM_HIP(hipMemcpy(buf.data(), base + L.k_off + round256(L.n * 4) * f, L.n * 4, hipMemcpyDeviceToHost));
        ok = std::fwrite(buf.data(), 1, L.n * 4, out.f) == L.n * 4;
    }
    if (!ok || std::fflush(out.f) != 0) return set_error(FVM_ERR_INVALID_ARGUMENT, std::string("write failed: ") + path);
    return FVM_OK;
}

int Engine::load_case(int id, const char* path) {
    HostCase* hc = get(id);
    if (!hc) return set_error(FVM_ERR_NOT_FOUND, "no such case");
    fvm_snapshot_header h;
    fvm_params p;
    int rc = read_snapshot_header(path, &h, &p);
    if (rc) return rc;
    if (h.nx != hc->p.nx || h.ny != hc->p.ny || h.elem_bytes != (fp64_ ? 8u : 4u) || (h.kahan != 0) != kahan_)
        return set_error(FVM_ERR_INVALID_ARGUMENT, "snapshot does not match the case (size, precision or carries)");
    File in(std::fopen(path, "rb"));
    if (!in.f || std::fseek(in.f, (long)(sizeof(h) + sizeof(fvm_params)), SEEK_SET) != 0)
        return set_error(FVM_ERR_INVALID_ARGUMENT, std::string("cannot read snapshot ") + path);
    FVM_HIP(hipStreamSynchronize(stream_));
    CaseLayout L = layout_for(hc->p, fp64_, kahan_, hc->nsets);
    char* base = (char*)hc->mem;
    const int cur = current_set(*hc);
    std::vector<char> buf(L.n * h.elem_bytes);
    for (int f = 0; f < 5; f++) {
        if (std::fread(buf.data(), 1, buf.size(), in.f) != buf.size())
            return set_error(FVM_ERR_INVALID_ARGUMENT, "snapshot state truncated");
        FVM_HIP(hipMemcpy(base + L.sets_off + L.plane * plane_index(cur, f), buf.data(), buf.size(), hipMemcpyHostToDevice));
    }
    for (int f = 0; kahan_ && f < 5; f++) {
        if (std::fread(buf.data(), 1, L.n * 4, in.f) != L.n * 4)
            return set_error(FVM_ERR_INVALID_ARGUMENT, "snapshot carries truncated");
        FVM_HIP(hipMemcpy(base + L.k_off + round256(L.n * 4) * f, buf.data(), L.n * 4, hipMemcpyHostToDevice));
    }
    long long steps = h.steps;
    unsigned bits[2] = {h.max_speed_bits, h.max_nu_bits};
    double m64[2] = {h.max_speed64, h.max_nu64};
    FVM_HIP(hipMemcpy(base + L.slots_off + kSlots64Off + phase_ * 16, m64, 16, hipMemcpyHostToDevice));
    FVM_HIP(hipMemcpy(base + L.t_off, &h.t, 8, hipMemcpyHostToDevice));
    FVM_HIP(hipMemcpy(base + L.steps_off, &steps, 8, hipMemcpyHostToDevice));
    FVM_HIP(hipMemcpy(base + L.slots_off + phase_ * 8, bits, 8, hipMemcpyHostToDevice));
    return FVM_OK;
}

}  // namespace fvm
