// fvm_kernels.hip -- hand-written gfx950 kernels of the FVM time step.
//
// Kernels of one SSP-RK2 step (src/solver.cu:3289-3412), all over a work list
// of (case, tile) items so that one launch serves one case or hundreds:
//   k_diffuse<final=false>   viscous half step D(dt/2) on U^n  -> U_d        (viscous only)
//   k_stage<STAGE1>          prims + MUSCL + Riemann flux + FV update + point-wise BC -> U*
//   k_stage<STAGE2>          immersed-boundary ghost fill of U* as the tile is staged,
//                            second stage + SSP blend (+ point-wise BC + per-tile CFL
//                            maxima when inviscid)
//   k_diffuse<final=true>    second half step + point-wise BC + per-tile CFL maxima (viscous only)
//   k_ghost<final>           ghost fill of the new state, CFL maxima of the ghost cells and
//                            the reduction of every case's per-tile maxima into its dt slot
// Each stage kernel stages its tile plus a 2-cell halo through LDS, derives
// the primitives once per cell, computes every face flux once per tile and
// keeps flux divergence in registers; no face array ever reaches HBM.
#include "fvm_device.hpp"
#include "fvm_kernels.hpp"

#include <hip/hip_ext.h>

#include <algorithm>
#include <type_traits>

namespace fvm {

constexpr int NT = 256;

// ---------------------------------------------------------------------------
// Obstacle marking (computeSDFKernel, src/solver.cu:536-589).  IEEE division
// and square root, no contraction: the mask must equal the CPU oracle's bit for
// bit.  Rotation and star vertices arrive precomputed from the host.
// ---------------------------------------------------------------------------
#pragma clang fp contract(off)
__device__ float dist_to_segment(float px, float py, float ax, float ay, float bx, float by) {
    float abx = bx - ax, aby = by - ay;
    float apx = px - ax, apy = py - ay;
    float t = __fdiv_rn(apx * abx + apy * aby, abx * abx + aby * aby + 1e-10f);
    t = fmaxf(0.0f, fminf(1.0f, t));
    float cx = ax + t * abx, cy = ay + t * aby;
    float dx = px - cx, dy = py - cy;
    return sqrtf(dx * dx + dy * dy);
}

__device__ __forceinline__ float cross2d(float ax, float ay, float bx, float by) { return ax * by - ay * bx; }

__device__ float sdf_quad(float px, float py, float v0x, float v0y, float v1x, float v1y,
                          float v2x, float v2y, float v3x, float v3y) {
    float d0 = dist_to_segment(px, py, v0x, v0y, v1x, v1y);
    float d1 = dist_to_segment(px, py, v1x, v1y, v2x, v2y);
    float d2 = dist_to_segment(px, py, v2x, v2y, v3x, v3y);
    float d3 = dist_to_segment(px, py, v3x, v3y, v0x, v0y);
    float m = fminf(fminf(d0, d1), fminf(d2, d3));
    float c0 = cross2d(v1x - v0x, v1y - v0y, px - v0x, py - v0y);
    float c1 = cross2d(v2x - v1x, v2y - v1y, px - v1x, py - v1y);
    float c2 = cross2d(v3x - v2x, v3y - v2y, px - v2x, py - v2y);
    float c3 = cross2d(v0x - v3x, v0y - v3y, px - v3x, py - v3y);
    bool inside = (c0 >= 0 && c1 >= 0 && c2 >= 0 && c3 >= 0) || (c0 <= 0 && c1 <= 0 && c2 <= 0 && c3 <= 0);
    return inside ? -m : m;
}

__device__ float shape_sdf(const Geom& k, float px, float py) {
    const float cx = k.obs_x, cy = k.obs_y, r = k.obs_r;
    float lx = px - cx, ly = py - cy;
    float qx = lx * k.cosR - ly * k.sinR;
    float qy = lx * k.sinR + ly * k.cosR;
    switch (k.shape) {
    default:
    case 0:  // circle
        return sqrtf((px - cx) * (px - cx) + (py - cy) * (py - cy)) - r;
    case 1: {  // five-point star: distance to the 10 edges, winding-number sign
        float minDist = 1e10f, winding = 0.0f;
        for (int i = 0; i < 10; i++) {
            int j = (i + 1) % 10;
            float ax = k.star[i * 2], ay = k.star[i * 2 + 1];
            float bx = k.star[j * 2], by = k.star[j * 2 + 1];
            minDist = fminf(minDist, dist_to_segment(qx, qy, ax, ay, bx, by));
            float eax = ax - qx, eay = ay - qy, ebx = bx - qx, eby = by - qy;
            winding += atan2f(cross2d(eax, eay, ebx, eby), eax * ebx + eay * eby);
        }
        float sign = (fabsf(winding) > 3.14159265358979323846f) ? -1.0f : 1.0f;
        return sign * minDist;
    }
    case 2: {  // diamond
        float ndx = __fdiv_rn(fabsf(qx), r), ndy = __fdiv_rn(fabsf(qy), r);
        return (ndx + ndy - 1.0f) * r * 0.7071f;
    }
    case 3: {  // capsule
        float hw = r * 1.5f, cr = r * 0.5f;
        float clampedX = fmaxf(-hw + cr, fminf(hw - cr, qx));
        float dx = qx - clampedX, dy = qy;
        return sqrtf(dx * dx + dy * dy) - cr;
    }
    case 4: {  // triangle pointing along +x
        const float s32 = 0.866025404f;
        float v0x = r, v0y = 0.0f, v1x = -r * 0.5f, v1y = r * s32, v2x = -r * 0.5f, v2y = -r * s32;
        float d0 = dist_to_segment(qx, qy, v0x, v0y, v1x, v1y);
        float d1 = dist_to_segment(qx, qy, v1x, v1y, v2x, v2y);
        float d2 = dist_to_segment(qx, qy, v2x, v2y, v0x, v0y);
        float m = fminf(d0, fminf(d1, d2));
        float c0 = cross2d(v1x - v0x, v1y - v0y, qx - v0x, qy - v0y);
        float c1 = cross2d(v2x - v1x, v2y - v1y, qx - v1x, qy - v1y);
        float c2 = cross2d(v0x - v2x, v0y - v2y, qx - v2x, qy - v2y);
        bool inside = (c0 >= 0 && c1 >= 0 && c2 >= 0) || (c0 <= 0 && c1 <= 0 && c2 <= 0);
        return inside ? -m : m;
    }
    case 5: {  // starship: circle body, fixed wing root, rotated wing (common.h:28-36)
        float distToCircle = sqrtf(qx * qx + qy * qy) - r;
        float scale = __fdiv_rn(r, 90.0f);
        float refX = __fdiv_rn(qx, scale);
        float refY = __fdiv_rn(fabsf(qy), scale);
        float dRoot = sdf_quad(refX, refY, -45.0f, 77.9422f, 10.9682f, 89.329f, 6.0f, 112.0f, -6.0f, 112.0f) * scale;
        float pX = -6.0f, pY = 112.0f;
        float dx1 = 6.0f - pX, dy1 = 112.0f - pY;
        float w1x = pX + dx1 * k.cosW - dy1 * k.sinW, w1y = pY + dx1 * k.sinW + dy1 * k.cosW;
        float dx2 = 3.0f - pX, dy2 = 186.0f - pY;
        float w2x = pX + dx2 * k.cosW - dy2 * k.sinW, w2y = pY + dx2 * k.sinW + dy2 * k.cosW;
        float dx3 = -3.0f - pX, dy3 = 186.0f - pY;
        float w3x = pX + dx3 * k.cosW - dy3 * k.sinW, w3y = pY + dx3 * k.sinW + dy3 * k.cosW;
        float dWing = sdf_quad(refX, refY, pX, pY, w1x, w1y, w2x, w2y, w3x, w3y) * scale;
        return fminf(distToCircle, fminf(dRoot, dWing));
    }
    }
}

__device__ __forceinline__ uint8_t classify(float d, int i, int nx, float band) {
    uint8_t t;
    if (d < -band) t = CT_SOLID;
    else if (d < band && d >= -band) t = CT_GHOST;
    else t = CT_FLUID;
    if (i == 0) t = CT_INFLOW;
    else if (i == nx - 1) t = CT_OUTFLOW;
    return t;
}

// Marks one case; with fixup != 0 it is updateSDFWithFixupKernel
// (solver.cu:607-664): cells that turn from solid/ghost into fluid are zeroed
// in buffer set `fix_set`.
template <class T>
__global__ __launch_bounds__(NT) void k_mark(const CaseDesc<T>* __restrict__ cases, int case_id, int fixup, int fix_set) {
    const CaseDesc<T>& cd = cases[case_id];
    const CaseConsts<T>& k = cd.c;
    size_t n = (size_t)k.nx * k.ny;
    for (size_t idx = (size_t)blockIdx.x * NT + threadIdx.x; idx < n; idx += (size_t)gridDim.x * NT) {
        int i = (int)(idx % k.nx), j = (int)(idx / k.nx);
        float x = ((float)i + 0.5f) * k.geo.dx;
        float y = ((float)j + 0.5f) * k.geo.dy;
        float d = shape_sdf(k.geo, x, y);
        uint8_t old = cd.b.ct[idx];
        uint8_t t = classify(d, i, k.nx, k.geo.band);
        cd.b.sdf[idx] = d;
        cd.b.ct[idx] = t;
        if (fixup && (old == CT_SOLID || old == CT_GHOST) && t == CT_FLUID) {
            for (int f = 0; f < 5; f++) ((T*)cd.b.U[fix_set][f])[idx] = T(0);
        }
    }
}

// Seeded counter-based perturbation: splitmix64 finaliser of
// (seed * golden + idx + 1); top 24 bits -> u in [0,1); eps = amp (2u - 1).
__device__ __forceinline__ float noise_unit_f(uint32_t seed, unsigned long long idx) {
    unsigned long long z = (unsigned long long)seed * 0x9E3779B97F4A7C15ull + idx + 1ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (float)(z >> 40) * (1.0f / 16777216.0f);
}

// initializeKernel (src/solver.cu:911-940) + perturbation; writes sets a and b.
template <class T>
__global__ __launch_bounds__(NT) void k_init(const CaseDesc<T>* __restrict__ cases, int case_id, int set_a, int set_b) {
    const CaseDesc<T>& cd = cases[case_id];
    const CaseConsts<T>& k = cd.c;
    size_t n = (size_t)k.nx * k.ny;
    for (size_t idx = (size_t)blockIdx.x * NT + threadIdx.x; idx < n; idx += (size_t)gridDim.x * NT) {
        T rho = (T)k.rho_inf, u = (T)k.u_inf, v = (T)k.v_inf;
        if (k.noise_field != 0) {
            T eps = (T)k.noise_amp * (T)(2.0f * noise_unit_f(k.noise_seed, idx) - 1.0f);
            if (k.noise_field == 1) rho = (T)k.rho_inf * (T(1) + eps);
            else v = (T)k.v_inf + (T)k.u_inf * eps;
        }
        T ke = T(0.5) * rho * (u * u + v * v);
        T e = (T)k.p_inf / (T)k.gm1;
        T vals[5] = {rho, rho * u, rho * v, e + ke, e};
        for (int f = 0; f < 5; f++) {
            ((T*)cd.b.U[set_a][f])[idx] = vals[f];
            if (set_b >= 0) ((T*)cd.b.U[set_b][f])[idx] = vals[f];
        }
    }
}
#pragma clang fp contract(on)

// Copy buffer set `src` to `dst` for one case.
template <class T>
__global__ __launch_bounds__(NT) void k_copy_set(const CaseDesc<T>* __restrict__ cases, int case_id, int src, int dst) {
    const CaseDesc<T>& cd = cases[case_id];
    size_t n = (size_t)cd.c.nx * cd.c.ny;
    for (size_t idx = (size_t)blockIdx.x * NT + threadIdx.x; idx < n; idx += (size_t)gridDim.x * NT)
        for (int f = 0; f < 5; f++) ((T*)cd.b.U[dst][f])[idx] = ((const T*)cd.b.U[src][f])[idx];
}

// Point-wise BC in place on buffer set r.out over a tile list (reset path).
template <class T, int TX, int TY>
__global__ __launch_bounds__(NT) void k_bc_inplace(const CaseDesc<T>* __restrict__ cases, const WorkItem* __restrict__ work,
                                                   const int* __restrict__ count, Roles r) {
    const int n_items = *count;
    for (int w = blockIdx.x; w < n_items; w += gridDim.x) {
        WorkItem it = work[w];
        if (it.index < 0) continue;
        const CaseDesc<T>& cd = cases[it.case_id];
        const CaseConsts<T>& k = cd.c;
        const int i0 = (it.index % cd.tiles_x) * TX, j0 = (it.index / cd.tiles_x) * TY;
        for (int q = threadIdx.x; q < TX * TY; q += NT) {
            int gi = i0 + q % TX, gj = j0 + q / TX;
            if (gi >= k.nx || gj >= k.ny) continue;
            size_t g = (size_t)gj * k.nx + gi;
            uint8_t c = cd.b.ct[g];
            if (c == CT_GHOST) continue;
            T v[5];
            for (int f = 0; f < 5; f++) v[f] = ((T*)cd.b.U[r.out][f])[g];
            if (bc_pointwise(k, gi, gj, c, v))
                for (int f = 0; f < 5; f++) ((T*)cd.b.U[r.out][f])[g] = v[f];
        }
    }
}

// ---------------------------------------------------------------------------
// Immersed-boundary ghost value (applyBoundaryConditionsKernel ghost branch,
// src/solver.cu:2045-2225).  A neighbour that is itself a ghost cell is read
// at its pre-BC value (set pre0, or the SSP blend of pre0 and pre1); any other
// neighbour at its post-BC value in set src.  Normal and ranked neighbours come
// from the ghost table (static per geometry).  IEEE arithmetic, no contraction,
// as in the oracle.
// ---------------------------------------------------------------------------
#pragma clang fp contract(off)
// Offsets of neighbour codes 0..7 (the 8-neighbourhood in the scan order of
// solver.cu:2072), packed two bits per code so that a per-lane code is decoded
// in registers rather than by a table load.
constexpr int NB_DX[8] = {-1, 1, 0, 0, -1, 1, -1, 1};
constexpr int NB_DY[8] = {0, 0, -1, 1, -1, -1, 1, 1};
constexpr unsigned pack_nb(const int (&d)[8]) {
    unsigned v = 0;
    for (int m = 0; m < 8; m++) v |= (unsigned)(d[m] + 1) << (2 * m);
    return v;
}
constexpr unsigned NB_DX_BITS = pack_nb(NB_DX), NB_DY_BITS = pack_nb(NB_DY);
constexpr bool nb_roundtrip(unsigned bits, const int (&d)[8]) {
    for (int m = 0; m < 8; m++)
        if ((int)((bits >> (2 * m)) & 3u) - 1 != d[m]) return false;
    return true;
}
static_assert(nb_roundtrip(NB_DX_BITS, NB_DX) && nb_roundtrip(NB_DY_BITS, NB_DY), "neighbour offset packing");
__device__ __forceinline__ int nb_dx(int m) { return (int)((NB_DX_BITS >> (2 * m)) & 3u) - 1; }
__device__ __forceinline__ int nb_dy(int m) { return (int)((NB_DY_BITS >> (2 * m)) & 3u) - 1; }

// Pre-BC value of field f at cell n: set pre0, or the mean of sets pre0 and
// pre1 when pre1 >= 0.  Both loads are issued unconditionally (pre0 stands in
// for a missing pre1), so no branch separates them.
template <class T>
__device__ __forceinline__ T pre_val(const CaseBufs& b, int pre0, int pre1, unsigned n, int f) {
    const T a = Plane<T>(b.U[pre0][f])[n];
    const T c = Plane<T>(b.U[pre1 >= 0 ? pre1 : pre0][f])[n];
    return pre1 >= 0 ? T(0.5) * a + T(0.5) * c : a;
}

// Index into the ghost table of cell (i, j), or -1 when it is not a ghost cell.
__device__ __forceinline__ int ghost_index(const CaseBufs& b, int i, int j) {
    const int li = i - b.gbox[0], lj = j - b.gbox[1];
    if (li < 0 || lj < 0 || li >= b.gbox[2] || lj >= b.gbox[3]) return -1;
    return Plane<int>(b.gmap)[(unsigned)(lj * b.gbox[2] + li)];
}

// EAGER (the final ghost launch, where the chain of loads is the kernel's
// whole cost): the first candidate's fields are loaded with its tag, before the
// density check.  Otherwise (the fill folded into stage 2's staging loop, where
// registers are short): the tag and density first, the winner's fields after.
template <class T, bool EAGER>
__device__ void ghost_value(const CaseDesc<T>& cd, int src, int pre0, int pre1, int gidx, T o[5]) {
    const CaseConsts<T>& k = cd.c;
    const CaseBufs& b = cd.b;
    const gmem<GhostInfo<T>>* gi = (const gmem<GhostInfo<T>>*)b.ghosts + gidx;
    const int nx = k.nx;
    const int idx = gi->idx;
    const uint32_t order = gi->order;
    const Plane<uint8_t> ct(b.ct);
    const T nxs = gi->nx, nys = gi->ny;
    const int cnt = (int)(order >> 24);
    // A ghost neighbour is read at its pre-BC value (set pre0, or the mean of
    // sets pre0 and pre1 when pre1 >= 0), any other neighbour from set src.
    // EAGER: a candidate's tag and both versions of its four fields are loaded
    // together and the version picked afterwards, one memory round trip per
    // candidate; the first listed one, which almost always passes the density
    // check, is loaded before the check (the cell itself stands in when none is
    // listed).
    const int s0 = pre0 >= 0 ? pre0 : src, s1 = pre1 >= 0 ? pre1 : s0;
    const bool mean = pre0 >= 0 && pre1 >= 0;
    auto candidate = [&](unsigned nidx, T cand[4]) {
        const bool g = ct[nidx] == CT_GHOST;
#pragma unroll
        for (int f = 0; f < 4; f++) {
            const T cur = Plane<T>(b.U[src][f])[nidx];
            const T a0 = Plane<T>(b.U[s0][f])[nidx];
            const T a1 = Plane<T>(b.U[s1][f])[nidx];
            const T pre = mean ? T(0.5) * a0 + T(0.5) * a1 : a0;
            cand[f] = g ? pre : cur;
        }
    };
    auto nb = [&](int q) {
        const int m = (order >> (3 * q)) & 7;
        return (unsigned)(idx + nb_dy(m) * nx + nb_dx(m));
    };
    T fv[4];
    bool found = false;
    if constexpr (EAGER) {
        candidate(cnt > 0 ? nb(0) : (unsigned)idx, fv);
        found = cnt > 0 && fv[0] >= k.min_rho;
        for (int q = 1; q < cnt && !found; q++) {
            T cand[4];
            candidate(nb(q), cand);
            if (cand[0] >= k.min_rho) {
#pragma unroll
                for (int f = 0; f < 4; f++) fv[f] = cand[f];
                found = true;
            }
        }
    } else {
        unsigned best = 0;
        bool best_ghost = false;
        for (int q = 0; q < cnt; q++) {
            const unsigned nidx = nb(q);
            const bool g = ct[nidx] == CT_GHOST;
            const T rn = g ? pre_val<T>(b, s0, mean ? pre1 : -1, nidx, 0) : Plane<T>(b.U[src][0])[nidx];
            if (rn >= k.min_rho) {
                best = nidx;
                best_ghost = g;
                found = true;
                break;
            }
        }
        if (found)
            for (int f = 0; f < 4; f++)
                fv[f] = best_ghost ? pre_val<T>(b, s0, mean ? pre1 : -1, best, f) : Plane<T>(b.U[src][f])[best];
    }
    if (!found) {
        o[0] = k.rho_inf; o[1] = T(0); o[2] = T(0); o[3] = k.e_inf; o[4] = k.e_inf;
        return;
    }
    T rho_f = fv[0];
    T u_f = fv[1] / rho_f, v_f = fv[2] / rho_f;
    T ke_f = T(0.5) * rho_f * (u_f * u_f + v_f * v_f);
    T p_f = k.gm1 * (fv[3] - ke_f);
    rho_f = tmax(rho_f, k.min_rho);
    p_f = tmax(p_f, k.min_p);
    T vn = u_f * nxs + v_f * nys;
    T vt = -v_f * nxs + u_f * nys;
    T T_f = p_f / (rho_f * k.R);
    T_f = tmin(T_f, k.max_T);
    T_f = tmax(T_f, k.min_T);
    p_f = rho_f * k.R * T_f;
    T ug, vg;
    if (k.visc) {
        ug = -u_f; vg = -v_f;
    } else {
        T slip = T(0);
        if (k.wall_mode == 1) slip = T(1);
        else if (vn > T(0)) slip = tmin(T(0.8f), vn / (vn + k.slip_ref));
        T us = -vn * nxs + vt * nys;
        T vs = -vn * nys - vt * nxs;
        ug = (T(1) - slip) * (-u_f) + slip * us;
        vg = (T(1) - slip) * (-v_f) + slip * vs;
    }
    T vel = ::sqrt(ug * ug + vg * vg);
    if (vel > k.vghost) {
        T s = k.vghost / vel;
        ug *= s; vg *= s;
    }
    T Tg;
    if (k.adiabatic) Tg = T_f;
    else {
        Tg = T(2) * k.T_wall - T_f;
        Tg = tmax(Tg, k.min_T);
        Tg = tmin(Tg, k.max_T);
    }
    T rg = tmax(p_f / (k.R * Tg), k.min_rho);
    T eg = p_f / k.gm1;
    T keg = T(0.5) * rg * (ug * ug + vg * vg);
    o[0] = rg; o[1] = rg * ug; o[2] = rg * vg; o[3] = eg + keg; o[4] = eg;
}
#pragma clang fp contract(on)

// ---------------------------------------------------------------------------
// Viscous half step pieces (fusedViscousDiffusionKernel, src/solver.cu:2251-2511),
// shared by the half-step kernel and the stage-1 kernel that fuses the first
// half step in.
// ---------------------------------------------------------------------------
// Stress and heat flux of one cell of the ring around the updated cells, into
// entry n of the five flux planes: Sutherland mu, k = mu cp / Pr (2317-2352),
// central gradients of the staged u, v, T (row stride W, cell s), one-sided at
// the domain edges (2360-2388); zero outside the domain and in solid cells.
template <class T, int W>
__device__ __forceinline__ void viscous_flux(const CaseConsts<T>& k, const ViscConsts<T>& vc, const T* s_u, const T* s_v,
                                             const T* s_T, const uint8_t* s_ct, int s, int gi, int gj, int n, T* s_txx,
                                             T* s_tyy, T* s_txy, T* s_qx, T* s_qy) {
    const int nx = k.nx, ny = k.ny;
    if (gi < 0 || gi >= nx || gj < 0 || gj >= ny || s_ct[s] == CT_SOLID) {
        s_txx[n] = T(0); s_tyy[n] = T(0); s_txy[n] = T(0); s_qx[n] = T(0); s_qy[n] = T(0);
        return;
    }
    const T i2x = (T)k.inv_2dx, i2y = (T)k.inv_2dy, ix = (T)k.inv_dx, iy = (T)k.inv_dy;
    T Tv = s_T[s];
    T mu = sutherland(k, Tv, vc.i_tref);
    T kk = mu * (T)k.cp * vc.i_pr;
    T dudx = (s_u[s + 1] - s_u[s - 1]) * i2x, dvdx = (s_v[s + 1] - s_v[s - 1]) * i2x;
    T dudy = (s_u[s + W] - s_u[s - W]) * i2y, dvdy = (s_v[s + W] - s_v[s - W]) * i2y;
    T dTdx = (s_T[s + 1] - s_T[s - 1]) * i2x, dTdy = (s_T[s + W] - s_T[s - W]) * i2y;
    if (gi == 0) {
        dudx = (s_u[s + 1] - s_u[s]) * ix; dvdx = (s_v[s + 1] - s_v[s]) * ix; dTdx = (s_T[s + 1] - s_T[s]) * ix;
    } else if (gi == nx - 1) {
        dudx = (s_u[s] - s_u[s - 1]) * ix; dvdx = (s_v[s] - s_v[s - 1]) * ix; dTdx = (s_T[s] - s_T[s - 1]) * ix;
    }
    if (gj == 0) {
        dudy = (s_u[s + W] - s_u[s]) * iy; dvdy = (s_v[s + W] - s_v[s]) * iy; dTdy = (s_T[s + W] - s_T[s]) * iy;
    } else if (gj == ny - 1) {
        dudy = (s_u[s] - s_u[s - W]) * iy; dvdy = (s_v[s] - s_v[s - W]) * iy; dTdy = (s_T[s] - s_T[s - W]) * iy;
    }
    T divv = dudx + dvdy;
    const T two3 = T(2) / T(3);
    s_txx[n] = mu * (T(2) * dudx - two3 * divv);
    s_tyy[n] = mu * (T(2) * dvdy - two3 * divv);
    s_txy[n] = mu * (dudy + dvdx);
    s_qx[n] = -kk * dTdx;
    s_qy[n] = -kk * dTdy;
}

// Half-step update of one interior fluid cell (2420-2505): momentum by div tau,
// E by -div q + div(tau v), rho e by -div q + Phi, from the ring's planes (row
// stride Q, entry t) and the staged velocities (row stride W, cell s).
template <class T, int W, int Q>
__device__ __forceinline__ void viscous_update(const CaseConsts<T>& k, const T* s_u, const T* s_v, int s,
                                               const T* s_txx, const T* s_tyy, const T* s_txy, const T* s_qx,
                                               const T* s_qy, int t, T hdt, T v[5]) {
    const T i2x = (T)k.inv_2dx, i2y = (T)k.inv_2dy;
    T dtxx = (s_txx[t + 1] - s_txx[t - 1]) * i2x;
    T dtxy_y = (s_txy[t + Q] - s_txy[t - Q]) * i2y;
    T dtxy_x = (s_txy[t + 1] - s_txy[t - 1]) * i2x;
    T dtyy = (s_tyy[t + Q] - s_tyy[t - Q]) * i2y;
    T dqx = (s_qx[t + 1] - s_qx[t - 1]) * i2x;
    T dqy = (s_qy[t + Q] - s_qy[t - Q]) * i2y;
    T dudx = (s_u[s + 1] - s_u[s - 1]) * i2x, dudy = (s_u[s + W] - s_u[s - W]) * i2y;
    T dvdx = (s_v[s + 1] - s_v[s - 1]) * i2x, dvdy = (s_v[s + W] - s_v[s - W]) * i2y;
    T Phi = s_txx[t] * dudx + s_tyy[t] * dvdy + s_txy[t] * (dudy + dvdx);
    T wxp = s_u[s + 1] * s_txx[t + 1] + s_v[s + 1] * s_txy[t + 1];
    T wxm = s_u[s - 1] * s_txx[t - 1] + s_v[s - 1] * s_txy[t - 1];
    T wyp = s_u[s + W] * s_txy[t + Q] + s_v[s + W] * s_tyy[t + Q];
    T wym = s_u[s - W] * s_txy[t - Q] + s_v[s - W] * s_tyy[t - Q];
    T rhsE = -dqx - dqy + (wxp - wxm) * i2x + (wyp - wym) * i2y;
    T rhse = -dqx - dqy + Phi;
    v[1] += hdt * (dtxx + dtxy_y);
    v[2] += hdt * (dtxy_x + dtyy);
    v[3] += hdt * rhsE;
    v[4] += hdt * rhse;
}

// ---------------------------------------------------------------------------
// Fused RK stage.  MODE 0: stage 1 (update + point-wise BC).  MODE 1: stage 2
// of an inviscid case (update + blend + point-wise BC + CFL maxima) or of a
// viscous case (update + blend only, no BC: the second half step follows).
// MODE 2: stage 1 with the first viscous half step fused in for viscous cases
// (inviscid cases as MODE 0): the half-step result on the tile and its 2-cell
// halo is recomputed from U^n on a 4-cell halo and staged in place of a
// loaded state; the tile's own cells of it are stored in set r.blend, where
// stage 2 reads its blend input.  The results equal the split launches bit
// for bit (the same cell arithmetic, shared with k_diffuse).
// KAHAN: the final blend is written as U^n + delta with a carried compensation.
// ri / rv: roles for inviscid / viscous cases.
// ---------------------------------------------------------------------------
// Small tiles whose x faces and y faces each fit in half the workgroup: both
// face sets in one pass, the x faces on the first two waves and the y faces on
// the other two (different SIMDs, so the two chains run side by side), into
// separate flux arrays.  Larger tiles: an x pass, then a y pass into the same
// array.
template <int TX, int TY>
constexpr bool split_faces() {
    return (TX + 1) * TY <= NT / 2 && TX * (TY + 1) <= NT / 2;
}

// Length of each flux component's LDS row in k_stage.
template <int TX, int TY>
constexpr int flux_row() {
    constexpr int FX = (TX + 1) * TY, FY = TX * (TY + 1);
    return split_faces<TX, TY>() ? FX + FY : (FX > FY ? FX : FY);
}

// LDS bytes of k_stage: staged (rho, u, v, p) + face fluxes + tags; MODE 2
// overlays the half step's (u, v, T) + tags on the 4-cell halo and its five
// stress / heat-flux planes on the 3-cell ring on the same space.
template <class T, int TX, int TY, int MODE>
constexpr int stage_lds() {
    constexpr int SX = TX + 4, SY = TY + 4;
    constexpr int stage = (4 * SX * SY + 5 * flux_row<TX, TY>()) * (int)sizeof(T) + SX * SY;
    constexpr int AXY = (TX + 8) * (TY + 8), BXY = (TX + 6) * (TY + 6);
    constexpr int d1 = MODE == 2 ? (3 * AXY + 5 * BXY) * (int)sizeof(T) + AXY : 0;
    return ((stage > d1 ? stage : d1) + 15) / 16 * 16;
}

// The fused first half step is built for FP32 tiles whose LDS fits a workgroup.
template <class T, int TX, int TY>
constexpr bool fused_d1_fits() {
    return sizeof(T) == 4 && stage_lds<T, TX, TY, 2>() + 64 <= 65536;
}

// Waves per SIMD asked of the register allocator for k_stage: seven for FP32
// (at most 72 VGPRs; the 64x7 stage kernels use all 72 without spills) and six for the
// compensated variant (80 VGPRs: at 72 it spills), unless the tile's LDS
// already admits fewer workgroups per CU (160 KiB of LDS, one wave per SIMD
// per 256-thread workgroup); FP64 keeps its registers.
template <class T, int TX, int TY, int MODE, bool KAHAN>
constexpr int stage_waves() {
    constexpr int lds = stage_lds<T, TX, TY, MODE>() + 64;
    constexpr int by_lds = 163840 / lds;
    constexpr int cap = (KAHAN || MODE == 2) ? 6 : 7;
    return sizeof(T) != 4 ? 1 : (by_lds < cap ? by_lds : cap);
}

template <class T, int TX, int TY, int MODE, bool KAHAN>
__global__ __launch_bounds__(NT, (stage_waves<T, TX, TY, MODE, KAHAN>())) void k_stage(const CaseDesc<T>* __restrict__ cases, const WorkItem* __restrict__ work,
                                              const int* __restrict__ count, Roles ri, Roles rv) {
    constexpr int SX = TX + 4, SY = TY + 4, CPT = (TX * TY + NT - 1) / NT;
    constexpr int FX = (TX + 1) * TY, FY = TX * (TY + 1);
    constexpr bool SPLIT = split_faces<TX, TY>();
    constexpr int FR = flux_row<TX, TY>();
    constexpr int YO = SPLIT ? FX : 0;   // offset of the y fluxes in a flux row
    static_assert(MODE != 2 || fused_d1_fits<T, TX, TY>(), "fused half step: FP32 tiles within 64 KiB of LDS");
    __shared__ __align__(16) char smem[stage_lds<T, TX, TY, MODE>()];
    T* const s_r = (T*)smem;
    T* const s_u = s_r + SX * SY;
    T* const s_v = s_u + SX * SY;
    T* const s_p = s_v + SX * SY;
    // five flux components per face
    T (*const s_f)[FR] = (T (*)[FR])(s_p + SX * SY);
    uint8_t* const s_ct = (uint8_t*)(s_p + SX * SY + 5 * FR);
    // CFL maxima: FP64 for the FP64 state and the compensated variant
    constexpr bool DT64 = KAHAN || sizeof(T) == 8;
    using MT = typename std::conditional<DT64, double, float>::type;
    __shared__ MT s_red[1][NT / 64];

    const int n_items = *count;
    for (int w = blockIdx.x; w < n_items; w += gridDim.x) {
        const WorkItem it = work[w];
        if (it.index < 0) continue;
        const CaseDesc<T>& cd = cases[it.case_id];
        const CaseConsts<T>& k = cd.c;
        const bool visc = k.visc != 0;
        const Roles r = visc ? rv : ri;
        const int nx = k.nx, ny = k.ny;
        const int i0 = (it.index % cd.tiles_x) * TX, j0 = (it.index / cd.tiles_x) * TY;
        // dt: the small tiles (at most two staging passes) form it once their
        // staging loads are in flight; the larger ones at the top, where it
        // costs a round trip but holds no registers across the staging loop
        // (formed late, the 64x7 stage kernels ran 1.2 us slower)
        constexpr bool LATE_DT = (SX * SY + NT - 1) / NT <= 2;
        const SlotPair<DT64> slot_pair = load_slot_pair<DT64>(cd, r.slot_r);
        double dtd = 0.0;
        if constexpr (!LATE_DT) dtd = dt_from_pair<DT64>(cd, slot_pair);

        // thread index opaque per tile (indices derived from it are not held
        // in registers across the tile loop)
        int tid = (int)threadIdx.x;
        asm volatile("" : "+v"(tid));
        __builtin_assume(tid >= 0 && tid < NT);   // tile indices divide as unsigned
        if (it.index == 0 && threadIdx.x == 0) {
            if (r.zero_slot) {
                cd.b.slots[r.slot_w * 2] = 0u; cd.b.slots[r.slot_w * 2 + 1] = 0u;
                if constexpr (DT64) { cd.slots64[r.slot_w * 2] = 0.0; cd.slots64[r.slot_w * 2 + 1] = 0.0; }
            }
            if (!LATE_DT && r.count_step) { cd.b.t[0] += dtd; cd.b.steps[0] += 1; }
        }

        const Plane<T> in0(cd.b.U[r.in][0]), in1(cd.b.U[r.in][1]), in2(cd.b.U[r.in][2]), in3(cd.b.U[r.in][3]),
            in4(cd.b.U[r.in][4]);
        const Plane<uint8_t> ctp(cd.b.ct);

        if (MODE == 2 && visc) {
            // ---- first half step on the tile + halo 2 from U^n on halo 4:
            // (u, v, T) and tags on halo 4, stress and heat flux on halo 3,
            // then the updated cells' (rho, u, v, p) staged as below
            constexpr int AX = TX + 8, AY = TY + 8, BX = TX + 6, BY = TY + 6, BXY = BX * BY;
            T* const a_u = (T*)smem;
            T* const a_v = a_u + AX * AY;
            T* const a_T = a_v + AX * AY;
            T* const b_f = a_T + AX * AY;
            uint8_t* const a_ct = (uint8_t*)(b_f + 5 * BXY);
            const ViscConsts<T> vc = visc_consts(k);
            for (int n = tid; n < AX * AY; n += NT) {
                int li = n % AX, lj = n / AX;
                int gi = min(max(i0 + li - 4, 0), nx - 1);
                int gj = min(max(j0 + lj - 4, 0), ny - 1);
                const unsigned g = cell_index(gi, gj, nx);
                T u, v, p, tt;
                primitives(k, in0[g], in1[g], in2[g], in3[g], in4[g], u, v, p, tt);
                a_u[n] = u; a_v[n] = v; a_T[n] = tt; a_ct[n] = ctp[g];
            }
            __syncthreads();
            for (int n = tid; n < BXY; n += NT) {
                int ti = n % BX, tj = n / BX;
                viscous_flux<T, AX>(k, vc, a_u, a_v, a_T, a_ct, (tj + 1) * AX + (ti + 1), i0 + ti - 3, j0 + tj - 3, n,
                                    b_f, b_f + BXY, b_f + 2 * BXY, b_f + 3 * BXY, b_f + 4 * BXY);
            }
            __syncthreads();
            const T hdt = (T)(dt_from_pair<DT64>(cd, slot_pair) * 0.5);
            // updated cells in registers until the half step's arrays are dead
            constexpr int SN = (SX * SY + NT - 1) / NT;
            T hq[SN][4];
            uint8_t hc[SN];
#pragma unroll
            for (int m = 0; m < SN; m++) {
                const int n = tid + m * NT;
                if (n >= SX * SY) break;
                const int li = n % SX, lj = n / SX;
                const int ci = i0 + li - 2, cj = j0 + lj - 2;
                const int gi = min(max(ci, 0), nx - 1), gj = min(max(cj, 0), ny - 1);
                const unsigned g = cell_index(gi, gj, nx);
                const int sa = (lj + 2) * AX + (li + 2);
                const uint8_t c = a_ct[sa];
                T v[5] = {in0[g], in1[g], in2[g], in3[g], in4[g]};
                const bool interior = !(c == CT_SOLID || c == CT_GHOST) && gi > 0 && gi < nx - 1 && gj > 0 && gj < ny - 1;
                if (interior)
                    viscous_update<T, AX, BX>(k, a_u, a_v, sa, b_f, b_f + BXY, b_f + 2 * BXY, b_f + 3 * BXY, b_f + 4 * BXY,
                                              (lj + 1) * BX + (li + 1), hdt, v);
                // the tile's own cells: the half-step result (rho unchanged:
                // that set's density plane is U^n's)
                if (li >= 2 && li < TX + 2 && lj >= 2 && lj < TY + 2 && ci < nx && cj < ny)
                    for (int f = 1; f < 5; f++) Plane<T>(cd.b.U[r.blend][f]).put(g, v[f]);
                T u, vv, p;
                primitives_uvp(k, v[0], v[1], v[2], v[3], v[4], u, vv, p);
                hq[m][0] = v[0]; hq[m][1] = u; hq[m][2] = vv; hq[m][3] = p;
                hc[m] = c;
            }
            __syncthreads();
#pragma unroll
            for (int m = 0; m < SN; m++) {
                const int n = tid + m * NT;
                if (n >= SX * SY) break;
                s_r[n] = hq[m][0]; s_u[n] = hq[m][1]; s_v[n] = hq[m][2]; s_p[n] = hq[m][3]; s_ct[n] = hc[m];
            }
        } else {
            // ---- stage tile + halo 2: conserved -> (rho, u, v, p) in LDS.
            // Cell n of the staged region from its loaded state q and tag c.
            auto stage_cell = [&](int n, T q[5], uint8_t c) {
                const int li = n % SX, lj = n / SX;
                if (MODE == 1 && c == CT_GHOST) {
                    const int gi = min(max(i0 + li - 2, 0), nx - 1);
                    const int gj = min(max(j0 + lj - 2, 0), ny - 1);
                    const unsigned g = cell_index(gi, gj, nx);
                    // Ghost fill of the stage-1 result, done here instead of in a
                    // launch of its own: ghost neighbours at their pre-BC value
                    // (the blend input), others as stage 1 left them.  The tile that
                    // owns the cell stores it for the later kernels: into the stage-1
                    // set (inviscid: the final ghost fill blends it), or blended into
                    // the output (viscous: the second half step reads it).
                    ghost_value<T, false>(cd, r.in, r.blend, -1, ghost_index(cd.b, gi, gj), q);
                    const bool own = li >= 2 && li < TX + 2 && lj >= 2 && lj < TY + 2 && i0 + li - 2 < nx && j0 + lj - 2 < ny;
                    if (own) {
                        if (!visc) {
                            for (int f = 0; f < 5; f++) Plane<T>(cd.b.U[r.in][f]).put(g, q[f]);
                        } else {
                            for (int f = 0; f < 5; f++) {
                                T bl = Plane<T>(cd.b.U[r.blend][f])[g];
                                Plane<T>(cd.b.U[r.out][f]).put(g, T(0.5) * bl + T(0.5) * q[f]);
                            }
                        }
                    }
                }
                T u, v, p;
                primitives_uvp(k, q[0], q[1], q[2], q[3], q[4], u, v, p);
                s_r[n] = q[0]; s_u[n] = u; s_v[n] = v; s_p[n] = p; s_ct[n] = c;
            };
            constexpr int SN = (SX * SY + NT - 1) / NT;
            if constexpr (SN <= 2) {
                // small tiles (on small grids one workgroup per CU, nothing to
                // hide latency behind): every staged load issued first
                T sq[SN][5];
                uint8_t sc[SN];
#pragma unroll
                for (int m = 0; m < SN; m++) {
                    const int n = tid + m * NT;
                    if (n < SX * SY) {
                        const int li = n % SX, lj = n / SX;
                        const int gi = min(max(i0 + li - 2, 0), nx - 1);
                        const int gj = min(max(j0 + lj - 2, 0), ny - 1);
                        const unsigned g = cell_index(gi, gj, nx);
                        sc[m] = ctp[g];
                        sq[m][0] = in0[g]; sq[m][1] = in1[g]; sq[m][2] = in2[g]; sq[m][3] = in3[g]; sq[m][4] = in4[g];
                    }
                }
#pragma unroll
                for (int m = 0; m < SN; m++) {
                    const int n = tid + m * NT;
                    if (n < SX * SY) stage_cell(n, sq[m], sc[m]);
                }
            } else {
                for (int n = tid; n < SX * SY; n += NT) {
                    const int li = n % SX, lj = n / SX;
                    const int gi = min(max(i0 + li - 2, 0), nx - 1);
                    const int gj = min(max(j0 + lj - 2, 0), ny - 1);
                    const unsigned g = cell_index(gi, gj, nx);
                    const uint8_t c = ctp[g];
                    T q[5] = {in0[g], in1[g], in2[g], in3[g], in4[g]};
                    stage_cell(n, q, c);
                }
            }
        }
        if constexpr (LATE_DT) {
            dtd = dt_from_pair<DT64>(cd, slot_pair);
            if (r.count_step && it.index == 0 && threadIdx.x == 0) { cd.b.t[0] += dtd; cd.b.steps[0] += 1; }
        }
        const T dt = (T)dtd;
        __syncthreads();

        // ---- x faces i+1/2 for local i in [-1, TX-1]
        auto x_face = [&](int n) {
            int fi = n % (TX + 1) - 1, fj = n / (TX + 1);
            int gi = i0 + fi, gj = j0 + fj;
            T o[5] = {T(0), T(0), T(0), T(0), T(0)};
            if (gi >= 0 && gi < nx - 1 && gj < ny) {
                int b = (fj + 2) * SX + (fi + 1);
                T rr[4], un[4], ut[4], pp[4];
                uint8_t cc[4];
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    rr[m] = s_r[b + m]; un[m] = s_u[b + m]; ut[m] = s_v[b + m]; pp[m] = s_p[b + m]; cc[m] = s_ct[b + m];
                }
                face_flux(k, rr, un, ut, pp, cc, o);
            }
#pragma unroll
            for (int c = 0; c < 5; c++) s_f[c][n] = o[c];
        };
        // ---- y faces j+1/2 for local j in [-1, TY-1]
        auto y_face = [&](int n) {
            int fi = n % TX, fj = n / TX - 1;
            int gi = i0 + fi, gj = j0 + fj;
            T o[5] = {T(0), T(0), T(0), T(0), T(0)};
            if (gj >= 0 && gj < ny - 1 && gi < nx) {
                int b = (fj + 1) * SX + (fi + 2);
                T rr[4], un[4], ut[4], pp[4];
                uint8_t cc[4];
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    int s = b + m * SX;
                    rr[m] = s_r[s]; un[m] = s_v[s]; ut[m] = s_u[s]; pp[m] = s_p[s]; cc[m] = s_ct[s];
                }
                face_flux(k, rr, un, ut, pp, cc, o);
            }
            s_f[0][YO + n] = o[0]; s_f[1][YO + n] = o[2]; s_f[2][YO + n] = o[1]; s_f[3][YO + n] = o[3];
            s_f[4][YO + n] = o[4];
        };
        if constexpr (SPLIT) {
            if (tid < NT / 2) {
                if (tid < FX) x_face(tid);
            } else if (tid - NT / 2 < FY) {
                y_face(tid - NT / 2);
            }
        } else {
            for (int n = tid; n < FX; n += NT) x_face(n);
        }
        __syncthreads();

        T res[CPT][5];
        // x-face mass fluxes of the -p div(v) source term
        T mxr[CPT], mxl[CPT];
#pragma unroll
        for (int q = 0; q < CPT; q++) {
            int n = tid + q * NT;
            if (n >= TX * TY) break;
            int li = n % TX, lj = n / TX;
            int a = lj * (TX + 1) + li;
#pragma unroll
            for (int c = 0; c < 5; c++) res[q][c] = (s_f[c][a + 1] - s_f[c][a]) * (T)k.inv_dx;
            mxr[q] = s_f[0][a + 1];
            mxl[q] = s_f[0][a];
        }
        if constexpr (!SPLIT) {
            __syncthreads();
            for (int n = tid; n < FY; n += NT) y_face(n);
            __syncthreads();
        }

        // the state the update starts from: set r.in, or the fused half step's
        // result stored above
        const int vs = (MODE == 2 && visc) ? r.blend : r.in;
        const Plane<T> st1(cd.b.U[vs][1]), st2(cd.b.U[vs][2]), st3(cd.b.U[vs][3]), st4(cd.b.U[vs][4]);
        MT vmax_local = MT(0);
#pragma unroll
        for (int q = 0; q < CPT; q++) {
            int n = tid + q * NT;
            if (n >= TX * TY) break;
            int li = n % TX, lj = n / TX;
            int gi = i0 + li, gj = j0 + lj;
            int a = lj * TX + li;
            T dfy[5];
#pragma unroll
            for (int c = 0; c < 5; c++) dfy[c] = (s_f[c][YO + a + TX] - s_f[c][YO + a]) * (T)k.inv_dy;
            if (gi >= nx || gj >= ny) continue;
            const unsigned g = cell_index(gi, gj, nx);
            const int sc = (lj + 2) * SX + (li + 2);
            const uint8_t c = s_ct[sc];
            T v0[5] = {s_r[sc], st1[g], st2[g], st3[g], st4[g]};
            T nv[5];
            if (c == CT_SOLID || c == CT_GHOST) {
#pragma unroll
                for (int f = 0; f < 5; f++) nv[f] = v0[f];
            } else {
                // updateKernel, src/solver.cu:1736-1868
                const bool ex = (gi == 0 || gi == nx - 1), ey = (gj == 0 || gj == ny - 1);
                T src = T(0);
                if (!ex && !ey) {
                    // interface velocities: face mass flux over the mean clamped
                    // density of the face's two cells (solver.cu:1787-1794)
                    const T rc = tmax(s_r[sc], (T)k.min_rho);
                    const T fxr = T(2) * mxr[q] * Num<T>::rcp(rc + tmax(s_r[sc + 1], (T)k.min_rho));
                    const T fxl = T(2) * mxl[q] * Num<T>::rcp(tmax(s_r[sc - 1], (T)k.min_rho) + rc);
                    const T fyt = T(2) * s_f[0][YO + a + TX] * Num<T>::rcp(rc + tmax(s_r[sc + SX], (T)k.min_rho));
                    const T fyb = T(2) * s_f[0][YO + a] * Num<T>::rcp(tmax(s_r[sc - SX], (T)k.min_rho) + rc);
                    T divv = (fxr - fxl) * (T)k.inv_dx + (fyt - fyb) * (T)k.inv_dy;
                    src = -s_p[sc] * divv;
                }
                T d[5];
#pragma unroll
                for (int f = 0; f < 5; f++) d[f] = (ex ? T(0) : res[q][f]) + (ey ? T(0) : dfy[f]);
                T nr = v0[0] - dt * d[0];
                T nru = v0[1] - dt * d[1];
                T nrv = v0[2] - dt * d[2];
                T nE = v0[3] - dt * d[3];
                T nre = v0[4] - dt * d[4] + dt * src;
                if (!Num<T>::finite(nr) || nr < (T)k.min_rho) nr = tmax(v0[0], (T)k.min_rho);
                if (!Num<T>::finite(nru)) nru = v0[1];
                if (!Num<T>::finite(nrv)) nrv = v0[2];
                if (!Num<T>::finite(nE)) nE = v0[3];
                if (!Num<T>::finite(nre) || nre < (T)k.min_p * (T)k.inv_gm1) nre = v0[4];
                nr = tmax(nr, (T)k.min_rho);
                T ir = Num<T>::rcp(nr);
                T un = nru * ir, vn = nrv * ir;
                T v2 = un * un + vn * vn;
                if (v2 > (T)k.vmax * (T)k.vmax) {
                    T s = (T)k.vmax * Num<T>::rsqrt(v2);
                    un *= s; vn *= s;
                    nru = nr * un; nrv = nr * vn;
                }
                T ke = T(0.5) * nr * (un * un + vn * vn);
                T e = (nE > T(0) && nre < T(0.001f) * nE) ? nre : nE - ke;
                T rRi = nr * (T)k.R * (T)k.inv_gm1;
                e = tmax(e, rRi * (T)k.min_T);
                e = tmin(e, rRi * (T)k.max_T);
                nv[0] = nr; nv[1] = nru; nv[2] = nrv; nv[3] = e + ke; nv[4] = e;
            }
            T out[5];
            T carry[5] = {T(0), T(0), T(0), T(0), T(0)};
            if (MODE != 1) {
#pragma unroll
                for (int f = 0; f < 5; f++) out[f] = nv[f];
            } else {
#pragma unroll
                for (int f = 0; f < 5; f++) {
                    T bl = Plane<T>(cd.b.U[r.blend][f])[g];
                    if (KAHAN) {
                        T delta = T(0.5) * (nv[f] - bl);
                        T y = delta - (T)Plane<float>(cd.b.K[f])[g];
                        T t = bl + y;
                        carry[f] = (t - bl) - y;
                        out[f] = t;
                    } else {
                        out[f] = T(0.5) * bl + T(0.5) * nv[f];
                    }
                }
            }
            if (MODE == 1 && c == CT_GHOST) {
                // stored by the staging loop
                if (KAHAN)
                    for (int f = 0; f < 5; f++) Plane<float>(cd.b.K[f]).put(g, 0.0f);
                continue;
            }
            const bool apply_bc = (MODE != 1) || !visc;
            if (apply_bc) {
                if (c == CT_GHOST) {
                    if (KAHAN && MODE == 1)
                        for (int f = 0; f < 5; f++) Plane<float>(cd.b.K[f]).put(g, 0.0f);
                    continue;  // filled by k_ghost
                }
                bool touched = bc_pointwise(k, gi, gj, c, out);
                if (MODE == 1) vmax_local = tmax(vmax_local, wave_speed_as<MT>(k, out));
                if (KAHAN && MODE == 1)
                    for (int f = 0; f < 5; f++) Plane<float>(cd.b.K[f]).put(g, touched ? 0.0f : (float)carry[f]);
            }
#pragma unroll
            for (int f = 0; f < 5; f++) Plane<T>(cd.b.U[r.out][f]).put(g, out[f]);
        }
        if (MODE == 1 && !visc) {
            const MT vm[1] = {vmax_local};
            if constexpr (DT64) block_max_store<NT, 1>(vm, s_red, cd.tile_max64, cd.tiles_x * cd.tiles_y, it.index);
            else block_max_store<NT, 1>(vm, s_red, cd.b.tile_max, cd.tiles_x * cd.tiles_y, it.index);
        }
        __syncthreads();
    }
}

#pragma clang fp contract(off)
// Normal and ranked neighbour list of each ghost cell (static per geometry).
template <class T>
__global__ __launch_bounds__(NT) void k_ghost_prep(const CaseDesc<T>* __restrict__ cases, const WorkItem* __restrict__ items,
                                                   int n_items) {
    const int w = blockIdx.x * NT + threadIdx.x;
    if (w >= n_items) return;
    const WorkItem it = items[w];
    const CaseDesc<T>& cd = cases[it.case_id];
    const CaseConsts<T>& k = cd.c;
    GhostInfo<T>* gi = (GhostInfo<T>*)cd.b.ghosts + it.index;
    const int nx = k.nx, ny = k.ny;
    const size_t idx = (size_t)gi->idx;
    const int i = (int)(idx % nx), j = (int)(idx / nx);
    const float* sdf = cd.b.sdf;
    float sxp = (i < nx - 1) ? sdf[idx + 1] : sdf[idx];
    float sxm = (i > 0) ? sdf[idx - 1] : sdf[idx];
    float syp = (j < ny - 1) ? sdf[idx + nx] : sdf[idx];
    float sym = (j > 0) ? sdf[idx - nx] : sdf[idx];
    T nxs = ((T)sxp - (T)sxm) / ((T)2 * k.dx);
    T nys = ((T)syp - (T)sym) / ((T)2 * k.dy);
    T nl = ::sqrt(nxs * nxs + nys * nys) + (T)1e-10f;
    nxs = nxs / nl;
    nys = nys / nl;
    T dots[8];
    int codes[8];
    int cnt = 0;
    for (int m = 0; m < 8; m++) {
        int ni = i + nb_dx(m), nj = j + nb_dy(m);
        if (ni < 0 || ni >= nx || nj < 0 || nj >= ny) continue;
        if (!(sdf[(size_t)nj * nx + ni] > 0.0f)) continue;
        T dxn = (T)nb_dx(m), dyn = (T)nb_dy(m);
        T len = ::sqrt(dxn * dxn + dyn * dyn);
        T dot = (dxn * nxs + dyn * nys) / len;
        if (!(dot > T(-1))) continue;
        // insertion: descending dot, earlier scan position first on ties
        int q = cnt++;
        while (q > 0 && dots[q - 1] < dot) {
            dots[q] = dots[q - 1];
            codes[q] = codes[q - 1];
            q--;
        }
        dots[q] = dot;
        codes[q] = m;
    }
    uint32_t order = (uint32_t)cnt << 24;
    for (int q = 0; q < cnt; q++) order |= (uint32_t)codes[q] << (3 * q);
    gi->order = order;
    gi->nx = nxs;
    gi->ny = nys;
}

// Immersed-boundary ghost fill (applyBoundaryConditionsKernel ghost branch,
// src/solver.cu:2045-2225) over a list of ghost cells.  A neighbour that is a
// ghost cell is read at its pre-BC value (set pre0, or the SSP blend of pre0
// and pre1); any other neighbour at its final value in set out.
// FINAL: also the CFL maxima of the filled cells.
// Wave-level maximum into a case's CFL slot counter: when every lane of the
// wave works on the same case, one atomic per wave; otherwise one per lane.
// Called by whole waves.
// FP32 maxima go to an unsigned bit-pattern slot, FP64 maxima to a double slot.
template <class V, class TG>
__device__ __forceinline__ void wave_slot_max(TG* target, bool active, int case_id, V v) {
    const int c0 = __shfl(case_id, 0, 64);
    const bool uniform = __all(!active || case_id == c0);
    if (uniform) {
        V m = active ? v : V(0);
        for (int off = 32; off > 0; off >>= 1) m = tmax(m, (V)__shfl_xor(m, off, 64));
        if ((threadIdx.x & 63) == 0 && m > V(0)) atomic_max_bits(target, m);
    } else if (active && v > V(0)) {
        atomic_max_bits(target, v);
    }
}

// Workgroups [gblocks, gridDim.x) of the final ghost launch: reduce the
// per-tile maxima (written by the step's last tiled kernel with plain stores)
// into each case's CFL slot, one item (case, chunk of REDUCE_CHUNK tiles) per
// workgroup pass and one atomic per item and quantity.  Each thread issues its
// loads of a chunk together, so a chunk costs one memory latency.
// One chunk of one case's per-tile maxima (V = float or double) into its slot.
template <class V, class TG>
__device__ __forceinline__ void reduce_tile_chunk(const V* tile_max, int nt, int chunk, TG* slot, V (*s_red)[NT / 64]) {
    constexpr int PER = REDUCE_CHUNK / NT;
    const Plane<V> tm(tile_max);
    const int t0 = chunk * REDUCE_CHUNK + threadIdx.x;
    V a0[PER], a1[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int t = t0 + q * NT;
        a0[q] = t < nt ? tm[(unsigned)t] : V(0);
        a1[q] = t < nt ? tm[(unsigned)(nt + t)] : V(0);
    }
    V m0 = V(0), m1 = V(0);
#pragma unroll
    for (int q = 0; q < PER; q++) {
        m0 = tmax(m0, a0[q]);
        m1 = tmax(m1, a1[q]);
    }
    for (int off = 32; off > 0; off >>= 1) {
        m0 = tmax(m0, (V)__shfl_xor(m0, off, 64));
        m1 = tmax(m1, (V)__shfl_xor(m1, off, 64));
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) { s_red[0][wave] = m0; s_red[1][wave] = m1; }
    __syncthreads();
    if (threadIdx.x == 0) {
        V a = s_red[0][0], c = s_red[1][0];
        for (int w = 1; w < NT / 64; w++) { a = tmax(a, s_red[0][w]); c = tmax(c, s_red[1][w]); }
        if (a > V(0)) atomic_max_bits(&slot[0], a);
        if (c > V(0)) atomic_max_bits(&slot[1], c);
    }
    __syncthreads();
}

// DT64: the engine keeps FP64 maxima (every case of an engine does the same).
template <class T, bool DT64>
__device__ void reduce_tile_maxima(const CaseDesc<T>* __restrict__ cases, const WorkItem* __restrict__ case_list,
                                   int nitems, int rblock, int nrblocks, Roles ri, Roles rv) {
    using MT = typename std::conditional<DT64, double, float>::type;
    __shared__ MT s_red[2][NT / 64];
    for (int ci = rblock; ci < nitems; ci += nrblocks) {
        const WorkItem it = case_list[ci];
        const CaseDesc<T>& cd = cases[it.case_id];
        const Roles r = cd.c.visc ? rv : ri;
        const int nt = cd.tiles_x * cd.tiles_y;
        if constexpr (DT64) reduce_tile_chunk<double>(cd.tile_max64, nt, it.index, &cd.slots64[r.slot_w * 2], s_red);
        else reduce_tile_chunk<float>(cd.b.tile_max, nt, it.index, &cd.b.slots[r.slot_w * 2], s_red);
    }
}

// One ghost cell per lane: fill, store, and (FINAL) its CFL maxima.
template <class T, bool FINAL, bool DT64>
__device__ __forceinline__ void ghost_item(const CaseDesc<T>& cd, const WorkItem it, bool active, Roles ri, Roles rv) {
    const CaseConsts<T>& k = cd.c;
    const Roles r = k.visc ? rv : ri;
    const CaseBufs& b = cd.b;
    const unsigned idx = (unsigned)((const gmem<GhostInfo<T>>*)b.ghosts)[it.index].idx;
    // the cell's own pre-BC state (mu of the CFL maxima) is loaded together
    // with the fill's loads
    T pv[5];
    if (FINAL)
        for (int f = 0; f < 5; f++) pv[f] = pre_val<T>(b, r.pre0, r.pre1, idx, f);
    T o[5];
    ghost_value<T, true>(cd, r.out, r.pre0, r.pre1, it.index, o);
    if (active)
        for (int f = 0; f < 5; f++) Plane<T>(b.U[r.out][f]).put(idx, o[f]);
    if (FINAL) {
        T u_, v_, p_, t_;
        primitives(k, pv[0], pv[1], pv[2], pv[3], pv[4], u_, v_, p_, t_);
        const T nu_c = sutherland(k, t_) * Num<T>::rcp(o[0] + T(1e-10f));
        // every lane takes part; a lane of another case keeps its own counter
        if constexpr (DT64) {
            const double s = active ? wave_speed_as<double>(k, o) : 0.0;
            const double nu = (active && k.visc) ? (double)nu_c : 0.0;
            wave_slot_max(&cd.slots64[r.slot_w * 2], active, it.case_id, s);
            wave_slot_max(&cd.slots64[r.slot_w * 2 + 1], active && k.visc, it.case_id, nu);
        } else {
            const float s = active ? wave_speed_as<float>(k, o) : 0.0f;
            const float nu = (active && k.visc) ? (float)nu_c : 0.0f;
            wave_slot_max(&b.slots[r.slot_w * 2], active, it.case_id, s);
            wave_slot_max(&b.slots[r.slot_w * 2 + 1], active && k.visc, it.case_id, nu);
        }
    }
}

template <class T, bool FINAL, bool DT64>
__global__ __launch_bounds__(NT) void k_ghost(const CaseDesc<T>* __restrict__ cases, const WorkItem* __restrict__ items,
                                              const int* __restrict__ count, Roles ri, Roles rv, int gblocks,
                                              const WorkItem* __restrict__ case_list, const int* __restrict__ ncases,
                                              int single_case) {
    if (FINAL && (int)blockIdx.x >= gblocks) {
        reduce_tile_maxima<T, DT64>(cases, case_list, *ncases, (int)blockIdx.x - gblocks, (int)gridDim.x - gblocks, ri, rv);
        return;
    }
    const int n_items = *count;
    // uniform trip count so that whole waves reach the slot reduction together
    for (int base = blockIdx.x * NT; base < n_items; base += gblocks * NT) {
        const int w = base + threadIdx.x;
        const bool active = w < n_items;
        if (single_case >= 0) {
            // one case: its list is (case, 0 .. n-1), no item to load
            ghost_item<T, FINAL, DT64>(cases[single_case], WorkItem{single_case, active ? w : base}, active, ri, rv);
            continue;
        }
        const WorkItem it = items[active ? w : base];
        // A wave whose lanes all hold one case reads the case table through
        // scalar loads with a wave-uniform index; per-lane indexing would turn
        // every buffer pointer into a vector load that the field load then
        // waits for.
        const int c0 = __builtin_amdgcn_readfirstlane(it.case_id);
        if (__all(it.case_id == c0)) ghost_item<T, FINAL, DT64>(cases[c0], it, active, ri, rv);
        else ghost_item<T, FINAL, DT64>(cases[it.case_id], it, active, ri, rv);
    }
}
#pragma clang fp contract(on)

// ---------------------------------------------------------------------------
// Viscous half step (fusedViscousDiffusionKernel, src/solver.cu:2251-2511)
// from set r.in to set r.out.  Stress and heat flux live in LDS on a 1-cell
// ring around the tile.  FINAL: + point-wise BC + CFL maxima (mu from the
// pre-diffusion temperature as in the reference, rho from the final state).
// ---------------------------------------------------------------------------
// Waves per SIMD asked of the register allocator for k_diffuse: as many
// workgroups as the tile's LDS admits per CU, at most seven, for FP32 (the
// final half step would otherwise take 79 VGPRs and run six waves where its
// LDS allows seven; measured ~2% faster with one spilled register); FP64 keeps
// its registers.
template <class T, int TX, int TY>
constexpr int diffuse_waves() {
    constexpr int SX = TX + 4, SY = TY + 4, QX = TX + 2, QY = TY + 2;
    constexpr int lds = 3 * SX * SY * (int)sizeof(T) + SX * SY + 5 * QX * QY * (int)sizeof(T) + 32;
    constexpr int by_lds = 163840 / lds;
    return sizeof(T) != 4 ? 1 : (by_lds < 7 ? by_lds : 7);
}

template <class T, int TX, int TY, bool FINAL>
__global__ __launch_bounds__(NT, (diffuse_waves<T, TX, TY>())) void k_diffuse(const CaseDesc<T>* __restrict__ cases, const WorkItem* __restrict__ work,
                                                const int* __restrict__ count, Roles r) {
    constexpr int SX = TX + 4, SY = TY + 4, QX = TX + 2, QY = TY + 2, CPT = (TX * TY + NT - 1) / NT;
    __shared__ T s_u[SY * SX], s_v[SY * SX], s_T[SY * SX];
    __shared__ uint8_t s_ct[SY * SX];
    __shared__ T s_txx[QY * QX], s_tyy[QY * QX], s_txy[QY * QX], s_qx[QY * QX], s_qy[QY * QX];
    // CFL maxima in the state's precision (FP64 state: dt64)
    __shared__ T s_red[2][NT / 64];

    const int n_items = *count;
    for (int w = blockIdx.x; w < n_items; w += gridDim.x) {
        const WorkItem it = work[w];
        if (it.index < 0) continue;
        const CaseDesc<T>& cd = cases[it.case_id];
        const CaseConsts<T>& k = cd.c;
        const int nx = k.nx, ny = k.ny;
        const int i0 = (it.index % cd.tiles_x) * TX, j0 = (it.index / cd.tiles_x) * TY;
        // dt: the first half step forms it once its staging loads are in
        // flight; the final one at the top (formed late it ran 0.4 us slower,
        // the first 0.5 us faster)
        constexpr bool LATE_DT = !FINAL;
        const SlotPair<(sizeof(T) == 8)> slot_pair = load_slot_pair<(sizeof(T) == 8)>(cd, r.slot_r);
        double dtd = 0.0;
        if constexpr (!LATE_DT) dtd = dt_from_pair<(sizeof(T) == 8)>(cd, slot_pair);
        // The thread index, opaque to the compiler per tile: every index
        // derived from it is recomputed per tile rather than held in registers
        // across the tile loop (which spilled them).
        int tid = (int)threadIdx.x;
        asm volatile("" : "+v"(tid));
        __builtin_assume(tid >= 0 && tid < NT);   // tile indices divide as unsigned
        if (it.index == 0 && threadIdx.x == 0) {
            if (r.zero_slot) {
                cd.b.slots[r.slot_w * 2] = 0u; cd.b.slots[r.slot_w * 2 + 1] = 0u;
                if constexpr (sizeof(T) == 8) { cd.slots64[r.slot_w * 2] = 0.0; cd.slots64[r.slot_w * 2 + 1] = 0.0; }
            }
            if (!LATE_DT && r.count_step) { cd.b.t[0] += dtd; cd.b.steps[0] += 1; }
        }
        const Plane<T> in[5] = {Plane<T>(cd.b.U[r.in][0]), Plane<T>(cd.b.U[r.in][1]), Plane<T>(cd.b.U[r.in][2]),
                                Plane<T>(cd.b.U[r.in][3]), Plane<T>(cd.b.U[r.in][4])};
        const Plane<T> outp[5] = {Plane<T>(cd.b.U[r.out][0]), Plane<T>(cd.b.U[r.out][1]), Plane<T>(cd.b.U[r.out][2]),
                                  Plane<T>(cd.b.U[r.out][3]), Plane<T>(cd.b.U[r.out][4])};
        const Plane<uint8_t> ctp(cd.b.ct);

        // Every staging load of the tile (the staged cells' states and tags)
        // is issued before the first is used, so that the tile waits for one
        // memory latency, not one per staging pass.
        constexpr int SN = (SX * SY + NT - 1) / NT;
        T sq[SN][5];
        uint8_t sc[SN];
#pragma unroll
        for (int m = 0; m < SN; m++) {
            const int n = tid + m * NT;
            if (n < SX * SY) {
                int li = n % SX, lj = n / SX;
                int gi = min(max(i0 + li - 2, 0), nx - 1);
                int gj = min(max(j0 + lj - 2, 0), ny - 1);
                const unsigned g = cell_index(gi, gj, nx);
#pragma unroll
                for (int f = 0; f < 5; f++) sq[m][f] = in[f][g];
                sc[m] = ctp[g];
            }
        }
#pragma unroll
        for (int m = 0; m < SN; m++) {
            const int n = tid + m * NT;
            if (n < SX * SY) {
                T u, v, p, tt;
                primitives(k, sq[m][0], sq[m][1], sq[m][2], sq[m][3], sq[m][4], u, v, p, tt);
                s_u[n] = u; s_v[n] = v; s_T[n] = tt; s_ct[n] = sc[m];
            }
        }
        if constexpr (LATE_DT) {
            dtd = dt_from_pair<(sizeof(T) == 8)>(cd, slot_pair);
            if (r.count_step && it.index == 0 && threadIdx.x == 0) { cd.b.t[0] += dtd; cd.b.steps[0] += 1; }
        }
        const T hdt = (T)(dtd * 0.5);
        const ViscConsts<T> vc = visc_consts(k);
        __syncthreads();

        for (int n = tid; n < QX * QY; n += NT) {
            int ti = n % QX, tj = n / QX;
            viscous_flux<T, SX>(k, vc, s_u, s_v, s_T, s_ct, (tj + 1) * SX + (ti + 1), i0 + ti - 1, j0 + tj - 1, n,
                                s_txx, s_tyy, s_txy, s_qx, s_qy);
        }
        __syncthreads();

        T smax = T(0), numax = T(0);
#pragma unroll
        for (int q = 0; q < CPT; q++) {
            int n = tid + q * NT;
            if (n >= TX * TY) break;
            int li = n % TX, lj = n / TX;
            int gi = i0 + li, gj = j0 + lj;
            if (gi >= nx || gj >= ny) continue;
            const unsigned g = cell_index(gi, gj, nx);
            int s = (lj + 2) * SX + (li + 2);
            int t = (lj + 1) * QX + (li + 1);
            uint8_t c = s_ct[s];
            T v[5];
            for (int f = 0; f < 5; f++) v[f] = in[f][g];
            T Tin = s_T[s];
            bool interior = !(c == CT_SOLID || c == CT_GHOST) && gi > 0 && gi < nx - 1 && gj > 0 && gj < ny - 1;
            if (interior) viscous_update<T, SX, QX>(k, s_u, s_v, s, s_txx, s_tyy, s_txy, s_qx, s_qy, t, hdt, v);
            if (FINAL) {
                if (c == CT_GHOST) continue;  // filled by k_ghost
                bc_pointwise(k, gi, gj, c, v);
                smax = tmax(smax, wave_speed_as<T>(k, v));
                T mu = (c == CT_SOLID) ? T(0) : sutherland(k, Tin, vc.i_tref);
                numax = tmax(numax, mu * Num<T>::rcp(v[0] + T(1e-10f)));
            }
            // the first half step leaves rho alone: set D's density plane is set A's
            for (int f = FINAL ? 0 : 1; f < 5; f++) outp[f].put(g, v[f]);
        }
        if (FINAL) {
            const T vm[2] = {smax, numax};
            if constexpr (sizeof(T) == 8) block_max_store<NT, 2>(vm, s_red, cd.tile_max64, cd.tiles_x * cd.tiles_y, it.index);
            else block_max_store<NT, 2>(vm, s_red, cd.b.tile_max, cd.tiles_x * cd.tiles_y, it.index);
        }
        __syncthreads();
    }
}

// CFL maxima of set r.in over a tile list with mu from the same state
// (reset: the reference writes mu from the post-BC initial state, solver.cu:3154).
template <class T, int TX, int TY>
__global__ __launch_bounds__(NT) void k_maxima(const CaseDesc<T>* __restrict__ cases, const WorkItem* __restrict__ work,
                                               const int* __restrict__ count, Roles r) {
    __shared__ double s_red[2][NT / 64];
    const int n_items = *count;
    for (int w = blockIdx.x; w < n_items; w += gridDim.x) {
        const WorkItem it = work[w];
        if (it.index < 0) continue;
        const CaseDesc<T>& cd = cases[it.case_id];
        const CaseConsts<T>& k = cd.c;
        const int i0 = (it.index % cd.tiles_x) * TX, j0 = (it.index / cd.tiles_x) * TY;
        // FP32 maxima (exact as doubles) or FP64 ones for dt64 cases
        double smax = 0.0, numax = 0.0;
        for (int q = threadIdx.x; q < TX * TY; q += NT) {
            int gi = i0 + q % TX, gj = j0 + q / TX;
            if (gi >= k.nx || gj >= k.ny) continue;
            size_t g = (size_t)gj * k.nx + gi;
            T v[5];
            for (int f = 0; f < 5; f++) v[f] = ((const T*)cd.b.U[r.in][f])[g];
            T tt;
            const T ws = wave_speed(k, v, &tt);
            T mu = (cd.b.ct[g] == CT_SOLID) ? T(0) : sutherland(k, tt);
            const T nu = mu * Num<T>::rcp(v[0] + T(1e-10f));
            if (cd.dt64) {
                smax = fmax(smax, wave_speed_as<double>(k, v));
                numax = fmax(numax, (double)nu);
            } else {
                smax = fmax(smax, (double)(float)ws);
                numax = fmax(numax, (double)(float)nu);
            }
        }
        if (cd.dt64) {
            block_max_atomic<NT>(smax, s_red[0], &cd.slots64[r.slot_w * 2]);
            block_max_atomic<NT>(numax, s_red[1], &cd.slots64[r.slot_w * 2 + 1]);
        } else {
            block_max_atomic<NT>((float)smax, (float*)s_red[0], &cd.b.slots[r.slot_w * 2]);
            block_max_atomic<NT>((float)numax, (float*)s_red[1], &cd.b.slots[r.slot_w * 2 + 1]);
        }
        __syncthreads();
    }
}

__global__ void k_zero_u32(unsigned* p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// Derived fields of set `set` (computeTemperature/Pressure/VelocityMag/Mach-
// DirectKernel, cellTypeToFloatKernel, density copy: solver.cu:3444-3671).
template <class T>
__global__ __launch_bounds__(NT) void k_field(const CaseDesc<T>* __restrict__ cases, int case_id, int set, int field, float* out) {
    const CaseDesc<T>& cd = cases[case_id];
    const CaseConsts<T>& k = cd.c;
    size_t n = (size_t)k.nx * k.ny;
    const T* rho = (const T*)cd.b.U[set][0];
    const T* ru = (const T*)cd.b.U[set][1];
    const T* rv = (const T*)cd.b.U[set][2];
    const T* re = (const T*)cd.b.U[set][4];
    for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < n; i += (size_t)gridDim.x * NT) {
        float o = 0.0f;
        switch (field) {
        case 0: { T e = re[i] / (rho[i] + T(1e-10f)); o = (float)((T)k.gm1 * e / (T)k.R); break; }
        case 1: o = (float)((T)k.gm1 * re[i]); break;
        case 2: o = (float)rho[i]; break;
        case 3: { T r_ = rho[i] + T(1e-10f); T u = ru[i] / r_, v = rv[i] / r_; o = (float)::sqrt(u * u + v * v); break; }
        case 4: {
            T r_ = rho[i] + T(1e-10f); T u = ru[i] / r_, v = rv[i] / r_;
            T sp = ::sqrt(u * u + v * v);
            T Tt = (T)k.gm1 * (re[i] / r_) / (T)k.R;
            T c = ::sqrt((T)k.gamma * (T)k.R * Tt);
            o = (float)(sp / (c + T(1e-10f)));
            break;
        }
        default: o = (float)cd.b.ct[i]; break;
        }
        out[i] = o;
    }
}

// Velocity arrows for the display overlay (generateVectorArrowsKernel,
// solver.cu:3689-3796): one candidate per sample cell (i, j) = (a step + step/2,
// b step + step/2); a cell moving at less than 1e-6 u_inf gets none.  An arrow
// is 8 vertices (x, y, r, g, b) in normalised device coordinates, drawn as 4
// lines: the shaft, the two head strokes (black) and a zero-length line.  Slots come from an atomic counter, one
// atomic per wave; vertices beyond max_vertices are not written (the counter
// still counts them), so the caller's buffer is never overrun.  The velocity
// is that of the primitives, u = rho_u / max(rho, rho_min).
template <class T>
__global__ __launch_bounds__(NT) void k_arrows(const CaseDesc<T>* __restrict__ cases, int case_id, int set, int step,
                                               int nax, int nay, float u_inf, float max_len, float cos_a, float sin_a,
                                               float head_ratio, int max_vertices, float* __restrict__ out,
                                               unsigned* counter) {
    const CaseDesc<T>& cd = cases[case_id];
    const CaseConsts<T>& k = cd.c;
    const int t = blockIdx.x * NT + threadIdx.x;
    const int a = t % nax, b = t / nax;
    const int i = a * step + step / 2, j = b * step + step / 2;
    bool have = t < nax * nay && i < k.nx && j < k.ny;
    float uu = 0.0f, vv = 0.0f, speed = 0.0f;
    if (have) {
        const size_t g = (size_t)j * k.nx + i;
        const T r = tmax(((const T*)cd.b.U[set][0])[g], (T)k.min_rho);
        uu = (float)(((const T*)cd.b.U[set][1])[g] / r);
        vv = (float)(((const T*)cd.b.U[set][2])[g] / r);
        speed = sqrtf(uu * uu + vv * vv);
        have = !(speed < 1e-6f * u_inf);
    }
    // wave-aggregated slot reservation: one atomic per wave, lanes in lane order
    const unsigned long long mask = __ballot(have);
    const int lane = threadIdx.x & 63;
    const unsigned before = (unsigned)__popcll(mask & ((1ull << lane) - 1ull));
    unsigned base = 0;
    if (lane == 0 && mask) base = atomicAdd(counter, 8u * (unsigned)__popcll(mask));
    base = (unsigned)__shfl((int)base, 0, 64);
    if (!have) return;
    const unsigned slot = base + 8u * before;
    if ((long long)slot + 8 > (long long)max_vertices) return;
    const float norm = fminf(speed / (u_inf * 1.5f), 1.0f);
    const float sx = (float)i / k.nx * 2.0f - 1.0f, sy = (float)j / k.ny * 2.0f - 1.0f;
    const float dx = uu / speed, dy = vv / speed;
    const float len = max_len * norm;
    const float ex = sx + dx * len, ey = sy + dy * len;
    const float hl = len * head_ratio;
    const float h1x = ex - hl * (dx * cos_a - dy * sin_a), h1y = ey - hl * (dx * sin_a + dy * cos_a);
    const float h2x = ex - hl * (dx * cos_a + dy * sin_a), h2y = ey - hl * (-dx * sin_a + dy * cos_a);
    // The reference reserves 8 vertices per arrow and writes the first 6, so the
    // fourth line it draws comes from whatever the buffer held; here it is a
    // zero-length line at the tip, which draws nothing.
    const float xy[8][2] = {{sx, sy}, {ex, ey}, {ex, ey}, {h1x, h1y}, {ex, ey}, {h2x, h2y}, {ex, ey}, {ex, ey}};
    float* o = out + (size_t)slot * 5;
    for (int q = 0; q < 8; q++) {
        o[q * 5 + 0] = xy[q][0];
        o[q * 5 + 1] = xy[q][1];
        o[q * 5 + 2] = 0.0f;
        o[q * 5 + 3] = 0.0f;
        o[q * 5 + 4] = 0.0f;
    }
}

// Statistics: max T of the primitives (mode 0) or max Mach |v|/(c+1e-10) (mode 1)
// (launchComputeMaxTemperature / launchComputeMaxMach, solver.cu:2782-2867).
template <class T>
__global__ __launch_bounds__(NT) void k_stat(const CaseDesc<T>* __restrict__ cases, int case_id, int set, int mode, unsigned* target) {
    __shared__ float s_red[NT / 64];
    const CaseDesc<T>& cd = cases[case_id];
    const CaseConsts<T>& k = cd.c;
    size_t n = (size_t)k.nx * k.ny;
    float m = 0.0f;
    for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < n; i += (size_t)gridDim.x * NT) {
        T v[5];
        for (int f = 0; f < 5; f++) v[f] = ((const T*)cd.b.U[set][f])[i];
        T u, vv, p, tt;
        primitives(k, v[0], v[1], v[2], v[3], v[4], u, vv, p, tt);
        if (mode == 0) m = fmaxf(m, (float)tt);
        else {
            T rr = v[0] + T(1e-10f);
            T c = ::sqrt((T)k.gamma * p / rr);
            m = fmaxf(m, (float)(::sqrt(u * u + vv * vv) / (c + T(1e-10f))));
        }
    }
    block_max_atomic<NT>(m, s_red, target);
}

// Domain totals of four planes (rho, rho_u, rho_v, E) of n cells in FP64:
// per-workgroup partials, then one workgroup sums them in a fixed order
// (deterministic for a given number of workgroups).
template <class T>
__global__ __launch_bounds__(NT) void k_cons_partial(const T* __restrict__ p0, const T* __restrict__ p1,
                                                     const T* __restrict__ p2, const T* __restrict__ p3, size_t n,
                                                     double* partial) {
    __shared__ double s[4][NT];
    const T* const pl[4] = {p0, p1, p2, p3};
    double acc[4] = {0, 0, 0, 0};
    for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < n; i += (size_t)gridDim.x * NT)
        for (int f = 0; f < 4; f++) acc[f] += (double)pl[f][i];
    for (int f = 0; f < 4; f++) s[f][threadIdx.x] = acc[f];
    __syncthreads();
    for (int off = NT / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            for (int f = 0; f < 4; f++) s[f][threadIdx.x] += s[f][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int f = 0; f < 4; f++) partial[blockIdx.x * 4 + f] = s[f][0];
}

// Pressure force on the obstacle, per unit depth.  computeFluxesKernel gives
// every face with a solid or ghost cell on either side the wall flux
// (0, p, 0, 0, 0) with p of the face's left (x) or lower (y) cell
// (solver.cu:1589-1598, 1673-1682).  The fluid cell on the other side of such a
// face loses that normal momentum (wall to its right/top) or gains it (wall to
// its left/bottom); the sum over all wall faces next to cells whose residual in
// that direction is applied (not on the domain edge, solver.cu:1744-1760) is the
// momentum the body takes from the fluid per unit time.  acc = (Fx, Fy, faces, 0).
template <class T>
__global__ __launch_bounds__(NT) void k_force_partial(const CaseDesc<T>* __restrict__ cases, int case_id, int set,
                                                      double* partial) {
    __shared__ double s[4][NT];
    const CaseDesc<T>& cd = cases[case_id];
    const CaseConsts<T>& k = cd.c;
    const int nx = k.nx, ny = k.ny;
    const size_t n = (size_t)nx * ny;
    const T* U0 = (const T*)cd.b.U[set][0];
    const T* U1 = (const T*)cd.b.U[set][1];
    const T* U2 = (const T*)cd.b.U[set][2];
    const T* U3 = (const T*)cd.b.U[set][3];
    const T* U4 = (const T*)cd.b.U[set][4];
    const uint8_t* ct = cd.b.ct;
    auto wall = [](uint8_t c) { return c == CT_SOLID || c == CT_GHOST; };
    double acc[4] = {0, 0, 0, 0};
    for (size_t g = (size_t)blockIdx.x * NT + threadIdx.x; g < n; g += (size_t)gridDim.x * NT) {
        const int i = (int)(g % nx), j = (int)(g / nx);
        const bool w = wall(ct[g]);
        int dir = 0;   // bit 0: x face i+1/2 is a wall face next to an updated cell; bit 1: y face j+1/2
        double sx = 0, sy = 0;
        if (i < nx - 1 && w != wall(ct[g + 1])) {
            const int fi = w ? i + 1 : i;
            if (fi != 0 && fi != nx - 1) { dir |= 1; sx = w ? -1.0 : 1.0; }
        }
        if (j < ny - 1 && w != wall(ct[g + nx])) {
            const int fj = w ? j + 1 : j;
            if (fj != 0 && fj != ny - 1) { dir |= 2; sy = w ? -1.0 : 1.0; }
        }
        if (dir) {
            T u, v, p;
            primitives_uvp<T, Precise<T>>(k, U0[g], U1[g], U2[g], U3[g], U4[g], u, v, p);
            acc[0] += sx * (double)p * (double)k.dy;
            acc[1] += sy * (double)p * (double)k.dx;
            acc[2] += (double)((dir & 1) + (dir >> 1));
        }
    }
    for (int f = 0; f < 4; f++) s[f][threadIdx.x] = acc[f];
    __syncthreads();
    for (int off = NT / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            for (int f = 0; f < 4; f++) s[f][threadIdx.x] += s[f][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int f = 0; f < 4; f++) partial[blockIdx.x * 4 + f] = s[f][0];
}

__global__ __launch_bounds__(64) void k_cons_final(const double* partial, int nblocks, double* out) {
    if (threadIdx.x < 4) {
        double a = 0;
        for (int b = 0; b < nblocks; b++) a += partial[b * 4 + threadIdx.x];
        out[threadIdx.x] = a;
    }
}

// ---------------------------------------------------------------------------
// Host-side launchers
// ---------------------------------------------------------------------------
// A launch, with start / stop events on the dispatch itself when the engine
// times its kernels (hipExtLaunchKernelGGL: the events bracket the kernel, not
// the queue gaps around it, as a kernel trace does).
template <typename F, typename... A>
static void go(F kernel, const KernelCall& kc, A... args) {
    if (kc.ev0) hipExtLaunchKernelGGL(kernel, kc.grid, dim3(NT), 0, kc.stream, kc.ev0, kc.ev1, 0, args...);
    else kernel<<<kc.grid, NT, 0, kc.stream>>>(args...);
}

template <class T, int TX, int TY>
struct Launch {
    using CD = const CaseDesc<T>*;
    // Stage 1 of the compensated variant is the KAHAN instantiation only for
    // its FP64 dt (the compensation itself is in the blend of stage 2).
    static void stage1(const KernelCall& kc) {
        if (kc.kahan && sizeof(T) == 4) go(k_stage<T, TX, TY, 0, true>, kc, (CD)kc.cases, kc.work, kc.count, kc.ri, kc.rv);
        else go(k_stage<T, TX, TY, 0, false>, kc, (CD)kc.cases, kc.work, kc.count, kc.ri, kc.rv);
    }
    // Stage 1 with the first viscous half step fused in (FP32, uncompensated).
    static bool stage1_fused(const KernelCall& kc) {
        if constexpr (fused_d1_fits<T, TX, TY>()) {
            if (kc.kahan) return false;
            go(k_stage<T, TX, TY, 2, false>, kc, (CD)kc.cases, kc.work, kc.count, kc.ri, kc.rv);
            return true;
        }
        return false;
    }
    static void stage2(const KernelCall& kc) {
        if (kc.kahan) go(k_stage<T, TX, TY, 1, true>, kc, (CD)kc.cases, kc.work, kc.count, kc.ri, kc.rv);
        else go(k_stage<T, TX, TY, 1, false>, kc, (CD)kc.cases, kc.work, kc.count, kc.ri, kc.rv);
    }
    static void diffuse(const KernelCall& kc, bool fin) {
        if (fin) go(k_diffuse<T, TX, TY, true>, kc, (CD)kc.cases, kc.work, kc.count, kc.rv);
        else go(k_diffuse<T, TX, TY, false>, kc, (CD)kc.cases, kc.work, kc.count, kc.rv);
    }
    static void bc_inplace(const KernelCall& kc) { go(k_bc_inplace<T, TX, TY>, kc, (CD)kc.cases, kc.work, kc.count, kc.ri); }
    static void maxima(const KernelCall& kc) { go(k_maxima<T, TX, TY>, kc, (CD)kc.cases, kc.work, kc.count, kc.ri); }
};

// Tile shapes compiled for the FP32 kernels (the autotuner chooses among them)
// and for the FP64 kernels.
#define FVM_TILE_TABLE(X) X(64, 4) X(64, 6) X(64, 7) X(64, 8) X(64, 12) X(64, 16) X(32, 2) X(32, 3) X(32, 4) X(32, 7) X(32, 8) X(32, 16) X(128, 4) X(128, 8) X(96, 8)
#define FVM_TILE_TABLE64(X) X(64, 8) X(32, 8)

template <class T>
static bool dispatch_tile(int tx, int ty, const KernelCall& kc);

#define FVM_TRY(A, B)                                                                              \
    if (tx == A && ty == B) {                                                                      \
        switch (kc.op) {                                                                           \
        case OP_STAGE1: Launch<T, A, B>::stage1(kc); break;                                        \
        case OP_STAGE1_FUSED: if (!Launch<T, A, B>::stage1_fused(kc)) return false; break;         \
        case OP_STAGE2: Launch<T, A, B>::stage2(kc); break;                                        \
        case OP_DIFFUSE: Launch<T, A, B>::diffuse(kc, false); break;                               \
        case OP_DIFFUSE_FINAL: Launch<T, A, B>::diffuse(kc, true); break;                          \
        case OP_BC_INPLACE: Launch<T, A, B>::bc_inplace(kc); break;                                \
        case OP_MAXIMA: Launch<T, A, B>::maxima(kc); break;                                        \
        default: return false;                                                                     \
        }                                                                                          \
        return true;                                                                               \
    }

template <>
bool dispatch_tile<float>(int tx, int ty, const KernelCall& kc) {
    using T = float;
    FVM_TILE_TABLE(FVM_TRY)
    return false;
}

template <>
bool dispatch_tile<double>(int tx, int ty, const KernelCall& kc) {
    using T = double;
    FVM_TILE_TABLE64(FVM_TRY)
    return false;
}
#undef FVM_TRY


bool launch_tiled(bool fp64, int tx, int ty, const KernelCall& kc) {
    return fp64 ? dispatch_tile<double>(tx, ty, kc) : dispatch_tile<float>(tx, ty, kc);
}

// Workgroups per CU that the occupancy calculator admits for a step kernel.
template <typename F>
static int resident(F kernel) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, NT, 0) != hipSuccess) return 0;
    return n;
}

template <class T, int TX, int TY>
static void tiled_residency(bool kahan, bool fused, int* stage, int* diffuse) {
    int s = kahan ? std::min(resident(k_stage<T, TX, TY, 0, true>), resident(k_stage<T, TX, TY, 1, true>))
                  : std::min(resident(k_stage<T, TX, TY, 0, false>), resident(k_stage<T, TX, TY, 1, false>));
    if constexpr (fused_d1_fits<T, TX, TY>())
        if (fused && !kahan) s = std::min(s, resident(k_stage<T, TX, TY, 2, false>));
    *stage = s;
    *diffuse = std::min(resident(k_diffuse<T, TX, TY, false>), resident(k_diffuse<T, TX, TY, true>));
}

bool tiled_resident_per_cu(bool fp64, int tx, int ty, bool kahan, bool fused, int* stage, int* diffuse) {
    *stage = *diffuse = 0;
#define FVM_RES(A, B) if (tx == A && ty == B) { tiled_residency<T, A, B>(kahan, fused, stage, diffuse); return true; }
    if (fp64) {
        using T = double;
        FVM_TILE_TABLE64(FVM_RES)
    } else {
        using T = float;
        FVM_TILE_TABLE(FVM_RES)
    }
#undef FVM_RES
    return false;
}

bool fused_d1_supported(int tx, int ty, bool fp64) {
    if (fp64) return false;
#define FVM_FITS(A, B) if (tx == A && ty == B) return fused_d1_fits<float, A, B>();
    FVM_TILE_TABLE(FVM_FITS)
#undef FVM_FITS
    return false;
}

// Smallest tile width and height among the compiled shapes (FP32 and FP64):
// ceil(nx / min width) * ceil(ny / min height) bounds any shape's tile count.
void tile_min_dims(int* tx, int* ty) {
    int mx = 1 << 30, my = 1 << 30;
#define FVM_MIN(A, B) mx = A < mx ? A : mx; my = B < my ? B : my;
    FVM_TILE_TABLE(FVM_MIN)
    FVM_TILE_TABLE64(FVM_MIN)
#undef FVM_MIN
    *tx = mx;
    *ty = my;
}

bool tile_supported(int tx, int ty, bool fp64) {
#define FVM_HAS(A, B) if (tx == A && ty == B) return true;
    if (fp64) {
        FVM_TILE_TABLE64(FVM_HAS)
    } else {
        FVM_TILE_TABLE(FVM_HAS)
    }
#undef FVM_HAS
    return false;
}

void launch_ghost(bool fp64, bool dt64, bool final_, dim3 grid, hipStream_t s, const void* c, const WorkItem* items,
                  const int* n, Roles ri, Roles rv, int gblocks, const WorkItem* case_list, const int* ncases,
                  int single_case, hipEvent_t ev0, hipEvent_t ev1) {
    if (!final_) gblocks = (int)grid.x;
    const CaseDesc<double>* c64 = (const CaseDesc<double>*)c;
    const CaseDesc<float>* c32 = (const CaseDesc<float>*)c;
    KernelCall kc{};
    kc.grid = grid;
    kc.stream = s;
    kc.ev0 = ev0;
    kc.ev1 = ev1;
    if (fp64) {
        if (final_) go(k_ghost<double, true, true>, kc, c64, items, n, ri, rv, gblocks, case_list, ncases, single_case);
        else go(k_ghost<double, false, true>, kc, c64, items, n, ri, rv, gblocks, case_list, ncases, single_case);
    } else if (dt64) {
        if (final_) go(k_ghost<float, true, true>, kc, c32, items, n, ri, rv, gblocks, case_list, ncases, single_case);
        else go(k_ghost<float, false, true>, kc, c32, items, n, ri, rv, gblocks, case_list, ncases, single_case);
    } else {
        if (final_) go(k_ghost<float, true, false>, kc, c32, items, n, ri, rv, gblocks, case_list, ncases, single_case);
        else go(k_ghost<float, false, false>, kc, c32, items, n, ri, rv, gblocks, case_list, ncases, single_case);
    }
}

void launch_ghost_prep(bool fp64, hipStream_t s, const void* c, const WorkItem* items, int n) {
    if (n <= 0) return;
    dim3 g((unsigned)((n + NT - 1) / NT));
    if (fp64) k_ghost_prep<double><<<g, NT, 0, s>>>((const CaseDesc<double>*)c, items, n);
    else k_ghost_prep<float><<<g, NT, 0, s>>>((const CaseDesc<float>*)c, items, n);
}

void launch_mark(bool fp64, hipStream_t s, const void* c, int case_id, size_t ncells, int fixup, int fix_set) {
    dim3 g((unsigned)std::min<size_t>((ncells + NT - 1) / NT, 65535));
    if (fp64) k_mark<double><<<g, NT, 0, s>>>((const CaseDesc<double>*)c, case_id, fixup, fix_set);
    else k_mark<float><<<g, NT, 0, s>>>((const CaseDesc<float>*)c, case_id, fixup, fix_set);
}

void launch_init(bool fp64, hipStream_t s, const void* c, int case_id, size_t ncells, int set_a, int set_b) {
    dim3 g((unsigned)std::min<size_t>((ncells + NT - 1) / NT, 65535));
    if (fp64) k_init<double><<<g, NT, 0, s>>>((const CaseDesc<double>*)c, case_id, set_a, set_b);
    else k_init<float><<<g, NT, 0, s>>>((const CaseDesc<float>*)c, case_id, set_a, set_b);
}

void launch_copy_set(bool fp64, hipStream_t s, const void* c, int case_id, size_t ncells, int src, int dst) {
    dim3 g((unsigned)std::min<size_t>((ncells + NT - 1) / NT, 65535));
    if (fp64) k_copy_set<double><<<g, NT, 0, s>>>((const CaseDesc<double>*)c, case_id, src, dst);
    else k_copy_set<float><<<g, NT, 0, s>>>((const CaseDesc<float>*)c, case_id, src, dst);
}

void launch_zero_u32(hipStream_t s, unsigned* p, int n) {
    k_zero_u32<<<(n + 255) / 256, 256, 0, s>>>(p, n);
}

void launch_field(bool fp64, hipStream_t s, const void* c, int case_id, size_t ncells, int set, int field, float* out) {
    dim3 g((unsigned)std::min<size_t>((ncells + NT - 1) / NT, 65535));
    if (fp64) k_field<double><<<g, NT, 0, s>>>((const CaseDesc<double>*)c, case_id, set, field, out);
    else k_field<float><<<g, NT, 0, s>>>((const CaseDesc<float>*)c, case_id, set, field, out);
}

void launch_arrows(bool fp64, hipStream_t s, const void* c, int case_id, int set, int step, int nax, int nay,
                   float u_inf, float max_len, float cos_a, float sin_a, float head_ratio, int max_vertices, float* out,
                   unsigned* counter) {
    const long long n = (long long)nax * nay;
    if (n <= 0) return;
    dim3 g((unsigned)((n + NT - 1) / NT));
    if (fp64)
        k_arrows<double><<<g, NT, 0, s>>>((const CaseDesc<double>*)c, case_id, set, step, nax, nay, u_inf, max_len, cos_a,
                                          sin_a, head_ratio, max_vertices, out, counter);
    else
        k_arrows<float><<<g, NT, 0, s>>>((const CaseDesc<float>*)c, case_id, set, step, nax, nay, u_inf, max_len, cos_a,
                                         sin_a, head_ratio, max_vertices, out, counter);
}

void launch_stat(bool fp64, hipStream_t s, const void* c, int case_id, size_t ncells, int set, int mode, unsigned* target) {
    dim3 g((unsigned)std::min<size_t>((ncells + NT - 1) / NT, 2048));
    if (fp64) k_stat<double><<<g, NT, 0, s>>>((const CaseDesc<double>*)c, case_id, set, mode, target);
    else k_stat<float><<<g, NT, 0, s>>>((const CaseDesc<float>*)c, case_id, set, mode, target);
}

void launch_cons(bool fp64, hipStream_t s, const void* const planes[4], size_t n, double* partial, int nblocks,
                 double* out) {
    if (fp64)
        k_cons_partial<double><<<nblocks, NT, 0, s>>>((const double*)planes[0], (const double*)planes[1],
                                                       (const double*)planes[2], (const double*)planes[3], n, partial);
    else
        k_cons_partial<float><<<nblocks, NT, 0, s>>>((const float*)planes[0], (const float*)planes[1],
                                                      (const float*)planes[2], (const float*)planes[3], n, partial);
    k_cons_final<<<1, 64, 0, s>>>(partial, nblocks, out);
}

void launch_forces(bool fp64, hipStream_t s, const void* c, int case_id, int set, double* partial, int nblocks, double* out) {
    if (fp64) k_force_partial<double><<<nblocks, NT, 0, s>>>((const CaseDesc<double>*)c, case_id, set, partial);
    else k_force_partial<float><<<nblocks, NT, 0, s>>>((const CaseDesc<float>*)c, case_id, set, partial);
    k_cons_final<<<1, 64, 0, s>>>(partial, nblocks, out);
}

}  // namespace fvm
