// fvm_device.hpp -- device data layout and per-cell physics of the FVM step.
//
// Layout in HBM: every case owns SoA planes (rho, rho_u, rho_v, E, rho_e) of
// nx*ny values, row-major idx = j*nx + i (src/solver.cu:547, 923), in up to
// four buffer sets (A = U^n, B = U*, C = U** / blend, D = U after the first
// viscous half step), one uint8 cell-tag plane, one FP32 signed-distance plane
// and a compact list of ghost cells.  Per-case constants sit in a CaseDesc
// table that kernels read with uniform (scalar) loads.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fvm {

enum : uint8_t { CT_FLUID = 0, CT_SOLID = 1, CT_GHOST = 2, CT_INFLOW = 3, CT_OUTFLOW = 4 };

// Obstacle geometry (transcendentals evaluated once on the host; FP32 always,
// the cell mask is compared bit for bit with the oracle).
struct Geom {
    int shape;
    float obs_x, obs_y, obs_r, cosR, sinR, cosW, sinW, band, dx, dy;
    float star[20];
};

// Per-case run-time constants (host-prepared from fvm_params in the precision
// of the state, T = float or double).
template <class T>
struct CaseConsts {
    int nx, ny;
    int visc, adiabatic, flux, limiter, wall_mode, dt_mode;
    int noise_field;
    uint32_t noise_seed;
    float noise_amp;
    float cfl, cfl_visc, min_dx, host_dt;
    T dx, dy, inv_dx, inv_dy, inv_2dx, inv_2dy;
    T gamma, gm1, inv_gm1, R, inv_R, cp, prandtl, two_over_gm1, gm1_over_4, inv_gamma, cs_fac;
    T min_rho, min_p, max_T, min_T, low_rho, vmax, vghost, slip_ref;
    T mu_ref, t_ref, s_su;
    T rho_inf, u_inf, v_inf, p_inf, e_inf, E_inf, T_wall;
    T p_gauge;                // subtracted from every normal-momentum face flux (FP32)
    T s_far, c_far;           // far-field entropy p rho^-gamma and sound speed
    Geom geo;
};

// Static part of a ghost cell's fill, computed once per geometry: the unit
// SDF normal and its candidate neighbours (codes 0..7 of the 8-neighbourhood
// in the reference's scan order, solver.cu:2072) sorted by descending
// alignment with the normal, ties kept in scan order; only neighbours inside
// the domain with sdf > 0 and alignment > -1 are listed (3 bits each, count in
// bits 24..27).  At run time the first listed neighbour whose density passes
// the check is the one the reference's scan selects.
template <class T>
struct GhostInfo {
    int idx;
    uint32_t order;
    T nx, ny;
};

struct CaseBufs {
    void* U[4][5];        // buffer set x field; float* or double*
    float* K[5];          // compensation carries (FP32 + Kahan variant)
    uint8_t* ct;
    float* sdf;
    void* ghosts;         // GhostInfo<T>[n_ghosts]
    int* gmap;            // ghost-table index of each cell of the ghost bounding box, -1 elsewhere
    int gbox[4];          // bounding box of the ghost cells: i0, j0, width, height
    unsigned* slots;      // [2 slots][2]: bits of max wave speed, max mu/rho (FP32 maxima)
    float* tile_max;      // [2][tiles]: per-tile max wave speed, max mu/rho of the step's final kernel
    double* t;            // simulated time
    long long* steps;     // steps taken
};

// The FP64 CFL maxima sit after the fields the stage kernels load, so that
// those keep the offsets (and scalar-load grouping) they have without them.
template <class T>
struct CaseDesc {
    CaseConsts<T> c;
    CaseBufs b;
    int tiles_x, tiles_y;
    // 1: the CFL maxima and the dt formula in FP64 (FP64 state, and the
    // compensated FP32 variant); 0: FP32 as in the reference
    int dt64;
    int pad;
    double* slots64;      // CaseBufs::slots in FP64 (dt64 cases; atomics on the 64-bit patterns)
    double* tile_max64;   // CaseBufs::tile_max in FP64 (dt64 cases)
};

// One work item: a tile of a case (stage / diffusion kernels) or a ghost cell.
struct WorkItem {
    int case_id;
    int index;
};

// Buffer roles and dt slots of one launch.
struct Roles {
    int in, out, blend, pre0, pre1;  // buffer-set indices (-1 = unused)
    int slot_r, slot_w;              // dt slot read / written (0,1)
    int zero_slot;                   // zero slot_w of every case before the writers run
    int count_step;                  // advance the case's time and step counter
};

// ------------------------------------------------------------------ math
template <class T> struct Num;

template <> struct Num<float> {
    // Throughput forms for the flow arithmetic: v_rcp_f32 / v_sqrt_f32 (about
    // 1 ulp), like the reference's --use_fast_math build (CMakeLists.txt:279).
    static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    static __device__ __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    static __device__ __forceinline__ float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
    static __device__ __forceinline__ float rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
    static __device__ __forceinline__ float exp(float x) { return __expf(x); }
    static __device__ __forceinline__ float log(float x) { return __logf(x); }
    static __device__ __forceinline__ bool finite(float x) { return __builtin_isfinite(x); }
};

template <> struct Num<double> {
    static __device__ __forceinline__ double rcp(double x) { return 1.0 / x; }
    static __device__ __forceinline__ double div(double a, double b) { return a / b; }
    static __device__ __forceinline__ double sqrt(double x) { return ::sqrt(x); }
    static __device__ __forceinline__ double rsqrt(double x) { return 1.0 / ::sqrt(x); }
    static __device__ __forceinline__ double exp(double x) { return ::exp(x); }
    static __device__ __forceinline__ double log(double x) { return ::log(x); }
    static __device__ __forceinline__ bool finite(double x) { return __builtin_isfinite(x); }
};

// Accurate forms (IEEE division and square root, ocml exp/log) for the
// O(perimeter) boundary work, where the characteristic reconstruction
// v = (R3 + R4)/2 amplifies sound-speed errors.
template <class T> struct Precise {
    static __device__ __forceinline__ T rcp(T x) { return T(1) / x; }
    static __device__ __forceinline__ T div(T a, T b) { return a / b; }
    static __device__ __forceinline__ T sqrt(T x) { return ::sqrt(x); }
    static __device__ __forceinline__ T exp(T x) { return ::exp(x); }
    static __device__ __forceinline__ T log(T x) { return ::log(x); }
    static __device__ __forceinline__ bool finite(T x) { return __builtin_isfinite(x); }
};

template <class T> __device__ __forceinline__ T tmax(T a, T b) { return a > b ? a : (b == b ? b : a); }
template <> __device__ __forceinline__ float tmax<float>(float a, float b) { return fmaxf(a, b); }
template <> __device__ __forceinline__ double tmax<double>(double a, double b) { return fmax(a, b); }
template <class T> __device__ __forceinline__ T tmin(T a, T b);
template <> __device__ __forceinline__ float tmin<float>(float a, float b) { return fminf(a, b); }
template <> __device__ __forceinline__ double tmin<double>(double a, double b) { return fmin(a, b); }

// One field plane (or any per-cell array) in HBM, indexed by cell number.
// The case table in device memory holds the buffer pointers, so the compiler
// sees them as generic pointers and would emit flat loads and stores; flat
// accesses count on the LDS counter as well (a wait for an LDS read then also
// waits for outstanding HBM loads) and each needs a 64-bit address.  Plane casts
// the pointer to the global address space (the address is unchanged) and
// indexes with a 32-bit byte offset, which gives global_load / global_store
// with the plane's base in SGPRs and one offset VGPR shared by every field of a
// cell.  The byte offset must fit in 32 bits: the engine rejects grids whose
// plane exceeds 4 GiB.
template <class T> using gmem = __attribute__((address_space(1))) T;

template <class T>
struct Plane {
    gmem<char>* base;
    __device__ __forceinline__ explicit Plane(const void* p) : base((gmem<char>*)p) {}
    __device__ __forceinline__ T operator[](unsigned i) const { return *(const gmem<T>*)(base + i * (unsigned)sizeof(T)); }
    __device__ __forceinline__ void put(unsigned i, T v) const { *(gmem<T>*)(base + i * (unsigned)sizeof(T)) = v; }
};

// Plane index of cell (i, j) on row stride nx, as one 24-bit multiply-add
// (grid dimensions are below 2^24 and a plane holds at most 2^30 cells).  The
// 64-bit multiply-add the compiler picks for j * nx + i reads a second
// register as the high half of its addend; when that register is the target
// of a load still in flight, every outstanding load is waited for first.
__device__ __forceinline__ unsigned cell_index(int i, int j, int nx) {
    return __umul24((unsigned)j, (unsigned)nx) + (unsigned)i;
}

template <class T> __device__ __forceinline__ T tabs(T a) { return a < T(0) ? -a : a; }
template <> __device__ __forceinline__ float tabs<float>(float a) { return fabsf(a); }
template <> __device__ __forceinline__ double tabs<double>(double a) { return fabs(a); }

// ------------------------------------------------------------------ physics
// computePrimitivesKernelInline, src/solver.cu:862-894
template <class T>
__device__ __forceinline__ void primitives(const CaseConsts<T>& k, T rho, T ru, T rv, T E, T re,
                                           T& u, T& v, T& p, T& Tt) {
    T r = tmax(rho, (T)k.min_rho);
    T ir = Num<T>::rcp(r);
    u = ru * ir;
    v = rv * ir;
    T ke = T(0.5) * r * (u * u + v * v);
    // dual-energy switch eta = rho_e / E < 1e-3 (E > 0), as a multiply-compare
    T e = (E > T(0) && re < T(0.001f) * E) ? re : E - ke;
    p = tmax((T)k.gm1 * e, (T)k.min_p);
    T tt = p * ir * (T)k.inv_R;
    tt = tmin(tt, (T)k.max_T);
    Tt = tmax(tt, (T)k.min_T);
}

// Pressure only (fluxes need rho, u, v, p; T is not used there).
template <class T, class M = Num<T>>
__device__ __forceinline__ void primitives_uvp(const CaseConsts<T>& k, T rho, T ru, T rv, T E, T re,
                                               T& u, T& v, T& p) {
    T r = tmax(rho, (T)k.min_rho);
    u = M::div(ru, r);
    v = M::div(rv, r);
    T ke = T(0.5) * r * (u * u + v * v);
    T e = (E > T(0) && re < T(0.001f) * E) ? re : E - ke;
    p = tmax((T)k.gm1 * e, (T)k.min_p);
}

// Limited MUSCL slope: minmod (solver.cu:110-115) in full, van Leer 2ab/(a+b)
// halved, ab/(a+b), so that it needs no doubling.  Face values are
// q +- w * slope with w = 1/2 for minmod and 1 for van Leer (0 for a
// first-order face), i.e. q +- (full slope)/2 either way; the halving is exact.
// Branch-free.  Minmod in FP32 is the median of (dL, dR, 0): the operand of
// smaller magnitude when the signs agree, else zero; adding +0 turns a -0
// operand into the +0 the reference's form gives.  Otherwise both candidates
// are formed and one selected, minmod as the operand of smaller magnitude
// (for dL dR > 0 that is copysign(min(|dL|, |dR|), dL)).
template <class T>
__device__ __forceinline__ T limited_slope(int limiter, T dL, T dR) {
    if constexpr (sizeof(T) == 4) {
        if (limiter != 1) return __builtin_amdgcn_fmed3f(dL, dR, 0.0f) + 0.0f;
    }
    const T prod = dL * dR;
    const T s = limiter == 1 ? Num<T>::div(prod, dL + dR) : (tabs(dL) < tabs(dR) ? dL : dR);
    return prod > T(0) ? s : T(0);
}

// Solid or ghost tag (CT_SOLID = 1, CT_GHOST = 2) in one unsigned compare.
__device__ __forceinline__ bool wall_tag(uint8_t c) { return (unsigned)c - 1u <= 1u; }

// Physical flux in face-normal form (computeFluxX, solver.cu:1005-1014).
// pg: gauge pressure removed from the momentum flux before rounding.
template <class T>
__device__ __forceinline__ void phys_flux(T rho, T un, T ut, T p, T E, T f[4], T pg = T(0)) {
    T m = rho * un;
    f[0] = m;
    f[1] = m * un + (p - pg);
    f[2] = m * ut;
    f[3] = (E + p) * un;
}

// HLLC with Roe-Davis wave speeds and the Harten-type push-out
// (hllcFluxX solver.cu:1112-1223; hllcFluxY 1315-1430 is the same with the
// velocity components exchanged).  Branch-free selection of the side.
template <class T>
__device__ __forceinline__ void hllc(const CaseConsts<T>& k, T rL, T uL, T vL, T pL, T EL,
                                     T rR, T uR, T vR, T pR, T ER, T f[4]) {
    // one reciprocal square root per side gives sqrt(rho) = rho rsqrt(rho);
    // the Roe weight of H = (E + p)/rho is then (E + p) rsqrt(rho)
    T rsL = Num<T>::rsqrt(rL), rsR = Num<T>::rsqrt(rR);
    T sL = rL * rsL, sR = rR * rsR;
    T cL = Num<T>::sqrt((T)k.gamma * pL) * rsL;
    T cR = Num<T>::sqrt((T)k.gamma * pR) * rsR;
    T id = Num<T>::rcp(sL + sR + T(1e-10f));
    T u_roe = (sL * uL + sR * uR) * id;
    T v_roe = (sL * vL + sR * vR) * id;
    T H_roe = ((EL + pL) * rsL + (ER + pR) * rsR) * id;
    T c2 = (T)k.gm1 * (H_roe - T(0.5) * (u_roe * u_roe + v_roe * v_roe));
    T c_roe = (c2 > T(0)) ? Num<T>::sqrt(c2) : T(0.5) * (cL + cR);
    T SL = tmin(uL - cL, u_roe - c_roe);
    T SR = tmax(uR + cR, u_roe + c_roe);
    T delta = T(0.1f) * tmax(cL, cR);
    if (tabs(SL) < delta) SL = -delta;
    if (tabs(SR) < delta) SR = delta;
    T dsl = SL - uL, dsr = SR - uR;
    T den = rL * dsl - rR * dsr;
    T safe = copysign(tmax(tabs(den), T(1e-10f)), den);
    T S_star = Num<T>::div(pR - pL + rL * uL * dsl - rR * uR * dsr, safe);
    // side K: left if SL >= 0, right if SR <= 0, else by the sign of S*
    bool useL = (SL >= T(0)) || (!(SR <= T(0)) && S_star >= T(0));
    bool star = !(SL >= T(0)) && !(SR <= T(0));
    T rK = useL ? rL : rR, uK = useL ? uL : uR, vK = useL ? vL : vR;
    T pK = useL ? pL : pR, EK = useL ? EL : ER, SK = useL ? SL : SR;
    phys_flux(rK, uK, vK, pK, EK, f, k.p_gauge);
    if (star) {
        T dsk = SK - uK;
        T cd = SK - S_star;
        if (tabs(cd) < T(1e-10f)) cd = (cd >= T(0)) ? T(1e-10f) : T(-1e-10f);
        T icd = Num<T>::rcp(cd);
        T rs = tmax(rK * dsk * icd, (T)k.min_rho);
        T ps = tmax(pK + rK * dsk * (S_star - uK), (T)k.min_p);
        T Es = (EK * dsk + ps * S_star - pK * uK) * icd;
        T dm = SK * (rs - rK);
        f[0] += dm;
        f[1] += SK * (rs * S_star - rK * uK);
        f[2] += dm * vK;
        f[3] += SK * (Es - EK);
    }
}

// AUSM+ (Liou 1996): split Mach/pressure polynomials, beta = 1/8, alpha = 3/16,
// interface sound speed from the critical speed c*^2 = 2(gamma-1)/(gamma+1) H.
template <class T>
__device__ __forceinline__ void ausm_plus(const CaseConsts<T>& k, T rL, T uL, T vL, T pL, T EL,
                                          T rR, T uR, T vR, T pR, T ER, T f[4]) {
    const T beta = T(0.125), alpha = T(0.1875);
    T HL = (EL + pL) * Num<T>::rcp(rL);
    T HR = (ER + pR) * Num<T>::rcp(rR);
    T csL = Num<T>::sqrt((T)k.cs_fac * HL);
    T csR = Num<T>::sqrt((T)k.cs_fac * HR);
    T chL = Num<T>::div(csL * csL, tmax(csL, uL));
    T chR = Num<T>::div(csR * csR, tmax(csR, -uR));
    T c12 = tmin(chL, chR);
    T ic = Num<T>::rcp(c12);
    T ML = uL * ic, MR = uR * ic;
    T Mp, Pp, Mm, Pm;
    if (tabs(ML) < T(1)) {
        T a = ML + T(1), b = ML * ML - T(1);
        Mp = T(0.25) * a * a + beta * b * b;
        Pp = T(0.25) * a * a * (T(2) - ML) + alpha * ML * b * b;
    } else {
        Mp = T(0.5) * (ML + tabs(ML));
        Pp = ML > T(0) ? T(1) : T(0);
    }
    if (tabs(MR) < T(1)) {
        T a = MR - T(1), b = MR * MR - T(1);
        Mm = -T(0.25) * a * a - beta * b * b;
        Pm = T(0.25) * a * a * (T(2) + MR) - alpha * MR * b * b;
    } else {
        Mm = T(0.5) * (MR - tabs(MR));
        Pm = MR < T(0) ? T(1) : T(0);
    }
    T m12 = Mp + Mm;
    // p12 - p_gauge, rearranged so the large common part cancels before rounding
    T p12 = Pp * (pL - k.p_gauge) + Pm * (pR - k.p_gauge) + k.p_gauge * ((Pp + Pm) - T(1));
    T mp = T(0.5) * (m12 + tabs(m12)) * c12;
    T mm = T(0.5) * (m12 - tabs(m12)) * c12;
    T aL = mp * rL, aR = mm * rR;
    f[0] = aL + aR;
    f[1] = aL * uL + aR * uR + p12;
    f[2] = aL * vL + aR * vR;
    f[3] = aL * HL + aR * HR;
}

// The face part of computeFluxesKernel (solver.cu:1503-1598 / 1603-1683) in
// face-normal form.  Cells m = 0..3 are i-1, i, i+1, i+2 along the normal.
// out = (mass, normal momentum, tangential momentum, energy, internal energy).
template <class T>
__device__ __forceinline__ void face_flux(const CaseConsts<T>& k, const T rho[4], const T un[4], const T ut[4],
                                          const T p[4], const uint8_t ct[4], T out[5]) {
    // The FP32 path carries the normal-momentum flux relative to a constant
    // gauge pressure (the freestream pressure): a constant added to every face
    // flux cancels in the divergence, and the face pressures ~p_inf no longer
    // swamp the small momentum changes of low-Mach flow in FP32 rounding.
    if (wall_tag(ct[1]) || wall_tag(ct[2])) {
        out[0] = T(0); out[1] = p[1] - k.p_gauge; out[2] = T(0); out[3] = T(0); out[4] = T(0);
        return;
    }
    // second order unless a +-2 neighbour is a wall or a face cell is near vacuum
    bool first = wall_tag(ct[0]) || wall_tag(ct[3]) || rho[1] < (T)k.low_rho || rho[2] < (T)k.low_rho;
    int lim = k.limiter;
    const T h = first ? T(0) : (lim == 1 ? T(1) : T(0.5));
    T rL = rho[1] + h * limited_slope(lim, rho[1] - rho[0], rho[2] - rho[1]);
    T uL = un[1] + h * limited_slope(lim, un[1] - un[0], un[2] - un[1]);
    T vL = ut[1] + h * limited_slope(lim, ut[1] - ut[0], ut[2] - ut[1]);
    T pL = p[1] + h * limited_slope(lim, p[1] - p[0], p[2] - p[1]);
    T rR = rho[2] - h * limited_slope(lim, rho[2] - rho[1], rho[3] - rho[2]);
    T uR = un[2] - h * limited_slope(lim, un[2] - un[1], un[3] - un[2]);
    T vR = ut[2] - h * limited_slope(lim, ut[2] - ut[1], ut[3] - ut[2]);
    T pR = p[2] - h * limited_slope(lim, p[2] - p[1], p[3] - p[2]);
    rL = tmax(rL, (T)k.min_rho);
    pL = tmax(pL, (T)k.min_p);
    rR = tmax(rR, (T)k.min_rho);
    pR = tmax(pR, (T)k.min_p);
    T reL = pL * (T)k.inv_gm1, reR = pR * (T)k.inv_gm1;
    T EL = reL + T(0.5) * rL * (uL * uL + vL * vL);
    T ER = reR + T(0.5) * rR * (uR * uR + vR * vR);
    T f[4];
    if (k.flux == 1) ausm_plus(k, rL, uL, vL, pL, EL, rR, uR, vR, pR, ER, f);
    else hllc(k, rL, uL, vL, pL, EL, rR, uR, vR, pR, ER, f);
    out[0] = f[0]; out[1] = f[1]; out[2] = f[2]; out[3] = f[3];
    out[4] = (uL + uR >= T(0)) ? reL * uL : reR * uR;
}

// Characteristic (non-reflecting) boundary state.  The reference builds the
// four Riemann invariants R1 = p rho^-gamma, R2 = u_t, R3/R4 = u_n +- 2c/(gamma-1)
// of the local and of the far-field state (solver.cu:680-841), takes each from
// the local state when its characteristic leaves the domain and from the far
// field otherwise (solver.cu:1925-2030), and inverts them.  The invariants are
// kept split into their (u_n, c) parts here, so that u_n = (R3 + R4)/2 and
// c = (R3 - R4)(gamma-1)/4 are formed without cancelling two terms of size
// 2c/(gamma-1) (in FP32 that cancellation alone leaves u_n with an absolute
// error near 1e-7 c); the result is the same in exact arithmetic.
template <class T>
__device__ __forceinline__ void nrbc_state(const CaseConsts<T>& k, T un_f, T ut_f, T rho_in, T un, T ut, T p,
                                           bool o1, bool o3, bool o4, T& rho_o, T& un_o, T& ut_o, T& p_o) {
    typedef Precise<T> M;
    T c_l = M::sqrt((T)k.gamma * p / rho_in);
    T s = o1 ? M::exp(-(T)k.gamma * M::log(rho_in)) * p : (T)k.s_far;
    ut_o = o1 ? ut : ut_f;
    T un3 = o3 ? un : un_f, c3 = o3 ? c_l : (T)k.c_far;
    T un4 = o4 ? un : un_f, c4 = o4 ? c_l : (T)k.c_far;
    un_o = (un3 + un4) * T(0.5) + (c3 - c4) * (T)k.inv_gm1;
    T c = (un3 - un4) * (T)k.gm1_over_4 + (c3 + c4) * T(0.5);
    T c2 = c * c;
    rho_o = M::exp((T)k.inv_gm1 * M::log(c2 * (T)k.inv_gamma / s));
    p_o = rho_o * c2 * (T)k.inv_gamma;
}

// Point-wise boundary conditions of applyBoundaryConditionsKernel
// (solver.cu:1913-2042) for a non-ghost cell; v = (rho, rho_u, rho_v, E, rho_e).
// Returns false for interior fluid cells (nothing to do).
template <class T>
__device__ __forceinline__ bool bc_pointwise(const CaseConsts<T>& k, int i, int j, uint8_t c, T v[5]) {
    if (c == CT_INFLOW || i == 0) {
        v[0] = (T)k.rho_inf; v[1] = (T)k.rho_inf * (T)k.u_inf; v[2] = (T)k.rho_inf * (T)k.v_inf;
        v[3] = (T)k.E_inf; v[4] = (T)k.e_inf;
        return true;
    }
    if (c == CT_OUTFLOW || i == k.nx - 1) {
        T rloc = tmax(v[0], (T)k.min_rho);
        T u, vv, p;
        primitives_uvp<T, Precise<T>>(k, v[0], v[1], v[2], v[3], v[4], u, vv, p);
        T cl = Precise<T>::sqrt((T)k.gamma * p / rloc);
        bool o1 = u > T(0), o3 = u + cl > T(0), o4 = u - cl > T(0);
        T rho, un, ut, pp;
        nrbc_state(k, (T)k.u_inf, (T)k.v_inf, v[0], u, vv, p, o1, o3, o4, rho, un, ut, pp);
        v[0] = rho; v[1] = rho * un; v[2] = rho * ut;
        v[4] = pp * (T)k.inv_gm1;
        v[3] = v[4] + T(0.5) * rho * (un * un + ut * ut);
        return true;
    }
    bool touched = false;
    for (int side = 0; side < 2; side++) {
        if ((side == 0 && j != 0) || (side == 1 && j != k.ny - 1)) continue;
        touched = true;
        T rloc = tmax(v[0], (T)k.min_rho);
        T u, vv, p;
        primitives_uvp<T, Precise<T>>(k, rloc, v[1], v[2], v[3], v[4], u, vv, p);
        T cl = Precise<T>::sqrt((T)k.gamma * p / rloc);
        bool o1, o3, o4;
        if (side == 0) { o1 = vv < T(0); o3 = vv + cl < T(0); o4 = vv - cl < T(0); }
        else { o1 = vv > T(0); o3 = vv + cl > T(0); o4 = vv - cl > T(0); }
        T rho, un, ut, pp;
        nrbc_state(k, (T)k.v_inf, (T)k.u_inf, rloc, vv, u, p, o1, o3, o4, rho, un, ut, pp);
        v[0] = rho; v[1] = rho * ut; v[2] = rho * un;
        v[4] = pp * (T)k.inv_gm1;
        v[3] = v[4] + T(0.5) * rho * (ut * ut + un * un);
    }
    if (c == CT_SOLID) {
        v[0] = (T)k.rho_inf; v[1] = T(0); v[2] = T(0); v[3] = (T)k.e_inf; v[4] = (T)k.e_inf;
        return true;
    }
    return touched;
}

// Sutherland viscosity (solver.cu:2317-2319).  i_tref = rcp(t_ref), formed
// once per tile by the kernels that evaluate it per cell (the compiler does
// not hoist the reciprocal of the constant out of their loops).
template <class T>
__device__ __forceinline__ T sutherland(const CaseConsts<T>& k, T Tv, T i_tref) {
    T tr = Tv * i_tref;
    return (T)k.mu_ref * (tr * Num<T>::sqrt(tr)) * ((T)k.t_ref + (T)k.s_su) * Num<T>::rcp(Tv + (T)k.s_su);
}
template <class T>
__device__ __forceinline__ T sutherland(const CaseConsts<T>& k, T Tv) {
    return sutherland(k, Tv, Num<T>::rcp((T)k.t_ref));
}

// The per-tile constants of the viscous terms: rcp(t_ref) and rcp(Pr).
template <class T>
struct ViscConsts {
    T i_tref, i_pr;
};
template <class T>
__device__ __forceinline__ ViscConsts<T> visc_consts(const CaseConsts<T>& k) {
    return ViscConsts<T>{Num<T>::rcp((T)k.t_ref), Num<T>::rcp((T)k.prandtl)};
}

// CFL wave speed |v| + c of a state (computeWaveSpeedKernel, solver.cu:2729-2739)
template <class T>
__device__ __forceinline__ T wave_speed(const CaseConsts<T>& k, const T v[5], T* Tout = nullptr) {
    T u, vv, p, tt;
    primitives(k, v[0], v[1], v[2], v[3], v[4], u, vv, p, tt);
    if (Tout) *Tout = tt;
    T c = Num<T>::sqrt((T)k.gamma * p * Num<T>::rcp(v[0] + T(1e-10f)));
    return Num<T>::sqrt(u * u + vv * vv) + c;
}

// The same wave speed evaluated in arithmetic type A from a state of type T
// (the FP64 CFL reduction of the compensated FP32 variant: the FP32 state is
// widened and every operation is FP64; for T = A it is wave_speed).
template <class A, class T>
__device__ __forceinline__ A wave_speed_as(const CaseConsts<T>& k, const T v[5]) {
    const A rho = (A)v[0], ru = (A)v[1], rv = (A)v[2], E = (A)v[3], re = (A)v[4];
    const A r = tmax(rho, (A)k.min_rho);
    const A ir = Num<A>::rcp(r);
    const A u = ru * ir, vv = rv * ir;
    const A ke = A(0.5) * r * (u * u + vv * vv);
    const A e = (E > A(0) && re < A(0.001f) * E) ? re : E - ke;
    const A p = tmax((A)k.gm1 * e, (A)k.min_p);
    const A c = Num<A>::sqrt((A)k.gamma * p * Num<A>::rcp(rho + A(1e-10f)));
    return Num<A>::sqrt(u * u + vv * vv) + c;
}

// The CFL maxima pair of a slot as loaded: FP64 values, or FP32 bit patterns.
template <bool DT64> struct SlotPair;
template <> struct SlotPair<true> { double2 v; };
template <> struct SlotPair<false> { uint2 v; };

// Issues the load of a slot's maxima pair, both values in one load (the pair
// is 16-byte aligned).  Kernels call it at a tile's top and form dt after
// issuing their staging loads, so that its round trip overlaps theirs.  The
// load is unconditional (every case's descriptor points at its slot buffer;
// host-dt mode ignores the value): a conditional one would be merged with a
// register copy that waits for it on the spot.
template <bool DT64, class T>
__device__ __forceinline__ SlotPair<DT64> load_slot_pair(const CaseDesc<T>& cd, int slot) {
    SlotPair<DT64> sp;
    if constexpr (DT64) sp.v = *(const double2*)(cd.slots64 + slot * 2);
    else sp.v = *(const uint2*)(cd.b.slots + slot * 2);
    return sp;
}

// dt from the maxima of the previous step (readTimeStepResult,
// solver.cu:3250-3265): FP32 as in the reference, or (DT64: the FP64 state
// and the compensated variant, whose engines set CaseDesc::dt64) FP64 from the FP64
// maxima.  Returned in FP64; for FP32 it is the widened FP32 value, so the
// FP32 arithmetic downstream is unchanged.
template <bool DT64, class T>
__device__ __forceinline__ double dt_from_pair(const CaseDesc<T>& cd, SlotPair<DT64> sp) {
    const CaseConsts<T>& k = cd.c;
    // both halves taken here: the compiler may not split the load and sink
    // its second half into the viscous branch (a second round trip)
    asm volatile("" : "+v"(sp.v.x), "+v"(sp.v.y));
    if (k.dt_mode == 0) return (double)k.host_dt;
    if constexpr (DT64) {
        const double ms = sp.v.x, mn = sp.v.y;
        const double mdx = (double)k.min_dx;
        double dt = (double)k.cfl * mdx / (ms + (double)1e-10f);
        if (k.visc) dt = fmin(dt, (double)k.cfl_visc * mdx * mdx / (mn + (double)1e-10f));
        return dt;
    } else {
        const float ms = __uint_as_float(sp.v.x), mn = __uint_as_float(sp.v.y);
        float dt = k.cfl * k.min_dx / (ms + 1e-10f);
        if (k.visc) {
            float dtv = k.cfl_visc * k.min_dx * k.min_dx / (mn + 1e-10f);
            dt = fminf(dt, dtv);
        }
        return (double)dt;
    }
}

// Atomic maximum of a non-negative value on its bit pattern (non-negative
// IEEE values order like their bits, as unsigned integers of the same width).
__device__ __forceinline__ void atomic_max_bits(unsigned* target, float v) { atomicMax(target, __float_as_uint(v)); }
__device__ __forceinline__ void atomic_max_bits(double* target, double v) {
    atomicMax((unsigned long long*)target, (unsigned long long)__double_as_longlong(v));
}

// Wave-wide max (64 lanes), then across the workgroup through LDS, then one
// atomic max on the bit pattern (non-negative floats order like their bits).
// The slot only grows during a kernel, so a workgroup whose maximum does not
// exceed the value already there skips the atomic: thousands of workgroups
// would otherwise serialise on one address.  Used off the step (maxima after a
// reset or a state upload, statistics); the step itself stores per-tile maxima.
__device__ __forceinline__ void slot_max(unsigned* target, float m) {
    if (!(m > 0.0f)) return;
    unsigned cur = __hip_atomic_load(target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (__float_as_uint(m) > cur) atomicMax(target, __float_as_uint(m));
}
__device__ __forceinline__ void slot_max(double* target, double m) {
    if (!(m > 0.0)) return;
    unsigned long long* t = (unsigned long long*)target;
    unsigned long long cur = __hip_atomic_load(t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((unsigned long long)__double_as_longlong(m) > cur) atomic_max_bits(target, m);
}

// Workgroup maxima of NQ (<= 2) quantities stored for one tile (plain stores,
// no atomics): tile_max[tile] and tile_max[ntiles + tile].  The step's last
// kernel reduces a case's tiles into its CFL slot.  Called by the whole
// workgroup; lds holds NQ x (NT / 64) floats.
template <int NT, int NQ, class V>
__device__ __forceinline__ void block_max_store(const V (&v)[NQ], V (*lds)[NT / 64], V* tile_max,
                                                int ntiles, int tile) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int q = 0; q < NQ; q++) {
        V m = v[q];
        for (int off = 32; off > 0; off >>= 1) m = tmax(m, (V)__shfl_xor(m, off, 64));
        if (lane == 0) lds[q][wave] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 0; q < 2; q++) {
            V m = V(0);
            if (q < NQ)
                for (int w = 0; w < NT / 64; w++) m = tmax(m, lds[q][w]);
            tile_max[q * ntiles + tile] = m;
        }
    }
}

template <int NT, class V, class TG>
__device__ __forceinline__ void block_max_atomic(V v, V* lds_scratch, TG* target) {
    for (int off = 32; off > 0; off >>= 1) v = tmax(v, (V)__shfl_xor(v, off, 64));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) lds_scratch[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        V m = lds_scratch[0];
        for (int w = 1; w < NT / 64; w++) m = tmax(m, lds_scratch[w]);
        slot_max(target, m);
    }
}

}  // namespace fvm
