"""Simulation parameters: the ctypes mirror of ``fvm_params`` (include/fvm/fvm.h).

The first block of fields is ``SimParams`` of the reference
(src/include/common.h:50-91) field for field; the second block holds the
reference's compile-time physical constants (src/solver.cu:10-14, 39-41, 52;
src/include/common.h:7-19) as run-time values; the third selects additions.

Also holds the workload builders for the synthetic configurations (a 256x128
Mach-3 Euler bow-shock case, the 2048x1024 Re-200 viscous cylinder, the
512-case mixed batch and the served request stream) and the SI <->
nondimensional unit map.
"""
from __future__ import annotations

import ctypes
import math

# enum values are ABI (common.h:39-47, 105-111)
CELL_FLUID, CELL_SOLID, CELL_GHOST, CELL_INFLOW, CELL_OUTFLOW = range(5)
SHAPE_CIRCLE, SHAPE_STAR, SHAPE_DIAMOND, SHAPE_CAPSULE, SHAPE_TRIANGLE, SHAPE_STARSHIP = range(6)
FLUX_HLLC, FLUX_AUSM_PLUS = 0, 1
LIMITER_MINMOD, LIMITER_VAN_LEER = 0, 1
WALL_REFERENCE, WALL_SLIP = 0, 1
PRECISION_FP32, PRECISION_FP32_KAHAN, PRECISION_FP64 = 0, 1, 2
DT_HOST, DT_DEVICE_CFL = 0, 1
SCHEDULE_AUTO, SCHEDULE_SPLIT, SCHEDULE_FUSED = 0, 1, 2
NOISE_NONE, NOISE_DENSITY, NOISE_TRANSVERSE_VELOCITY = 0, 1, 2
FIELD_TEMPERATURE, FIELD_PRESSURE, FIELD_DENSITY, FIELD_VELOCITY_MAG, FIELD_MACH, FIELD_CELL_TYPE = range(6)


class FvmParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
        ("dx", ctypes.c_float), ("dy", ctypes.c_float),
        ("domain_width", ctypes.c_float), ("domain_height", ctypes.c_float),
        ("mach", ctypes.c_float), ("T_inf", ctypes.c_float), ("p_inf", ctypes.c_float),
        ("rho_inf", ctypes.c_float), ("u_inf", ctypes.c_float), ("v_inf", ctypes.c_float),
        ("c_inf", ctypes.c_float),
        ("obstacle_x", ctypes.c_float), ("obstacle_y", ctypes.c_float),
        ("obstacle_r", ctypes.c_float), ("obstacle_rotation", ctypes.c_float),
        ("obstacle_shape", ctypes.c_int32),
        ("wing_rotation", ctypes.c_float),
        ("cfl", ctypes.c_float), ("dt", ctypes.c_float), ("t_current", ctypes.c_float),
        ("step", ctypes.c_int32),
        ("enable_viscosity", ctypes.c_int32),
        ("cfl_visc", ctypes.c_float),
        ("adiabatic_wall", ctypes.c_int32),
        ("T_wall", ctypes.c_float),
        ("paused", ctypes.c_int32),
        ("gamma", ctypes.c_float), ("r_gas", ctypes.c_float), ("mu_ref", ctypes.c_float),
        ("t_ref", ctypes.c_float), ("s_sutherland", ctypes.c_float), ("prandtl", ctypes.c_float),
        ("min_density", ctypes.c_float), ("min_pressure", ctypes.c_float),
        ("max_temperature", ctypes.c_float), ("min_temperature", ctypes.c_float),
        ("low_density_threshold", ctypes.c_float), ("max_velocity_limit", ctypes.c_float),
        ("max_ghost_velocity", ctypes.c_float), ("slip_reference_velocity", ctypes.c_float),
        ("flux_scheme", ctypes.c_int32), ("limiter", ctypes.c_int32),
        ("wall_mode", ctypes.c_int32), ("precision", ctypes.c_int32),
        ("dt_mode", ctypes.c_int32), ("noise_field", ctypes.c_int32),
        ("noise_amplitude", ctypes.c_float), ("noise_seed", ctypes.c_uint32),
        ("reserved", ctypes.c_int32 * 8),
    ]

    def copy(self) -> "FvmParams":
        out = FvmParams()
        ctypes.pointer(out)[0] = self
        return out

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


def f32(x: float) -> float:
    """Round a Python float to the nearest IEEE single."""
    return ctypes.c_float(x).value


def default_params() -> FvmParams:
    """SimParams defaults (common.h:53-91) with the reference's SI constants."""
    p = FvmParams()
    p.nx, p.ny = 512, 256
    p.dx = p.dy = 1.0
    p.domain_width, p.domain_height = 10.0, 5.0
    p.mach, p.T_inf, p.p_inf = 3.0, 300.0, 101325.0
    p.obstacle_x = p.obstacle_y = 2.5
    p.obstacle_r = 0.5
    p.obstacle_shape = SHAPE_CIRCLE
    p.cfl, p.dt, p.cfl_visc = 0.5, 1e-6, 0.4
    p.adiabatic_wall, p.T_wall, p.paused = 1, 300.0, 1
    p.gamma, p.r_gas, p.mu_ref, p.t_ref = 1.4, 287.05, 1.716e-5, 273.15
    p.s_sutherland, p.prandtl = 110.4, 0.71
    p.min_density, p.min_pressure = 1e-4, 1e-2
    p.max_temperature, p.min_temperature = 5000.0, 50.0
    p.low_density_threshold, p.max_velocity_limit = 0.01, 5000.0
    p.max_ghost_velocity, p.slip_reference_velocity = 2500.0, 100.0
    return p


def compute_derived(p: FvmParams) -> FvmParams:
    """SimParams::computeDerived (common.h:93-101), FP32 like the reference."""
    p.rho_inf = f32(f32(p.p_inf) / f32(f32(p.r_gas) * f32(p.T_inf)))
    p.c_inf = f32(math.sqrt(f32(f32(f32(p.gamma) * f32(p.r_gas)) * f32(p.T_inf))))
    p.u_inf = f32(f32(p.mach) * p.c_inf)
    p.dx = f32(f32(p.domain_width) / f32(float(p.nx)))
    p.dy = f32(f32(p.domain_height) / f32(float(p.ny)))
    p.obstacle_y = f32(f32(p.domain_height) / 2.0)
    return p


# ---------------------------------------------------------------------------
# SI <-> nondimensional unit map.  Scales from the reference defaults
# (T = 300 K, p = 101325 Pa, common.h:62-63): rho_s = p/(R T), c_s = sqrt(gamma R T),
# p_s = rho_s c_s^2, T_s = 300 K, L_s = one cell.  With them R* = 1/gamma and T_inf* = 1.
# ---------------------------------------------------------------------------
SI_R, SI_T, SI_P, GAMMA = 287.05, 300.0, 101325.0, 1.4
RHO_S = SI_P / (SI_R * SI_T)
C_S = math.sqrt(GAMMA * SI_R * SI_T)
P_S = RHO_S * C_S * C_S
T_S = SI_T


def set_nondimensional_units(p: FvmParams) -> FvmParams:
    """Scale every dimensional constant of the reference to the units above."""
    p.gamma = GAMMA
    p.r_gas = 1.0 / GAMMA
    p.t_ref = 273.15 / T_S
    p.s_sutherland = 110.4 / T_S
    p.min_density = 1e-4 / RHO_S
    p.min_pressure = 1e-2 / P_S
    p.max_temperature = 5000.0 / T_S
    p.min_temperature = 50.0 / T_S
    p.low_density_threshold = 0.01 / RHO_S
    p.max_velocity_limit = 5000.0 / C_S
    p.max_ghost_velocity = 2500.0 / C_S
    p.slip_reference_velocity = 100.0 / C_S
    p.T_inf = 1.0
    p.p_inf = 1.0 / GAMMA
    p.T_wall = 1.0
    return p


def sutherland_mu_ref_for(mu_inf: float, T_inf: float, t_ref: float, s: float) -> float:
    """mu_ref such that mu(T_inf) = mu_inf under mu = mu_ref (T/T_ref)^1.5 (T_ref+S)/(T+S)."""
    tr = T_inf / t_ref
    return mu_inf / (tr * math.sqrt(tr) * (t_ref + s) / (T_inf + s))


def euler_bow_shock(nx: int = 256, ny: int = 128, flux: int = FLUX_HLLC, limiter: int = LIMITER_MINMOD,
                    seed: int = 1, mach: float = 3.0, radius: float = 16.0, center=(64.0, 64.0)) -> FvmParams:
    """Inviscid Euler, Mach 3 along +x, rho=1, p=1/gamma, circle r=16 at (64,64), dx=dy=1,
    fixed inflow left, NRBC elsewhere, slip/no-slip blend wall, CFL 0.5, FP32,
    +-1e-3 relative density noise."""
    p = set_nondimensional_units(default_params())
    p.nx, p.ny = nx, ny
    p.domain_width, p.domain_height = float(nx), float(ny)
    p.mach = mach
    p.obstacle_x = center[0]
    p.obstacle_r = radius
    p.obstacle_shape = SHAPE_CIRCLE
    p.enable_viscosity = 0
    p.cfl = 0.5
    p.flux_scheme, p.limiter = flux, limiter
    p.noise_field, p.noise_amplitude, p.noise_seed = NOISE_DENSITY, 1e-3, seed
    compute_derived(p)
    p.obstacle_y = center[1]
    p.dt = 0.125  # placeholder; set from the CFL limit by the caller
    return p


def viscous_cylinder(nx: int = 2048, ny: int = 1024, diameter: float = 64.0, reynolds: float = 200.0,
                     mach: float = 0.3, seed: int = 2, flux: int = FLUX_HLLC,
                     limiter: int = LIMITER_VAN_LEER, center=None) -> FvmParams:
    """Navier-Stokes, Mach 0.3, Re 200 on the cylinder diameter, Sutherland viscosity in
    nondimensional form (T_ref = 273.15/300, S = 110.4/300), no-slip adiabatic wall,
    HLLC + van Leer, CFL 0.4, cfl_visc 0.4, +-1e-3 transverse-velocity noise."""
    p = set_nondimensional_units(default_params())
    p.nx, p.ny = nx, ny
    p.domain_width, p.domain_height = float(nx), float(ny)
    p.mach = mach
    cx, cy = center if center is not None else (nx / 4.0, ny / 2.0)
    p.obstacle_x = cx
    p.obstacle_r = diameter / 2.0
    p.obstacle_shape = SHAPE_CIRCLE
    p.enable_viscosity = 1
    p.adiabatic_wall = 1
    p.cfl, p.cfl_visc = 0.4, 0.4
    p.flux_scheme, p.limiter = flux, limiter
    p.noise_field, p.noise_amplitude, p.noise_seed = NOISE_TRANSVERSE_VELOCITY, 1e-3, seed
    mu_inf = 1.0 * mach * 1.0 * diameter / reynolds  # rho_inf u_inf D / Re with c_inf = 1
    p.mu_ref = sutherland_mu_ref_for(mu_inf, 1.0, p.t_ref, p.s_sutherland)
    compute_derived(p)
    p.obstacle_y = cy
    p.dt = 0.3
    return p


def mixed_case(rng, n_lo=64 * 64, n_hi=512 * 256, nx_range=(64, 512), ny_range=(64, 256),
               shapes=(SHAPE_CIRCLE, SHAPE_STAR, SHAPE_DIAMOND, SHAPE_CAPSULE, SHAPE_TRIANGLE)) -> FvmParams:
    """One case of the mixed batch: cells log-uniform in [n_lo, n_hi], aspect U[1,2],
    shape uniform over circle/star/diamond/capsule/triangle, size 8-25% of the height,
    random rotation, centre (W/4, H/2), Mach U[0.3, 4], viscous (Re U[100,1000]) or
    inviscid 50/50, HLLC + minmod, CFL 0.5 / 0.4, device CFL dt, FP32."""
    n = math.exp(rng.uniform(math.log(n_lo), math.log(n_hi)))
    a = rng.uniform(1.0, 2.0)
    ny = int(round(math.sqrt(n / a)))
    nx = int(round(a * ny))
    nx = min(max(nx, nx_range[0]), nx_range[1])
    ny = min(max(ny, ny_range[0]), ny_range[1])
    shape = int(shapes[rng.integers(len(shapes))])
    s = rng.uniform(0.08, 0.25)
    rot = rng.uniform(0.0, 2 * math.pi)
    mach = rng.uniform(0.3, 4.0)
    visc = bool(rng.integers(2))
    seed = int(rng.integers(1, 2 ** 31))
    if visc:
        re = rng.uniform(100.0, 1000.0)
        p = viscous_cylinder(nx=nx, ny=ny, diameter=2.0 * s * ny, reynolds=re, mach=mach, seed=seed,
                             limiter=LIMITER_MINMOD, center=(nx / 4.0, ny / 2.0))
    else:
        p = euler_bow_shock(nx=nx, ny=ny, seed=seed, mach=mach, radius=s * ny, center=(nx / 4.0, ny / 2.0))
    p.obstacle_shape = shape
    p.obstacle_rotation = rot
    p.cfl, p.cfl_visc = 0.5, 0.4
    p.dt_mode = DT_DEVICE_CFL
    return p


def mixed_batch(n_cases: int = 512, seed: int = 4) -> list:
    import numpy as np
    rng = np.random.default_rng(seed)
    return [mixed_case(rng) for _ in range(n_cases)]


def served_request(rng) -> tuple:
    """A served solve request: grid log-uniform 128x64 .. 2048x1024, other
    parameters as the mixed batch, steps uniform in 500..5000."""
    p = mixed_case(rng, n_lo=128 * 64, n_hi=2048 * 1024, nx_range=(128, 2048), ny_range=(64, 1024))
    steps = int(rng.integers(500, 5001))
    return p, steps
