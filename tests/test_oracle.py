"""The CPU oracle pinned by closed forms, textbook solutions and brute force.

The reference ships no tests, fixtures or golden vectors for this path, so
every check here comes from mathematics the reference's algorithm must obey:
an exact Riemann solution (Toro) for the full step in x and in y, Riemann-flux
known answers, brute-force circle masks, geometric facts of each shape,
freestream preservation, the CFL closed form and a viscous shear identity.
"""
import math
import os

import numpy as np
import pytest

import oracle as O
from fvm_amd import params as P
from exact_riemann import sample
from states import hydrostatic_state


def sod_params(nx, ny, length, along_y=False):
    p = P.default_params()
    p.gamma, p.r_gas = 1.4, 1.0
    p.min_density, p.min_pressure = 1e-8, 1e-8
    p.min_temperature, p.max_temperature = 1e-6, 1e6
    p.max_velocity_limit = p.max_ghost_velocity = 1e6
    p.low_density_threshold = 1e-6
    p.nx, p.ny = nx, ny
    h = length / (ny if along_y else nx)
    p.domain_width, p.domain_height = nx * h, ny * h
    p.dx = p.dy = h
    p.T_inf, p.p_inf, p.mach = 1.0, 1.0, 0.0
    p.rho_inf, p.u_inf, p.v_inf, p.c_inf = 1.0, 0.0, 0.0, math.sqrt(1.4)
    p.obstacle_x, p.obstacle_y, p.obstacle_r = -1000.0, -1000.0, 0.1  # no obstacle cells
    p.enable_viscosity = 0
    return p


def run_sod(along_y, flux=P.FLUX_HLLC):
    # domain [0, 2] along the tube, discontinuity at 1, t = 0.2 (Toro test 1)
    n_long, n_cross = 400, 160
    nx, ny = (n_cross, n_long) if along_y else (n_long, n_cross)
    p = sod_params(nx, ny, 2.0, along_y)
    p.flux_scheme = flux
    h = p.dx
    o = O.Oracle(p, fp64=True)
    s = o.state()
    xs = (np.arange(n_long) + 0.5) * h
    left = xs < 1.0
    rho = np.where(left, 1.0, 0.125)
    pr = np.where(left, 1.0, 0.1)
    prof = np.stack([rho, np.zeros_like(rho), np.zeros_like(rho), pr / 0.4, pr / 0.4])
    if along_y:
        s[:] = prof[:, :, None]
    else:
        s[:] = prof[:, None, :]
    o.set_state(s)
    t_end, dt = 0.2, 0.4 * h / 2.5
    nsteps = int(round(t_end / dt))
    p.dt = t_end / nsteps
    o.step(nsteps, p)
    st = o.state()
    mid = n_cross // 2
    line = st[:, :, mid] if along_y else st[:, mid, :]
    r = line[0]
    un = (line[2] if along_y else line[1]) / r
    ut = (line[1] if along_y else line[2]) / r
    return xs, r, un, ut, line


@pytest.mark.parametrize("flux", [P.FLUX_HLLC, P.FLUX_AUSM_PLUS])
@pytest.mark.parametrize("along_y", [False, True])
def test_sod_exact_solution(along_y, flux):
    xs, r, un, ut, line = run_sod(along_y, flux)
    ex = sample(xs, 0.2, 1.0, (1.0, 0.0, 1.0), (0.125, 0.0, 0.1))
    win = (xs > 0.6) & (xs < 1.4)
    h = xs[1] - xs[0]
    l1_rho = np.sum(np.abs(r[win] - ex[0][win])) * h
    l1_u = np.sum(np.abs(un[win] - ex[1][win])) * h
    # second-order MUSCL + HLLC or AUSM+ on 400 cells: L1 errors of a few 1e-3 (shock and contact smeared)
    assert l1_rho < 6e-3, l1_rho
    assert l1_u < 8e-3, l1_u
    # u = u* between the rarefaction tail (x~0.99) and the shock (x~1.35)
    plateau = (xs > 1.04) & (xs < 1.30)
    assert np.allclose(un[plateau], 0.92745, atol=5e-3)
    # the transverse NRBC rows (80 cells away) leave only a small trace on the centre row
    assert np.all(np.abs(ut) < 1e-3)


def test_sod_x_and_y_agree():
    """The y sweep is the x sweep with the velocity components exchanged; away
    from the ends they agree; what remains (~1e-5) is the trace of the
    different cross-direction boundaries (characteristic rows vs inflow/outflow
    columns) 80 cells away."""
    xs, rx, ux, _, _ = run_sod(False)
    _, ry, uy, _, _ = run_sod(True)
    win = (xs > 0.5) & (xs < 1.5)
    assert np.allclose(rx[win], ry[win], rtol=0, atol=5e-5)
    assert np.allclose(ux[win], uy[win], rtol=0, atol=5e-5)


def kat_params(flux=P.FLUX_HLLC):
    p = P.euler_bow_shock()
    p.flux_scheme = flux
    return p


@pytest.mark.parametrize("flux", [P.FLUX_HLLC, P.FLUX_AUSM_PLUS])
@pytest.mark.parametrize("state", [(1.0, 0.3, -0.2, 0.7), (0.5, 2.5, 0.4, 0.3), (1.3, -1.7, 0.1, 1.1), (2.0, 0.0, 0.0, 1.0)])
def test_riemann_flux_consistency(flux, state):
    """F(U, U) is the physical flux."""
    p = kat_params(flux)
    g = P.f32(1.4)  # the oracle works with the FP32 parameter value
    r, u, v, pr = state
    E = pr / (g - 1) + 0.5 * r * (u * u + v * v)
    phys = np.array([r * u, r * u * u + pr, r * u * v, (E + pr) * u])
    f = O.riemann_flux(p, state, state, fp64=True)
    assert np.allclose(f, phys, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("flux", [P.FLUX_HLLC, P.FLUX_AUSM_PLUS])
def test_riemann_flux_supersonic_upwind(flux):
    """All waves move right: the flux is the left physical flux."""
    p = kat_params(flux)
    L, R = (1.0, 3.0, 0.2, 0.7), (0.8, 2.8, -0.1, 0.6)
    g = P.f32(1.4)  # the oracle works with the FP32 parameter value
    r, u, v, pr = L
    E = pr / (g - 1) + 0.5 * r * (u * u + v * v)
    phys = np.array([r * u, r * u * u + pr, r * u * v, (E + pr) * u])
    assert np.allclose(O.riemann_flux(p, L, R), phys, rtol=1e-12)


@pytest.mark.parametrize("flux", [P.FLUX_HLLC, P.FLUX_AUSM_PLUS])
def test_riemann_flux_mirror_symmetry(flux):
    """Reflecting x swaps the states and flips the normal velocity."""
    p = kat_params(flux)
    L, R = (1.0, 0.4, 0.3, 1.0), (0.6, -0.2, -0.5, 0.5)
    f = O.riemann_flux(p, L, R)
    mL = (R[0], -R[1], R[2], R[3])
    mR = (L[0], -L[1], L[2], L[3])
    fm = O.riemann_flux(p, mL, mR)
    assert np.allclose(fm, [-f[0], f[1], -f[2], -f[3]], rtol=1e-12, atol=1e-14)


def test_hllc_stationary_contact():
    """HLLC keeps a stationary contact exact: zero mass/energy flux, pressure flux p."""
    p = kat_params(P.FLUX_HLLC)
    f = O.riemann_flux(p, (1.0, 0.0, 0.3, 0.8), (0.2, 0.0, -0.4, 0.8))
    assert np.allclose(f, [0.0, 0.8, 0.0, 0.0], atol=1e-14)


def test_limiters():
    for lim in (P.LIMITER_MINMOD, P.LIMITER_VAN_LEER):
        assert O.limited_slope(lim, 0.0, 1.0, 0.5) == 0.0          # extremum: slopes of opposite sign
        assert O.limited_slope(lim, 0.0, 1.0, 2.0) == pytest.approx(1.0)  # linear data: exact slope
    assert O.limited_slope(P.LIMITER_MINMOD, 0.0, 1.0, 4.0) == pytest.approx(1.0)     # min(1, 3)
    assert O.limited_slope(P.LIMITER_MINMOD, 0.0, -2.0, -2.5) == pytest.approx(-0.5)
    assert O.limited_slope(P.LIMITER_VAN_LEER, 0.0, 1.0, 4.0) == pytest.approx(2 * 3 / 4)  # harmonic mean


def test_primitives_and_clamps():
    p = P.euler_bow_shock()
    g = P.f32(1.4)  # the oracle works with the FP32 parameter value
    r, u, v, pr = 1.2, 0.5, -0.25, 0.9
    E = pr / (g - 1) + 0.5 * r * (u * u + v * v)
    out = O.primitives(p, [r, r * u, r * v, E, pr / (g - 1)])
    Rg = p.r_gas
    assert np.allclose(out, [u, v, pr, pr / (r * Rg)], rtol=1e-12)
    # negative internal energy -> p = min_pressure, T clamped at min_temperature
    out = O.primitives(p, [1.0, 3.0, 0.0, 1.0, 1.0])
    assert out[2] == pytest.approx(p.min_pressure, rel=1e-6)
    assert out[3] == pytest.approx(p.min_temperature, rel=1e-6)
    # dual-energy switch: eta = rho_e/E < 1e-3 uses rho_e
    out = O.primitives(p, [1.0, 0.0, 0.0, 10.0, 0.005])
    assert out[2] == pytest.approx((g - 1) * 0.005, rel=1e-12)


def test_riemann_invariant_roundtrip():
    p = P.euler_bow_shock()
    for prim in [(1.0, 0.3, 0.1, 0.7142857), (0.3, 2.0, -0.5, 0.2), (2.5, -0.4, 0.9, 3.0)]:
        assert np.allclose(O.riemann_roundtrip(p, prim), prim, rtol=1e-12)


def test_circle_mask_brute_force():
    """Cell (i,j) is solid iff its centre is closer than 14.5 to (64,64), ghost
    iff within [14.5, 17.5); squared distances are integers + 0.5, never tied."""
    p = P.euler_bow_shock()
    o = O.Oracle(p, fp64=False)
    ct = o.cell_types()
    j, i = np.mgrid[0:p.ny, 0:p.nx]
    d2 = (2 * i + 1 - 128) ** 2 + (2 * j + 1 - 128) ** 2  # (2*dist)^2, integers
    exp = np.where(d2 < 29 ** 2, P.CELL_SOLID, np.where(d2 < 35 ** 2, P.CELL_GHOST, P.CELL_FLUID))
    exp[:, 0] = P.CELL_INFLOW
    exp[:, -1] = P.CELL_OUTFLOW
    assert np.array_equal(ct, exp.astype(np.uint8))


@pytest.mark.parametrize("shape", [P.SHAPE_DIAMOND, P.SHAPE_CAPSULE, P.SHAPE_TRIANGLE, P.SHAPE_STAR, P.SHAPE_STARSHIP])
def test_shape_sdf_geometry(shape):
    p = P.euler_bow_shock()
    p.obstacle_shape = shape
    p.obstacle_r = 20.0
    cx, cy, r = p.obstacle_x, p.obstacle_y, p.obstacle_r
    f = lambda x, y: O.sdf_point(p, x, y)
    assert f(cx, cy) < 0  # centre inside
    if shape == P.SHAPE_DIAMOND:
        assert f(cx + r, cy) == pytest.approx(0.0, abs=1e-4)               # vertex on the boundary
        assert f(cx, cy) == pytest.approx(-r * 0.7071, rel=1e-6)
    if shape == P.SHAPE_CAPSULE:
        assert f(cx + 1.5 * r, cy) == pytest.approx(0.0, abs=1e-4)          # cap tip
        assert f(cx, cy) == pytest.approx(-0.5 * r, rel=1e-6)
        assert f(cx, cy + r) == pytest.approx(0.5 * r, rel=1e-6)
    if shape == P.SHAPE_TRIANGLE:
        assert f(cx, cy) == pytest.approx(-0.5 * r, rel=1e-5)               # inradius = r/2
        assert f(cx + r, cy) == pytest.approx(0.0, abs=1e-4)                # tip vertex
        assert f(cx + 2 * r, cy) == pytest.approx(r, rel=1e-5)
    if shape == P.SHAPE_STAR:
        assert f(cx, cy - 2 * r) == pytest.approx(r, rel=1e-5)              # nearest point: the tip at -90 deg
        assert f(cx, cy - r) == pytest.approx(0.0, abs=1e-3)
    if shape == P.SHAPE_STARSHIP:
        assert f(cx + 2 * r, cy) == pytest.approx(r, rel=1e-5)              # body circle dominates
    # rotation by 90 degrees maps the geometry
    q = p.copy()
    q.obstacle_rotation = math.pi / 2
    for (dx, dy) in [(0.3 * r, 0.1 * r), (1.2 * r, -0.4 * r), (-0.7 * r, 0.9 * r)]:
        a = O.sdf_point(p, cx + dx, cy + dy)
        b = O.sdf_point(q, cx - dy, cy + dx)
        assert a == pytest.approx(b, abs=1e-3 * r)


def test_freestream_preservation():
    """Uniform flow with no obstacle stays uniform."""
    p = P.euler_bow_shock(nx=64, ny=32)
    p.noise_field = P.NOISE_NONE
    p.obstacle_x = p.obstacle_y = -1000.0
    p.dt = 0.1
    o = O.Oracle(p, fp64=True)
    s0 = o.state()
    o.step(20, p)
    s1 = o.state()
    assert np.allclose(s1, s0, rtol=1e-12, atol=1e-12)


def test_cfl_dt_closed_form():
    """Uniform Mach-3 flow with c = 1: max(|v| + c) = 4, dt = 0.5 * 1 / 4."""
    p = P.euler_bow_shock()
    p.noise_field = P.NOISE_NONE
    o = O.Oracle(p, fp64=True)
    ms, _ = o.maxima()
    assert ms == pytest.approx(4.0, rel=1e-6)
    assert o.stable_dt() == pytest.approx(0.125, rel=1e-6)


def test_viscous_linear_shear():
    """u = a*y, v = 0, uniform rho and p: MUSCL reconstructs linear data
    exactly, so convection is inert; diffusion leaves rho_u unchanged in the
    interior (div tau = 0) and heats by Phi = mu a^2 per unit time."""
    p = P.viscous_cylinder(nx=48, ny=48, diameter=4.0, reynolds=50.0, mach=0.3)
    p.obstacle_x = p.obstacle_y = -1000.0
    p.noise_field = P.NOISE_NONE
    p.dt = 0.05
    o = O.Oracle(p, fp64=True)
    s = o.state()
    a = 0.004
    y = (np.arange(p.ny) + 0.5)[:, None] * np.ones((1, p.nx))
    u = a * (y - 24.0)
    rho, pr = 1.0, 1.0 / 1.4
    s[0] = rho
    s[1] = rho * u
    s[2] = 0.0
    s[4] = pr / 0.4
    s[3] = s[4] + 0.5 * rho * u * u
    o.set_state(s)
    o.step(1, p)
    s1 = o.state()
    core = (slice(20, 28), slice(20, 28))
    assert np.allclose(s1[1][core], s[1][core], atol=1e-13)
    mu = 1.0 * 0.3 * 4.0 / 50.0
    heat = p.dt * mu * a * a
    assert np.allclose(s1[4][core] - s[4][core], heat, rtol=2e-3)
    assert np.allclose(s1[3][core] - s[3][core], heat, rtol=2e-3)


def test_fp32_fp64_agree_on_bow_shock():
    p = P.euler_bow_shock()
    p.dt = 0.1
    a = O.Oracle(p, fp64=True)
    b = O.Oracle(p, fp64=False)
    a.step(50, p.copy())
    b.step(50, p.copy())
    sa, sb = a.state(), b.state()
    for c in range(5):
        assert np.abs(sa[c] - sb[c]).sum() / np.abs(sa[c]).sum() < 1e-5


def test_noise_generator():
    u = np.array([O.noise_unit(7, i) for i in range(20000)])
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.01
    assert O.noise_unit(7, 3) == O.noise_unit(7, 3)
    assert O.noise_unit(7, 3) != O.noise_unit(8, 3)


@pytest.mark.parametrize("shape", [0, 2, 4])
@pytest.mark.parametrize("along_y", [False, True])
def test_forces_discrete_buoyancy(shape, along_y):
    """A linear pressure field p = 1 + b x on fluid at rest: the wall-flux force
    is the discrete buoyancy -b * (solid + ghost cells) * dx * dy along the
    gradient and exactly 0 across it (each row of a one-run body contributes
    p(first fluid cell) - p(last wall cell) = -b * run length * dx)."""
    p = P.euler_bow_shock(nx=96, ny=64, radius=10.0, center=(40.0, 32.0))
    p.obstacle_shape = shape
    p.obstacle_rotation = 0.3
    P.compute_derived(p)
    b = 0.01
    o = O.Oracle(p, fp64=True)
    o.set_state(hydrostatic_state(p, b, along_y))
    f = o.forces()
    ct = o.cell_types()
    nwall = int(np.count_nonzero((ct == P.CELL_SOLID) | (ct == P.CELL_GHOST)))
    expect = -b * nwall * p.dx * p.dy
    along, across = (f[1], f[0]) if along_y else (f[0], f[1])
    assert along == pytest.approx(expect, rel=1e-12)
    assert abs(across) < 1e-12
    assert f[2] > 0
    o.close()


def test_forces_zero_at_uniform_pressure():
    p = P.viscous_cylinder(nx=128, ny=64, diameter=16.0, reynolds=100.0)
    o = O.Oracle(p, fp64=True)
    f = o.forces()
    assert abs(f[0]) < 1e-12 and abs(f[1]) < 1e-12
    o.close()


def test_obstacle_update_zeroes_exposed_cells():
    """updateSDFWithFixupKernel (solver.cu:607-664): cells that turn from
    solid/ghost into fluid get a zero state; interior fluid cells are not
    touched by the boundary conditions, so exactly those cells read zero after
    the update, and the new mask equals a fresh marking of the new geometry."""
    p = P.euler_bow_shock(nx=128, ny=64, radius=12.0, center=(40.0, 32.0))
    p.dt = 0.1
    o = O.Oracle(p, fp64=True)
    o.step(5, p.copy())
    old = o.cell_types()
    p2 = p.copy()
    p2.obstacle_x += 6.0
    o.update_obstacle(p2)
    new = o.cell_types()
    fresh = O.Oracle(p2, fp64=True)
    assert np.array_equal(new, fresh.cell_types())
    exposed = np.isin(old, (P.CELL_SOLID, P.CELL_GHOST)) & (new == P.CELL_FLUID)
    assert exposed.sum() > 0
    s = o.state()
    zero = np.all(s == 0.0, axis=0)
    assert np.array_equal(zero, exposed)
    o.close()
    fresh.close()


def test_openmp_oracle_is_thread_count_independent():
    """Jacobi boundary semantics and per-cell arithmetic make the oracle's
    result independent of how cells are split over threads."""
    p = P.euler_bow_shock(nx=96, ny=48, radius=8.0, center=(30.0, 24.0))
    p.dt = 0.1
    out = []
    for threads in (1, 3, 8):
        O.set_threads(threads)
        o = O.Oracle(p, fp64=False)
        o.step(30, p.copy())
        out.append(o.state())
        o.close()
    O.set_threads(os.cpu_count() or 1)
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])
