"""Pins of the oracle that do not use the reference: closed forms, invariants,
brute force, and the reference's own result files (step counts).

Each test is chosen so that a plausible slip (a dropped term, a wrong sign or
index, a transposed axis) fails it.
"""

import math

import numpy as np
import pytest

from oracle import hit as H
from oracle import nonperiodic as N
from oracle import periodic as P
from oracle import riemann as R
from oracle.equations import Gas, conservative, viscous_rhs
from oracle.rk import ssprk3


def _modified_wavenumber(w):
    a, b, c, al = 25 / 16, 1 / 5, -1 / 80, 3 / 8
    return (a * np.sin(w) + (b / 2) * np.sin(2 * w) + (c / 3) * np.sin(3 * w)) / (1 + 2 * al * np.cos(w))


@pytest.mark.parametrize("n,k", [(32, 1), (32, 5), (48, 11), (24, 3)])
def test_compact_periodic_fourier_symbol(n, k):
    """Smooth interface flux + circulant solve = Lele's alpha=3/8 scheme:
    d/dx e^{ikx} -> i k'(k dx)/dx e^{ikx}."""
    L = 2 * np.pi
    h = L / n
    x = h * np.arange(n)
    f = np.sin(k * x)
    d = P.compact_solve(P.smooth_flux(f), h)
    kp = _modified_wavenumber(k * h) / h
    np.testing.assert_allclose(d, kp * np.cos(k * x), atol=1e-12)


def test_compact_scheme_is_eighth_order():
    w = 0.5
    assert abs(_modified_wavenumber(w) - w) < 2e-7  # error ~ w^9
    assert abs(_modified_wavenumber(0.25) - 0.25) < 5e-10


@pytest.mark.parametrize("n", [8, 13, 32])
def test_thomas_matches_dense_solve(n):
    rng = np.random.default_rng(n)
    F = rng.standard_normal((3, 2, n + 1))
    h = 0.1
    got = N.compact_solve(F, h)
    A = np.eye(n)
    for i in range(1, n - 1):
        A[i, i - 1] = A[i, i + 1] = 3 / 8
    raw = (F[..., 1:] - F[..., :-1]) / h
    want = np.linalg.solve(A, raw.reshape(-1, n).T).T.reshape(raw.shape)
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13)


def test_periodic_solve_matches_dense_cyclic():
    n = 17
    rng = np.random.default_rng(3)
    F = rng.standard_normal((2, n))
    A = np.eye(n)
    for i in range(n):
        A[i, (i - 1) % n] = A[i, (i + 1) % n] = 3 / 8
    raw = (F - np.roll(F, 1, axis=-1)) / 0.2
    np.testing.assert_allclose(P.compact_solve(F, 0.2), np.linalg.solve(A, raw.T).T, atol=1e-13)


def test_weno7_invariants():
    rng = np.random.default_rng(0)
    # constants map to themselves
    v = rng.uniform(0.5, 2.0, 50)
    np.testing.assert_allclose(P.weno7(*(v,) * 7), v, rtol=1e-14)
    # the 544 quirk: IS0 of constant data is -3 v^2, the others are 0
    _, (b0, b1, b2, b3) = P.weno7_parts(*(v,) * 7)
    np.testing.assert_allclose(b0, -3 * v**2, rtol=1e-10)
    for b in (b1, b2, b3):
        np.testing.assert_allclose(b, 0, atol=1e-9 * np.max(v**2))
    # linear data j+10: IS0 = -60, IS1..3 = 240 (not translation-invariant IS0)
    j = np.arange(7.0) + 10
    _, bs = P.weno7_parts(*[np.array([t]) for t in j])
    np.testing.assert_allclose([b[0] for b in bs], [-60, 240, 240, 240])
    # mirror symmetry IS2(v) = IS1(reversed v) and IS3 = IS0 with 547
    vv = rng.standard_normal((7, 40))
    _, (c0, c1, c2, c3) = P.weno7_parts(*vv)
    _, (r0, r1, r2, r3) = P.weno7_parts(*vv[::-1])
    np.testing.assert_allclose(c2, r1, rtol=1e-10, atol=1e-9)
    np.testing.assert_allclose(c3, r0 + 3 * vv[6] ** 2, rtol=1e-10, atol=1e-9)
    # tiny-amplitude data: weights -> linear weights, output -> upwind-7 formula
    s = 1e-12 * rng.standard_normal((7, 40))
    up = np.array([-3, 25, -101, 319, 214, -38, 4]) / 420
    np.testing.assert_allclose(P.weno7(*s), up @ s, rtol=1e-6, atol=1e-20)
    # a step inside stencil 0 only: output follows the smooth stencils
    step = [np.array([x]) for x in (5.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0)]
    assert abs(P.weno7(*step)[0] - 1.0) < 1e-6


def test_periodic_sensor_special_cases():
    gas = Gas()
    nx, ny = 40, 24
    h = 2 * np.pi / nx
    ones = np.ones((ny, nx))
    q = conservative(ones, 0 * ones, 0 * ones, ones / 1.4, gas)
    assert not P.sensor(q, gas, h, h, P.SensorParams()).any()
    rho = ones.copy()
    rho[:, 10:20] = 2.0  # two x-discontinuities (between 9|10 and 19|20)
    q = conservative(rho, 0 * ones, 0 * ones, ones / 1.4, gas)
    m = P.sensor(q, gas, h, h, P.SensorParams(width=4))
    cols = np.flatnonzero(m.all(axis=0))
    assert m.any(axis=0).sum() == cols.size == 20
    assert set(cols) == set(range(9 - 4, 10 + 5)) | set(range(19 - 4, 20 + 5))
    # transposed: rows not columns
    qt = conservative(rho.T.copy(), 0 * ones.T, 0 * ones.T, ones.T / 1.4, gas)
    mt = P.sensor(qt, gas, h, h, P.SensorParams(width=4))
    np.testing.assert_array_equal(mt, m.T)


def test_riemann_sensor_boundary_band():
    gas = Gas()
    n = 30
    ones = np.ones((n, n))
    q = conservative(ones, 0 * ones, 0 * ones, ones, gas)
    m = N.sensor(q, gas, 1 / n, 1 / n, N.SensorParams(width=4, shear_threshold=None, boundary_guard=4))
    want = np.zeros((n, n), bool)
    want[:8, :] = want[-8:, :] = want[:, :8] = want[:, -8:] = True
    np.testing.assert_array_equal(m, want)


def test_hyperviscosity_fourier_symbol():
    nx, ny = 32, 24
    dx, dy = 2 * np.pi / nx, 2 * np.pi / ny
    X, Y = np.meshgrid(dx * np.arange(nx), dy * np.arange(ny))
    kx, ky = 3, 2
    f = np.cos(kx * X + ky * Y)
    sym = -(4 / dx**2) * np.sin(kx * dx / 2) ** 2 - (4 / dy**2) * np.sin(ky * dy / 2) ** 2
    np.testing.assert_allclose(P._lap(P._lap(f, dx, dy), dx, dy), sym**2 * f, atol=1e-9 * sym**2)
    # masked filter touches only active nodes (reference tests/test_hit_run_metadata_and_hyperviscosity.py:60-82)
    gas = Gas()
    rho = 1.0 + 0.01 * np.sin(X) * np.cos(Y)
    q = conservative(rho, 0.02 * np.sin(Y), 0.02 * np.cos(X), np.full_like(rho, 1 / 1.4), gas)
    act = np.zeros((ny, nx), bool)
    act[3:7, 4:9] = True
    d = P.hyperviscosity(q, gas, dx, dy, 0.002, act) - q
    assert np.any(np.abs(d[:, act]) > 0) and np.allclose(d[:, ~act], 0)
    with pytest.raises(ValueError):
        P.hyperviscosity(q, gas, dx, dy, 0.002, None)
    with pytest.raises(ValueError):
        N.hyperviscosity(q, act[:, :-1], gas, dx, dy, 0.001)


def test_viscous_shear_wave():
    nx, ny = 16, 32
    dx, dy = 2 * np.pi / nx, 2 * np.pi / ny
    X, Y = np.meshgrid(dx * np.arange(nx), dy * np.arange(ny))
    A, mu = 0.1, 0.01
    gas = Gas(viscosity=mu)
    ones = np.ones_like(X)
    q = conservative(ones, A * np.sin(Y), 0 * ones, ones, gas)
    r = viscous_rhs(q, gas, dx, dy)
    # discrete symbol of the nested central difference: -sin(dy)^2/dy^2
    np.testing.assert_allclose(r[1], -mu * A * np.sin(Y) * np.sin(dy) ** 2 / dy**2, atol=1e-15)
    np.testing.assert_allclose(r[0], 0, atol=0)
    np.testing.assert_allclose(r[2], 0, atol=1e-15)
    # energy: d/dy(u tau_xy) = mu A^2 (sin dy/dy) d/dy(sin y cos y)
    e = mu * A**2 * (np.sin(dy) / dy) * 0.5 * np.cos(2 * Y) * np.sin(2 * dy) / dy
    np.testing.assert_allclose(r[3], e, atol=1e-15)


@pytest.mark.parametrize("z", [-0.3, -1.2, 0.4])
def test_ssprk3_stability_polynomial(z):
    q = np.array([1.0, 2.0])
    out = ssprk3(q, lambda s: z * s, 1.0)
    np.testing.assert_allclose(out, q * (1 + z + z**2 / 2 + z**3 / 6), rtol=1e-15)


def test_hit_initial_state_closed_forms():
    p = H.HITParams(nx=64, ny=48, target_mach=0.25, initial_kmin=3, initial_kmax=5)
    q = H.initial_state(p, np.random.default_rng(1234))
    rho, u, v = q[0], q[1] / q[0], q[2] / q[0]
    assert abs(0.5 * np.mean(rho * (u**2 + v**2)) - 0.25**2 / 2) < 1e-15
    kx, ky, _ = H.wavenumbers(p.nx, p.ny, p.dx, p.dy)
    div = np.fft.ifft2(1j * kx * np.fft.fft2(u) + 1j * ky * np.fft.fft2(v)).real
    assert np.sqrt(np.mean(div**2)) < 1e-12
    uh = np.abs(np.fft.fft2(u))
    k = np.sqrt(kx**2 + ky**2)
    assert np.all(uh[(k < 3 - 1e-9) | (k > 5 + 1e-9)] < 1e-10 * uh.max())


def test_forcing_is_solenoidal_shell_field():
    p = H.HITParams(nx=32, ny=24)
    f = H.ShellOUForcing(p.nx, p.ny, p.dx, p.dy, 3.0, 5.0, seed=5, target_power=None)
    ones = np.ones((p.ny, p.nx))
    fx, fy, info = f.update(1e-3, ones, 0 * ones, 0 * ones)
    kx, ky, k2 = H.wavenumbers(p.nx, p.ny, p.dx, p.dy)
    fxh, fyh = np.fft.fft2(fx), np.fft.fft2(fy)
    assert np.max(np.abs(kx * fxh + ky * fyh)) < 1e-10 * np.max(np.abs(fxh))
    k = np.sqrt(k2)
    off = (k < 3 - 1e-9) | (k > 5 + 1e-9)
    assert np.max(np.abs(fxh[off])) < 1e-10 * np.max(np.abs(fxh))
    assert abs(np.sqrt(np.mean(fx**2 + fy**2)) - 1.0) < 1e-12  # force_rms, alpha = 1
    assert info["alpha"] == 1.0


# Reference result files (results/*/*/diagnostics.json:15) fix these step counts.
@pytest.mark.parametrize("factory,kw,steps", [
    (R.config3, dict(cfl=0.4), 1343),
    (R.config3, dict(cfl=0.25), 2149),
    (R.config3_08, dict(tfinal=0.85), 3805),
    (R.config6, dict(), 1158),
])
def test_riemann_step_counts_512(factory, kw, steps):
    p = factory(nx=512, ny=512, **kw)
    q = N.outflow_guard(R.initial_state(p), 2)
    dt, n = N.fixed_dt(q, p.gas(), p.dx, p.dy, p.cfl, p.tfinal)
    assert n == steps
    assert math.isclose(dt * n, p.tfinal, rel_tol=1e-15)


def test_riemann_initial_states_reference_tests():
    """reference tests/test_riemann_config3_08.py:17-34 and test_riemann_config6.py:11-43."""
    p = R.config3_08(nx=10, ny=10)
    q = R.initial_state(p)
    from oracle.equations import primitives

    rho, u, v, pr = primitives(q, p.gas())
    for (i, j), st in {(0, 0): (0.138, 1.206, 1.206, 0.029), (0, 9): (0.5323, 0.0, 1.206, 0.3),
                       (9, 0): (0.5323, 1.206, 0.0, 0.3), (9, 9): (1.5, 0.0, 0.0, 1.5)}.items():
        np.testing.assert_allclose((rho[i, j], u[i, j], v[i, j], pr[i, j]), st, rtol=0, atol=1e-14)
    assert (rho == 0.138).sum() == 64 and (rho == 0.5323).sum() == 32 and (rho == 1.5).sum() == 4
    p6 = R.config6(nx=4, ny=4)
    rho, u, v, pr = primitives(R.initial_state(p6), p6.gas())
    np.testing.assert_array_equal(rho, [[1, 1, 3, 3], [1, 1, 3, 3], [2, 2, 1, 1], [2, 2, 1, 1.0]])
    np.testing.assert_array_equal(u, [[-0.75] * 4, [-0.75] * 4, [0.75] * 4, [0.75] * 4])
    np.testing.assert_array_equal(v, [[0.5, 0.5, -0.5, -0.5]] * 4)
    np.testing.assert_array_equal(pr, np.ones((4, 4)))


# ---------------------------------------------------------------- diagnostics
@pytest.mark.parametrize("nx,ny", [(32, 24), (33, 25)])
def test_spectral_derivatives_of_a_plane_wave(nx, ny):
    """d/dx, d/dy of sin(3x + 2y) on the 2 pi box are 3 cos, 2 cos (even and odd n)."""
    L = 2.0 * np.pi
    dx, dy = L / nx, L / ny
    X, Y = np.meshgrid(dx * (np.arange(nx) + 0.5), dy * (np.arange(ny) + 0.5), indexing="xy")
    gx, gy = H.spectral_derivatives(np.sin(3 * X + 2 * Y), nx, ny, dx, dy)
    assert np.abs(gx - 3 * np.cos(3 * X + 2 * Y)).max() < 1e-12
    assert np.abs(gy - 2 * np.cos(3 * X + 2 * Y)).max() < 1e-12


def _shear_state(n, A, B, C, p0, gamma=1.4):
    L = 2.0 * np.pi
    h = L / n
    X, Y = np.meshgrid(h * (np.arange(n) + 0.5), h * (np.arange(n) + 0.5), indexing="xy")
    u = A * np.sin(Y) + C * np.sin(3 * X)
    v = B * np.sin(X)
    return conservative(np.ones_like(u), u, v, np.full_like(u, p0), Gas(gamma=gamma)), h


def test_diagnostics_closed_forms():
    """rho = 1, u = A sin y + C sin 3x, v = B sin x, p = p0: every diagnostic
    in closed form (moments, Mt, divergence, vorticity, low-k fraction with the
    |k| = 3 mode outside the band, drag and cooling powers)."""
    n, A, B, C, p0, g = 32, 0.3, 0.2, 0.1, 0.8, 1.4
    q, h = _shear_state(n, A, B, C, p0, g)
    gas = Gas(gamma=g)
    L = 2.0 * np.pi
    frac = H.large_scale_kinetic_fraction(q, gas, n, n, h, h, 2.0)
    assert frac == pytest.approx((A**2 + B**2) / (A**2 + B**2 + C**2), rel=1e-12)
    drag = H.source_power(H.drag_source(q, gas, n, n, h, h, 0.1, 2.0))
    assert drag == pytest.approx(-0.1 * (A**2 + B**2) / 2, rel=1e-12)
    cool = H.source_power(H.cooling_source(q, gas, 5.0))
    assert cool == pytest.approx(-(p0 - 1 / g) / (g - 1) / 5.0, rel=1e-12)
    d = H.compute_diagnostics(q, gas, n, n, h, h, L, L, initial_mass=L * L * 0.5)
    uu, vv = (A**2 + C**2) / 2, B**2 / 2
    assert d["kinetic_energy"] == pytest.approx((uu + vv) / 2, rel=1e-12)
    assert d["Kx"] == pytest.approx(uu / 2, rel=1e-12)
    assert d["Ky"] == pytest.approx(vv / 2, rel=1e-12)
    assert d["A_K"] == pytest.approx(abs(uu - vv) / (uu + vv), rel=1e-12)
    assert abs(d["C_uv"]) < 1e-14
    rms = np.sqrt(uu + vv)
    assert d["rms_velocity"] == pytest.approx(rms, rel=1e-12)
    assert d["turbulent_mach"] == pytest.approx(rms / np.sqrt(g * p0), rel=1e-12)
    assert d["mean_density"] == pytest.approx(1.0, rel=1e-14)
    assert d["mean_pressure"] == pytest.approx(p0, rel=1e-12)
    assert d["divergence_rms"] == pytest.approx(3 * C / np.sqrt(2), rel=1e-12)  # du/dx = 3C cos 3x
    # vorticity = dv/dx - du/dy = B cos x - A cos y - 0
    assert d["vorticity_rms"] == pytest.approx(np.sqrt((A**2 + B**2) / 2), rel=1e-12)
    assert d["mass"] == pytest.approx(L * L, rel=1e-14)
    assert d["mass_error"] == pytest.approx(L * L * 0.5, rel=1e-12)
    assert d["mach_control_mach"] == d["turbulent_mach"] and d["P_in"] == 0.0
    f = H.snapshot_fields(q, gas, n, n, h, h)
    X, Y = np.meshgrid(h * (np.arange(n) + 0.5), h * (np.arange(n) + 0.5), indexing="xy")
    assert np.abs(f["vorticity"] - (B * np.cos(X) - A * np.cos(Y))).max() < 1e-12
    assert np.abs(f["divergence"] - 3 * C * np.cos(3 * X)).max() < 1e-12
