"""Closed-form pins of the oracle's spectra and correlations
(oracle/postprocess.py), each chosen so that a plausible slip fails it:

* a single Fourier mode lands in shell rint(|k|/dk) -- (2, 2) has
  |k|/dk = 2.83, so floor binning would put it in shell 2, not 3 -- with
  energy A^2/4 for a cos mode of amplitude A in u (the 1/N^2 normalisation
  and the factor 1/2), and enstrophy |k|^2 times that;
* Parseval: the shell energies of a band-limited field sum to
  0.5 mean(rho (u'^2 + v'^2)) (the sqrt(rho) weighting);
* Favre and Reynolds fluctuations differ by the density-weighted mean;
* the autocorrelation of cos(k x) is cos(k r), and the unnormalised
  cross-correlation of (cos x, sin x) is mean(cos(x + r) sin x) = -sin(r)/2
  (which side of the product is conjugated);
* the integral scale of R = 1 - r/L on a grid crossing zero at L is L/2;
  the Taylor microscale of 1 - a r^2 is sqrt(1/(2a)).
"""

import numpy as np
import pytest

from oracle import postprocess as PP


def grid(nx, ny, L=2 * np.pi):
    dx, dy = L / nx, L / ny
    x = np.arange(nx) * dx
    y = np.arange(ny) * dy
    X, Y = np.meshgrid(x, y, indexing="xy")
    return X, Y, dx, dy


@pytest.mark.parametrize("mx,my,shell", [(1, 0, 1), (2, 2, 3), (3, 1, 3), (0, 5, 5)])
def test_single_mode_shell(mx, my, shell):
    nx, ny = 32, 24
    X, Y, dx, dy = grid(nx, ny)
    A = 0.7
    u = A * np.cos(mx * X + my * Y)
    v = np.zeros_like(u)
    k, e, z = PP.snapshot_spectra(np.ones_like(u), u, v, dx, dy)
    assert int(np.argmax(e)) == shell
    np.testing.assert_allclose(e[shell], A * A / 4, rtol=1e-12)
    assert np.sum(np.abs(np.delete(e, shell))) < 1e-25
    # u only: omega_hat = -i ky u_hat, so Z = ky^2 E (a my-only factor)
    np.testing.assert_allclose(z[shell], my * my * A * A / 4, rtol=1e-12, atol=1e-25)
    assert k[shell] == pytest.approx(shell * 1.0)


def test_parseval_and_density_weighting():
    nx, ny = 40, 36
    X, Y, dx, dy = grid(nx, ny)
    rng = np.random.default_rng(3)
    u = sum(rng.normal() * np.cos(a * X + b * Y + rng.uniform(0, 6)) for a, b in [(1, 2), (3, 0), (2, 5)])
    v = sum(rng.normal() * np.sin(a * X + b * Y + rng.uniform(0, 6)) for a, b in [(0, 1), (4, 3)])
    rho = 1.0 + 0.3 * np.cos(X - 2 * Y)
    _, e, _ = PP.snapshot_spectra(rho, u, v, dx, dy)
    uf, vf = PP.fluctuations(rho, u, v)
    assert np.sum(e) == pytest.approx(0.5 * np.mean(rho * (uf**2 + vf**2)), rel=1e-12)


def test_favre_fluctuations():
    rho = np.array([[1.0, 2.0], [3.0, 4.0]])
    u = np.array([[1.0, 0.0], [0.0, 1.0]])
    v = np.zeros_like(u)
    uf, _ = PP.fluctuations(rho, u, v, "favre")
    np.testing.assert_allclose(uf, u - (1.0 + 4.0) / 10.0)
    uf, _ = PP.fluctuations(rho, u, v, "reynolds")
    np.testing.assert_allclose(uf, u - 0.5)


def test_correlation_closed_forms():
    nx, ny = 32, 16
    X, Y, dx, dy = grid(nx, ny)
    R = PP.autocorrelation(np.cos(3 * X))
    np.testing.assert_allclose(R[0, : nx // 2 + 1], np.cos(3 * X[0, : nx // 2 + 1]), atol=1e-13)
    np.testing.assert_allclose(R[:, 0], 1.0, atol=1e-13)
    C = PP.cross_correlation(np.cos(X), np.sin(X))
    np.testing.assert_allclose(C[0], -np.sin(X[0]) / 2, atol=1e-14)


def test_length_scales():
    r = np.linspace(0.0, 2.0, 21)
    L = 1.0
    assert PP.integral_length_scale(r, 1.0 - r / L) == pytest.approx(L / 2, rel=1e-12)
    a = 0.3
    assert PP.taylor_microscale(r, 1.0 - a * r * r) == pytest.approx(np.sqrt(1.0 / (2 * a)), rel=1e-12)


def test_pileup_ratio():
    flat = np.ones(40)
    assert PP.cutoff_pileup_ratio(flat) == pytest.approx(1.0)
    rising = np.concatenate([[0.0], np.ones(36), 3.0 * np.ones(3)])  # last tenth (3 shells) 3x the tenth before
    assert PP.cutoff_pileup_ratio(rising) == pytest.approx(3.0)
    assert PP.cutoff_pileup_ratio(np.ones(4)) == 0.0


def test_longitudinal_transverse_composition():
    """u = cos x, v = cos y on a square grid: R_uu(rx, 0) = cos rx and
    R_vv(0, ry) = cos ry, so R_LL = cos r; R_vv(rx, 0) = R_uu(0, ry) = 1, so
    R_NN = 1; r = (rx + ry)/2 over the half separations."""
    n = 24
    X, Y, dx, dy = grid(n, n)
    rho = np.ones_like(X)
    out = PP.snapshot_correlations(rho, np.cos(X), np.cos(Y), 1.0 + 0.1 * np.cos(X + Y), dx, dy)
    r = dx * np.arange(n // 2 + 1)
    np.testing.assert_allclose(out["r"], r)
    np.testing.assert_allclose(out["R_LL"], np.cos(r), atol=1e-13)
    np.testing.assert_allclose(out["R_NN"], 1.0, atol=1e-13)
    np.testing.assert_allclose(out["Ruv_2d"], 0.0, atol=1e-13)
