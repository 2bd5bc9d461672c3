"""HIT post-processing on snapshot planes: shell spectra and two-point
correlations.

TEST INFRASTRUCTURE (part of the oracle package): the checker of the device
spectra and correlations, never called by the product.

Restates reference OOP/postprocess/spectra.py:48-149, 203-225
(HIT2DSpectra.compute, _fluctuations, _cutoff_pileup_ratio) and
OOP/postprocess/two_point_correlation.py:107-139, 141-287, 363-485
(autocorrelation, cross_correlation, compute, integral_length_scale,
taylor_microscale, _velocity_fluctuations, _half_separations,
_directional_curves, _normalize_curve, _divergence_and_vorticity) on arrays
instead of snapshot files, with NumPy FFTs as the reference uses them.

Pinned by closed forms (tests/test_oracle_pins.py): a single Fourier mode
lands in the one shell rint(|k|/dk) with energy A^2/4 per component; the
shell sums add up to the field's energy (Parseval) when no energy lies beyond
the last shell; enstrophy of a mode is |k|^2 times its energy; the
autocorrelation of cos(k x) is cos(k r), the cross-correlation of cos and
sin is -sin(k r)/2 (unnormalised); the integral scale of a linear ramp and
the Taylor microscale of 1 - a r^2 are known.
"""

from __future__ import annotations

import numpy as np


def fluctuations(rho, u, v, kind="reynolds"):
    """spectra.py:203-214 / two_point_correlation.py:406-422."""
    if kind == "favre":
        rho_mean = float(np.mean(rho))
        u_mean = float(np.mean(rho * u) / rho_mean)
        v_mean = float(np.mean(rho * v) / rho_mean)
    else:
        u_mean = float(np.mean(u))
        v_mean = float(np.mean(v))
    return u - u_mean, v - v_mean


def shell_geometry(nx, ny, dx, dy):
    """Wavenumbers, shell index of every mode, dk and the last shell
    (spectra.py:97-103)."""
    kx = 2.0 * np.pi * np.fft.fftfreq(nx, d=dx)
    ky = 2.0 * np.pi * np.fft.fftfreq(ny, d=dy)
    KX, KY = np.meshgrid(kx, ky, indexing="xy")
    kmag = np.sqrt(KX**2 + KY**2)
    dk = min(2.0 * np.pi / (nx * dx), 2.0 * np.pi / (ny * dy))
    return KX, KY, np.rint(kmag / dk).astype(int), dk, min(nx, ny) // 2


def snapshot_spectra(rho, u, v, dx, dy, kind="reynolds"):
    """One snapshot's shell spectra (spectra.py:90-131): returns (k, energy,
    enstrophy) with k = dk * arange(max_shell + 1)."""
    ny, nx = rho.shape
    uf, vf = fluctuations(rho, u, v, kind)
    KX, KY, shell, dk, max_shell = shell_geometry(nx, ny, dx, dy)
    wu = np.fft.fft2(np.sqrt(rho) * uf)
    wv = np.fft.fft2(np.sqrt(rho) * vf)
    norm = float((nx * ny) ** 2)
    energy_mode = 0.5 * (np.abs(wu) ** 2 + np.abs(wv) ** 2) / norm
    uh = np.fft.fft2(uf)
    vh = np.fft.fft2(vf)
    omega_hat = 1j * (KX * vh - KY * uh)
    enstrophy_mode = 0.5 * np.abs(omega_hat) ** 2 / norm
    e = np.bincount(shell.ravel(), weights=energy_mode.ravel(), minlength=max_shell + 1)[: max_shell + 1]
    z = np.bincount(shell.ravel(), weights=enstrophy_mode.ravel(), minlength=max_shell + 1)[: max_shell + 1]
    return dk * np.arange(max_shell + 1), e, z


def cutoff_pileup_ratio(values):
    """spectra.py:216-225."""
    nonzero = values[1:]
    width = max(3, nonzero.size // 10)
    if nonzero.size < 2 * width:
        return 0.0
    reference = float(np.mean(nonzero[-2 * width: -width]))
    cutoff = float(np.mean(nonzero[-width:]))
    if float(np.sum(nonzero[-width:])) < 1.0e-8 * float(np.sum(nonzero)):
        return 0.0
    return cutoff / max(reference, np.finfo(float).tiny)


def spectra(snapshots, dx, dy, kind="reynolds"):
    """spectra.py:84-149 over a list of (rho, u, v): mean and std of the
    shell spectra and the pile-up ratios."""
    es, zs, k = [], [], None
    for rho, u, v in snapshots:
        k, e, z = snapshot_spectra(rho, u, v, dx, dy, kind)
        es.append(e)
        zs.append(z)
    E, Z = np.stack(es), np.stack(zs)
    em, zm = np.mean(E, axis=0), np.mean(Z, axis=0)
    return dict(k=k, energy_mean=em, energy_std=np.std(E, axis=0), enstrophy_mean=zm,
                enstrophy_std=np.std(Z, axis=0), high_k_energy_ratio=cutoff_pileup_ratio(em),
                high_k_enstrophy_ratio=cutoff_pileup_ratio(zm))


def autocorrelation(field, normalize=True):
    """two_point_correlation.py:107-121."""
    values = np.asarray(field, dtype=float)
    spectrum = np.fft.fft2(values)
    corr = np.fft.ifft2(spectrum * np.conj(spectrum)).real / values.size
    if normalize:
        zero = float(corr[0, 0])
        if abs(zero) > np.finfo(float).eps:
            corr = corr / zero
        else:
            corr = np.zeros_like(corr)
            corr[0, 0] = 1.0
    return corr


def cross_correlation(a, b, normalize=False):
    """two_point_correlation.py:123-139."""
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    corr = np.fft.ifft2(np.fft.fft2(a) * np.conj(np.fft.fft2(b))).real / a.size
    if normalize:
        scale = np.sqrt(float(np.mean(a**2) * np.mean(b**2)))
        corr = corr / scale if scale > np.finfo(float).eps else np.zeros_like(corr)
    return corr


def divergence_and_vorticity(u, v, dx, dy):
    """two_point_correlation.py:472-485 (second-order central differences)."""
    du_dx = (np.roll(u, -1, axis=1) - np.roll(u, 1, axis=1)) / (2.0 * dx)
    du_dy = (np.roll(u, -1, axis=0) - np.roll(u, 1, axis=0)) / (2.0 * dy)
    dv_dx = (np.roll(v, -1, axis=1) - np.roll(v, 1, axis=1)) / (2.0 * dx)
    dv_dy = (np.roll(v, -1, axis=0) - np.roll(v, 1, axis=0)) / (2.0 * dy)
    return du_dx + dv_dy, dv_dx - du_dy


def normalize_curve(values):
    zero = float(values[0])
    if abs(zero) <= np.finfo(float).eps:
        return np.full_like(values, np.nan)
    return values / zero


def snapshot_correlations(rho, u, v, p, dx, dy, kind="reynolds", vorticity=None, divergence=None):
    """The per-snapshot maps and curves of two_point_correlation.py:162-232, 446-470."""
    ny, nx = rho.shape
    rx, ry = dx * np.arange(nx // 2 + 1), dy * np.arange(ny // 2 + 1)
    uf, vf = fluctuations(rho, u, v, kind)
    rf = rho - float(np.mean(rho))
    pf = p - float(np.mean(p))
    if vorticity is None or divergence is None:
        divergence, vorticity = divergence_and_vorticity(u, v, dx, dy)
    wf = vorticity - float(np.mean(vorticity))
    tf = divergence - float(np.mean(divergence))
    out = dict(Ruu_2d=autocorrelation(uf), Rvv_2d=autocorrelation(vf), Ruv_2d=cross_correlation(uf, vf, True),
               Rrho_2d=autocorrelation(rf), Rp_2d=autocorrelation(pf), Romega_2d=autocorrelation(wf),
               Rtheta_2d=autocorrelation(tf), Rrhop_2d=cross_correlation(rf, pf, True),
               Rp_theta_2d=cross_correlation(pf, tf, True))
    nxh, nyh = rx.size, ry.size
    Ruu, Rvv = out["Ruu_2d"], out["Rvv_2d"]
    curves = dict(Ruu_x=Ruu[0, :nxh], Ruu_y=Ruu[:nyh, 0], Rvv_x=Rvv[0, :nxh], Rvv_y=Rvv[:nyh, 0])
    common = min(curves["Ruu_x"].size, curves["Rvv_y"].size)
    out.update(curves)
    out["r"] = 0.5 * (rx[:common] + ry[:common])
    out["R_LL"] = normalize_curve(0.5 * (curves["Ruu_x"][:common] + curves["Rvv_y"][:common]))
    out["R_NN"] = normalize_curve(0.5 * (curves["Rvv_x"][:common] + curves["Ruu_y"][:common]))
    out["rx"], out["ry"] = rx, ry
    return out


def integral_length_scale(r, R_LL):
    """two_point_correlation.py:363-385 (without the no-crossing warning)."""
    if r.size < 2 or R_LL.size < 2:
        return float("nan")
    negative = np.flatnonzero(R_LL < 0.0)
    if negative.size == 0:
        return float(np.trapezoid(R_LL, r))
    first = int(negative[0])
    if first == 0:
        return 0.0
    rs, Rs = list(r[:first]), list(R_LL[:first])
    r0, r1, R0, R1 = r[first - 1], r[first], R_LL[first - 1], R_LL[first]
    if R1 != R0:
        rs.append(float(r0 - R0 * (r1 - r0) / (R1 - R0)))
        Rs.append(0.0)
    return float(np.trapezoid(np.asarray(Rs), np.asarray(rs)))


def taylor_microscale(r, R_LL, max_points=6):
    """two_point_correlation.py:387-404 (NaN where the reference warns)."""
    if r.size < 4 or R_LL.size < 4:
        return float("nan")
    count = min(max_points, r.size - 1)
    x = r[1: count + 1] ** 2
    y = 1.0 - R_LL[1: count + 1]
    ok = np.isfinite(x) & np.isfinite(y) & (x > 0.0)
    if np.count_nonzero(ok) < 3:
        return float("nan")
    slope = float(np.dot(x[ok], y[ok]) / np.dot(x[ok], x[ok]))
    if slope <= 0.0 or not np.isfinite(slope):
        return float("nan")
    return float(np.sqrt(1.0 / (2.0 * slope)))
