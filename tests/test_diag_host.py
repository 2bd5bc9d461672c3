"""Host composition of the HIT diagnostics (no GPU): the dictionary built
from the device's sums (HITX_DIAG_* slots) equals the oracle's
compute_diagnostics when the sums are formed on the CPU from the same
fields, and the loop helpers keep the reference's formats."""

import numpy as np
import pytest

from oracle import hit as OH
from oracle.equations import Gas, primitives


def sums_of(q, gas, n, h, kmax, coef, mask):
    """The HITX_DIAG_* sums of q, formed with the oracle's FFT fields."""
    from hit_solver_amd import _lib

    rho, u, v, p = primitives(q, gas)
    div, vort = OH.divergence_and_vorticity(u, v, n, n, h, h)
    uf, vf = u - np.sum(u) / u.size, v - np.sum(v) / v.size
    ul = OH.band_filter(uf, n, n, h, h, 0.0, kmax)
    vl = OH.band_filter(vf, n, n, h, h, 0.0, kmax)
    drag = OH.drag_source(q, gas, n, n, h, h, coef, kmax)
    s = np.zeros(_lib.DIAG_COUNT)
    s[_lib.DIAG_RHO], s[_lib.DIAG_U], s[_lib.DIAG_V], s[_lib.DIAG_P] = rho.sum(), u.sum(), v.sum(), p.sum()
    s[_lib.DIAG_C] = np.sqrt(gas.gamma * p / rho).sum()
    s[_lib.DIAG_U2V2] = (u**2 + v**2).sum()
    s[_lib.DIAG_RUU], s[_lib.DIAG_RVV] = (rho * uf**2).sum(), (rho * vf**2).sum()
    s[_lib.DIAG_RUV] = (rho * uf * vf).sum()
    s[_lib.DIAG_RKE] = (rho * (uf**2 + vf**2)).sum()
    s[_lib.DIAG_DIV2], s[_lib.DIAG_VORT2] = (div**2).sum(), (vort**2).sum()
    s[_lib.DIAG_LOW] = (rho * (ul**2 + vl**2)).sum()
    s[_lib.DIAG_DRAG_E] = drag[3].sum()
    s[_lib.DIAG_MASK] = mask.sum()
    return s


def test_dictionary_from_sums_matches_oracle():
    from hit_solver_amd import hit2d
    from hit_solver_amd.domain import Domain2D
    from hit_solver_amd.equations import EulerEquation2D

    n = 48
    p = OH.HITParams(nx=n, ny=n, target_mach=0.25, initial_kmin=3, initial_kmax=5)
    q = OH.initial_state(p, np.random.default_rng(7))
    q[0] *= 1.0 + 0.05 * np.sin(np.linspace(0, 2 * np.pi, n))[None, :]
    gas, eq = Gas(gamma=1.4), EulerEquation2D(gamma=1.4)
    d = Domain2D(0.0, 2 * np.pi, 0.0, 2 * np.pi, n, n)
    h = d.dx
    mask = np.zeros((n, n), bool)
    mask[3:9, 10:20] = True
    s = sums_of(q, gas, n, h, 2.0, 0.1, mask)
    N = n * n
    frac = hit2d._kinetic_fraction(s, N)
    assert frac == pytest.approx(OH.large_scale_kinetic_fraction(q, gas, n, n, h, h, 2.0), rel=1e-12)
    cool = hit2d._cooling_power(s, N, eq, 5.0, None)
    assert cool == pytest.approx(OH.source_power(OH.cooling_source(q, gas, 5.0)), rel=1e-12)
    assert hit2d._cooling_power(s, N, eq, None, None) == 0.0
    info = {"injected_power": 1e-3, "Fxx": 0.5, "Fyy": 0.4, "Fxy": 0.01, "A_F": 0.1, "alpha": 0.9}
    got = hit2d._diagnostics_from_sums(s, d, 38.0, info, float(s[13] / N), cool, frac, float(mask.mean()))
    want = OH.compute_diagnostics(q, gas, n, n, h, h, d.x_max - d.x_min, d.y_max - d.y_min, 38.0, info,
                                  OH.source_power(OH.drag_source(q, gas, n, n, h, h, 0.1, 2.0)), cool, frac,
                                  float(np.mean(mask)))
    assert list(got) == list(want)
    for k in want:
        assert got[k] == pytest.approx(want[k], rel=1e-12, abs=1e-16), k


def test_reference_key_order_and_print_line(capsys):
    """The diagnostics dictionary keys, in the reference's order
    (OOP/hit2d.py:495-547), and the progress line fields (:618-637)."""
    from hit_solver_amd import _lib, hit2d
    from hit_solver_amd.domain import Domain2D

    d = Domain2D(0.0, 1.0, 0.0, 1.0, 8, 8)
    s = np.ones(_lib.DIAG_COUNT)
    got = hit2d._diagnostics_from_sums(s, d)
    assert list(got) == [
        "kinetic_energy", "Kx", "Ky", "A_K", "C_uv", "rms_velocity", "turbulent_mach", "mean_density",
        "mean_pressure", "divergence_rms", "vorticity_rms", "mass", "mass_error", "large_scale_kinetic_fraction",
        "drag_power", "cooling_power", "weno_fraction", "P_in", "Fxx", "Fyy", "Fxy", "A_F", "forcing_alpha",
        "forcing_target_power", "mach_control_mach", "mach_control_target", "mach_control_power_desired",
    ]
    hit2d._print_diagnostics(5, 0.5, 1e-3, got)
    line = capsys.readouterr().out
    assert line.startswith("step=     5, t=0.500000, dt=1.000e-03, KE=") and "mass_err=" in line
    rec = hit2d._diagnostic_record(5, 0.5, 1e-3, got)
    assert list(rec)[:3] == ["step", "time", "dt"] and rec["step"] == 5.0


def test_warnings_and_nonfinite(capsys):
    from hit_solver_amd import _lib, hit2d
    from hit_solver_amd.domain import Domain2D
    from hit_solver_amd.hit2d import HIT2DConfig

    d = Domain2D(0.0, 1.0, 0.0, 1.0, 8, 8)
    good = hit2d._diagnostics_from_sums(np.ones(_lib.DIAG_COUNT), d, weno_fraction=0.9)
    hit2d._warn_diagnostic_state(good, [], HIT2DConfig())
    assert "WENO is active over a large domain fraction" in capsys.readouterr().out
    bad = dict(good, kinetic_energy=float("nan"))
    with pytest.raises(FloatingPointError):
        hit2d._warn_diagnostic_state(bad, [], HIT2DConfig())


def test_run_config_json(tmp_path):
    import json

    from hit_solver_amd import hit2d
    from hit_solver_amd.hit2d import HIT2DConfig

    path = hit2d.save_run_config(tmp_path / "out", HIT2DConfig.isotropic_128(output_dir=tmp_path / "out"))
    conf = json.loads(path.read_text())
    assert conf["forcing_type"] == "solenoidal_shell_ou" and conf["hyperviscosity_policy"] == "compact_nodes_only"
    assert conf["output_dir"] == str(tmp_path / "out") and conf["tfinal"] == 15.0
    assert list(conf) == sorted(conf)
