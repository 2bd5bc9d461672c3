"""HIT post-processing on the device: shell spectra and two-point correlations.

API of the reference's OOP/postprocess/spectra.py (HIT2DSpectra,
SpectrumResults2D) and OOP/postprocess/two_point_correlation.py
(TwoPointCorrelation2D, CorrelationResults2D).  The grid work -- mean
removal, the 2-D transforms, the shell sums, the spectral products, the
inverse transforms and the normalisations -- runs in the HIP library
(hitx_post_fields, hitx_dft2, hitx_shell_spectra, hitx_correlations); the
host reads the snapshot files, averages the per-snapshot results (a few
arrays) and forms the 1-D curves and scalars from them, as the reference
does after its FFTs.  Figures are left to the reference's plot methods,
which read the saved .npz files.
"""

from __future__ import annotations

import ctypes
import warnings
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np
import torch

from ._lib import check, lib

DEV = "cuda"


def _p(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _plane(a):
    """A float64 device plane from a numpy array or tensor."""
    if isinstance(a, torch.Tensor):
        return a.to(device=DEV, dtype=torch.float64).contiguous()
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(DEV)


class _Workspace:
    """Device workspace of the post-processing calls, grown on demand."""

    def __init__(self):
        self.buf = None

    def get(self, nx, ny, planes):
        need = int(lib().hitx_post_workspace_bytes(int(nx), int(ny), int(planes)))
        if need == 0:
            raise ValueError("grid too small for post-processing")
        if self.buf is None or self.buf.numel() < need:
            self.buf = torch.empty(need, dtype=torch.uint8, device=DEV)
        return self.buf


_WS = _Workspace()


def fluctuation_planes(rho, u, v, dx, dy, p=None, vorticity=None, divergence=None, fluctuation_type="reynolds"):
    """Device planes [8, ny, nx]: sqrt(rho) u', sqrt(rho) v', u', v', rho',
    p', omega', theta' (p' zero when p is None; omega, theta from central
    differences of u, v when not given)."""
    if fluctuation_type not in ("reynolds", "favre"):
        raise ValueError("fluctuation_type must be 'reynolds' or 'favre'")
    r, a, b = _plane(rho), _plane(u), _plane(v)
    ny, nx = r.shape
    pp = _plane(p) if p is not None else None
    w = _plane(vorticity) if vorticity is not None and divergence is not None else None
    t = _plane(divergence) if w is not None else None
    out = torch.zeros((8, ny, nx), dtype=torch.float64, device=DEV)
    ws = _WS.get(nx, ny, 1)
    check(lib().hitx_post_fields(nx, ny, float(dx), float(dy), _p(r), _p(a), _p(b), _p(pp), _p(w), _p(t),
                                 int(fluctuation_type == "favre"), _p(out), _p(ws), ws.numel(), _stream()))
    return out


def dft2(planes, inverse=False):
    """numpy fft2 / ifft2 of real or complex device planes [count, ny, nx]
    (complex result as [count, ny, nx, 2] float64)."""
    x = planes
    cplx = x.dim() == 4
    count, ny, nx = x.shape[:3]
    out = torch.empty((count, ny, nx, 2), dtype=torch.float64, device=DEV)
    ws = _WS.get(nx, ny, count)
    check(lib().hitx_dft2(nx, ny, count, _p(x.contiguous()), int(cplx), int(inverse), _p(out), _p(ws), ws.numel(),
                          _stream()))
    return out


def snapshot_spectra(rho, u, v, dx, dy, fluctuation_type="reynolds"):
    """One snapshot's shell spectra (reference spectra.py:84-131): (k,
    energy, enstrophy) as numpy arrays."""
    f = fluctuation_planes(rho, u, v, dx, dy, fluctuation_type=fluctuation_type)
    ny, nx = f.shape[1:]
    hat = dft2(f[:4])
    nshell = min(nx, ny) // 2 + 1
    e = torch.empty(nshell, dtype=torch.float64, device=DEV)
    z = torch.empty(nshell, dtype=torch.float64, device=DEV)
    ws = _WS.get(nx, ny, 1)
    check(lib().hitx_shell_spectra(nx, ny, float(dx), float(dy), _p(hat), nshell, _p(e), _p(z), _p(ws), ws.numel(),
                                   _stream()))
    dk = min(2.0 * np.pi / (nx * dx), 2.0 * np.pi / (ny * dy))
    return dk * np.arange(nshell), e.cpu().numpy(), z.cpu().numpy()


@dataclass
class SpectrumResults2D:
    k: np.ndarray
    energy_mean: np.ndarray
    energy_std: np.ndarray
    enstrophy_mean: np.ndarray
    enstrophy_std: np.ndarray
    number_of_snapshots: int
    selected_time_interval: tuple
    high_k_energy_ratio: float
    high_k_enstrophy_ratio: float


def _load_npz(path, names):
    with np.load(path, allow_pickle=False) as data:
        return {k: np.asarray(data[k]) for k in names if k in data}


class HIT2DSpectra:
    """Shell-averaged kinetic-energy and enstrophy spectra from HIT2D
    snapshots (reference OOP/postprocess/spectra.py:24-149, same arguments,
    errors and warnings)."""

    def __init__(self, run_dir, fluctuation_type: str = "reynolds", start_time=None, end_time=None, stride: int = 1,
                 output_dir=None):
        self.run_dir = Path(run_dir)
        self.fluctuation_type = fluctuation_type.lower()
        if self.fluctuation_type not in {"reynolds", "favre"}:
            raise ValueError("fluctuation_type must be 'reynolds' or 'favre'")
        if stride <= 0:
            raise ValueError("stride must be positive")
        self.start_time, self.end_time, self.stride = start_time, end_time, stride
        self.output_dir = Path(output_dir) if output_dir is not None else self.run_dir / "postprocess"
        self.results = None

    def _selected(self):
        paths = sorted(self.run_dir.glob("hit2d_step*.npz"))
        if not paths:
            raise FileNotFoundError(f"No HIT2D snapshots found in {self.run_dir}")
        sel = []
        for path in paths:
            d = _load_npz(path, ("time", "x", "y", "rho", "u", "v"))
            t = float(d["time"])
            if self.start_time is not None and t < self.start_time:
                continue
            if self.end_time is not None and t > self.end_time:
                continue
            sel.append((t, d))
        sel = sel[:: self.stride]
        if not sel:
            raise ValueError("No snapshots are available in the selected spectral interval")
        if len(sel) < 5:
            warnings.warn("Few snapshots selected for spectral averaging", RuntimeWarning)
        return sel

    def compute(self) -> SpectrumResults2D:
        sel = self._selected()
        snaps = []
        for t, d in sel:
            x, y = np.asarray(d["x"], float), np.asarray(d["y"], float)
            snaps.append((t, d["rho"], d["u"], d["v"], float(x[1] - x[0]), float(y[1] - y[0])))
        return self.compute_arrays(snaps)

    def compute_arrays(self, snapshots) -> SpectrumResults2D:
        """The same from in-memory snapshots: (time, rho, u, v, dx, dy) tuples
        (numpy arrays or device tensors)."""
        es, zs, k = [], [], None
        for _t, rho, u, v, dx, dy in snapshots:
            k, e, z = snapshot_spectra(rho, u, v, dx, dy, self.fluctuation_type)
            es.append(e)
            zs.append(z)
        E, Z = np.stack(es), np.stack(zs)
        em, zm = np.mean(E, axis=0), np.mean(Z, axis=0)
        he, hz = self._cutoff_pileup_ratio(em), self._cutoff_pileup_ratio(zm)
        if he > 2.0:
            warnings.warn("Kinetic-energy spectrum may be piling up near k_max", RuntimeWarning)
        if hz > 2.0:
            warnings.warn("Enstrophy spectrum may be piling up near k_max", RuntimeWarning)
        times = [s[0] for s in snapshots]
        self.results = SpectrumResults2D(k=k, energy_mean=em, energy_std=np.std(E, axis=0), enstrophy_mean=zm,
                                         enstrophy_std=np.std(Z, axis=0), number_of_snapshots=len(snapshots),
                                         selected_time_interval=(float(times[0]), float(times[-1])),
                                         high_k_energy_ratio=he, high_k_enstrophy_ratio=hz)
        return self.results

    def save(self, output_path=None) -> Path:
        r = self._require_results()
        path = Path(output_path) if output_path is not None else self.run_dir / "spectra_diagnostics.npz"
        path.parent.mkdir(parents=True, exist_ok=True)
        np.savez_compressed(path, k=r.k, energy_mean=r.energy_mean, energy_std=r.energy_std,
                            enstrophy_mean=r.enstrophy_mean, enstrophy_std=r.enstrophy_std,
                            number_of_snapshots=np.asarray(r.number_of_snapshots),
                            selected_time_interval=np.asarray(r.selected_time_interval),
                            high_k_energy_ratio=np.asarray(r.high_k_energy_ratio),
                            high_k_enstrophy_ratio=np.asarray(r.high_k_enstrophy_ratio),
                            fluctuation_type=np.asarray(self.fluctuation_type))
        return path

    @staticmethod
    def _cutoff_pileup_ratio(values) -> float:
        """Mean of the last tenth of the spectrum over the tenth before it
        (0 when there is too little or no energy there)."""
        tail = values[1:]
        width = max(3, tail.size // 10)
        if tail.size < 2 * width:
            return 0.0
        ref = float(np.mean(tail[-2 * width: -width]))
        cut = float(np.mean(tail[-width:]))
        if float(np.sum(tail[-width:])) < 1.0e-8 * float(np.sum(tail)):
            return 0.0
        return cut / max(ref, np.finfo(float).tiny)

    def _require_results(self):
        if self.results is None:
            raise RuntimeError("Call compute() before save() or plot()")
        return self.results


# ---------------------------------------------------------------- correlations
# planes of fluctuation_planes: 2 u', 3 v', 4 rho', 5 p', 6 omega', 7 theta'
_MAPS = (  # name, plane a, plane b, normalisation (1 auto; 2 cross by the two auto maps named)
    ("Ruu_2d", 2, 2, (1, 0, 0)), ("Rvv_2d", 3, 3, (1, 0, 0)), ("Ruv_2d", 2, 3, (2, 0, 1)),
    ("Rrho_2d", 4, 4, (1, 0, 0)), ("Rp_2d", 5, 5, (1, 0, 0)), ("Romega_2d", 6, 6, (1, 0, 0)),
    ("Rtheta_2d", 7, 7, (1, 0, 0)), ("Rrhop_2d", 4, 5, (2, 3, 4)), ("Rp_theta_2d", 5, 7, (2, 4, 6)),
)


def correlation_maps(hat, maps):
    """Device correlation maps of transformed planes: maps is a list of (a,
    b, (mode, ref0, ref1)); returns [len(maps), ny, nx] float64."""
    nplanes, ny, nx = hat.shape[:3]
    out = torch.empty((len(maps), ny, nx), dtype=torch.float64, device=DEV)
    pairs = np.ascontiguousarray(np.array([[a, b] for a, b, _ in maps], dtype=np.int32))
    norms = np.ascontiguousarray(np.array([list(n) for _, _, n in maps], dtype=np.int32))
    ws = _WS.get(nx, ny, 3 * len(maps))
    check(lib().hitx_correlations(nx, ny, len(maps), _p(hat), nplanes, pairs.ctypes.data_as(ctypes.c_void_p),
                                  norms.ctypes.data_as(ctypes.c_void_p), _p(out), _p(ws), ws.numel(), _stream()))
    return out


@dataclass
class CorrelationResults2D:
    r: np.ndarray
    rx: np.ndarray
    ry: np.ndarray
    Ruu_2d: np.ndarray
    Rvv_2d: np.ndarray
    Ruv_2d: np.ndarray
    Rrhop_2d: np.ndarray | None
    Rp_theta_2d: np.ndarray | None
    Rrho_2d: np.ndarray | None
    Rp_2d: np.ndarray | None
    Romega_2d: np.ndarray | None
    Rtheta_2d: np.ndarray | None
    Ruu_x: np.ndarray
    Ruu_y: np.ndarray
    Rvv_x: np.ndarray
    Rvv_y: np.ndarray
    R_LL: np.ndarray
    R_NN: np.ndarray
    R_LL_std: np.ndarray
    R_NN_std: np.ndarray
    L_integral: float
    lambda_taylor: float
    fluctuation_type: str
    number_of_snapshots: int
    metadata: dict = field(default_factory=dict)


class TwoPointCorrelation2D:
    """Periodic two-point correlations of 2-D HIT snapshots (reference
    OOP/postprocess/two_point_correlation.py:42-485, same arguments and
    results; maps computed on the device)."""

    def __init__(self, run_dir=None, snapshot_paths=None, fluctuation_type: str = "reynolds", Lx=None, Ly=None,
                 output_dir=None):
        self.run_dir = Path(run_dir) if run_dir is not None else None
        self.snapshot_paths = [Path(p) for p in snapshot_paths] if snapshot_paths else []
        self.fluctuation_type = fluctuation_type.lower()
        if self.fluctuation_type not in {"reynolds", "favre"}:
            raise ValueError("fluctuation_type must be 'reynolds' or 'favre'")
        self.Lx, self.Ly = Lx, Ly
        self.output_dir = Path(output_dir) if output_dir is not None else self.run_dir
        self.results = None

    def load_snapshots(self, start=None, stop=None, stride: int = 1, max_snapshots=None):
        if self.run_dir is None:
            if not self.snapshot_paths:
                raise ValueError("Provide run_dir or snapshot_paths before loading")
            selected = self.snapshot_paths
        else:
            selected = sorted(self.run_dir.glob("hit2d_step*.npz"))
            if not selected:
                raise FileNotFoundError(f"No HIT2D .npz snapshots found in {self.run_dir}")
        selected = selected[slice(start, stop, max(stride, 1))]
        if max_snapshots is not None:
            selected = selected[:max_snapshots]
        if not selected:
            raise ValueError("Snapshot selection is empty")
        self.snapshot_paths = selected
        return selected

    @staticmethod
    def autocorrelation(field, normalize: bool = True):
        """Periodic autocorrelation of one plane on the device (numpy in,
        numpy out)."""
        f = _plane(field)[None]
        hat = dft2(f)
        return correlation_maps(hat, [(0, 0, (1 if normalize else 0, 0, 0))])[0].cpu().numpy()

    @staticmethod
    def cross_correlation(a, b, normalize: bool = False):
        """Periodic cross-correlation <a(x) b(x + r)> on the device."""
        a, b = _plane(a), _plane(b)
        if a.shape != b.shape:
            raise ValueError("cross-correlation fields must have the same shape")
        hat = dft2(torch.stack([a, b]))
        if not normalize:
            return correlation_maps(hat, [(0, 1, (0, 0, 0))])[0].cpu().numpy()
        maps = correlation_maps(hat, [(0, 0, (0, 0, 0)), (1, 1, (0, 0, 0)), (0, 1, (2, 0, 1))])
        return maps[2].cpu().numpy()

    @staticmethod
    def _load_snapshot_fields(path):
        with np.load(path, allow_pickle=False) as data:
            required = ("x", "y", "rho", "u", "v", "pressure")
            missing = [n for n in required if n not in data]
            if missing:
                raise KeyError(f"{path} is missing {missing}")
            out = {n: np.asarray(data[n], dtype=float) for n in required}
            for n in ("vorticity", "divergence"):
                if n in data:
                    out[n] = np.asarray(data[n], dtype=float)
        return out

    @staticmethod
    def _spacing(values) -> float:
        return float(values[1] - values[0]) if values.size > 1 else 1.0

    def snapshot_maps(self, rho, u, v, p, dx, dy, vorticity=None, divergence=None):
        """The nine correlation maps of one snapshot (numpy, by name)."""
        planes = fluctuation_planes(rho, u, v, dx, dy, p=p, vorticity=vorticity, divergence=divergence,
                                    fluctuation_type=self.fluctuation_type)
        hat = dft2(planes[2:8])  # u', v', rho', p', omega', theta' -> planes 0..5 of hat
        maps = correlation_maps(hat, [(a - 2, b - 2, nrm) for _, a, b, nrm in _MAPS]).cpu().numpy()
        return {name: maps[i] for i, (name, _, _, _) in enumerate(_MAPS)}

    def compute(self) -> CorrelationResults2D:
        if not self.snapshot_paths:
            self.load_snapshots()
        per = {name: [] for name, _, _, _ in _MAPS}
        per["R_LL"], per["R_NN"] = [], []
        meta = rx = ry = r = None
        for path in self.snapshot_paths:
            f = self._load_snapshot_fields(path)
            x, y = f["x"], f["y"]
            dx, dy = self._spacing(x), self._spacing(y)
            if meta is None:
                Lx = self.Lx if self.Lx is not None else dx * x.size
                Ly = self.Ly if self.Ly is not None else dy * y.size
                meta = {"nx": int(x.size), "ny": int(y.size), "Lx": float(Lx), "Ly": float(Ly), "dx": float(dx),
                        "dy": float(dy), "run_dir": str(self.run_dir) if self.run_dir is not None else ""}
                rx, ry = dx * np.arange(x.size // 2 + 1), dy * np.arange(y.size // 2 + 1)
            maps = self.snapshot_maps(f["rho"], f["u"], f["v"], f["pressure"], dx, dy, f.get("vorticity"),
                                      f.get("divergence"))
            for name in maps:
                per[name].append(maps[name])
            cur = self._directional_curves(maps["Ruu_2d"], maps["Rvv_2d"], rx.size, ry.size)
            common = min(cur["Ruu_x"].size, cur["Rvv_y"].size)
            r = 0.5 * (rx[:common] + ry[:common])
            per["R_LL"].append(self._normalize_curve(0.5 * (cur["Ruu_x"][:common] + cur["Rvv_y"][:common])))
            per["R_NN"].append(self._normalize_curve(0.5 * (cur["Rvv_x"][:common] + cur["Ruu_y"][:common])))
        if meta is None:
            raise RuntimeError("No snapshots were processed")
        LL, NN = np.stack(per["R_LL"]), np.stack(per["R_NN"])
        LLm, NNm = np.mean(LL, axis=0), np.mean(NN, axis=0)
        mean = {name: np.mean(np.stack(per[name]), axis=0) for name, _, _, _ in _MAPS}

        def curve(key, name):
            return np.mean(np.stack([self._directional_curves(m, m, rx.size, ry.size)[key] for m in per[name]]),
                           axis=0)

        self.results = CorrelationResults2D(
            r=r, rx=rx, ry=ry, Ruu_x=curve("Ruu_x", "Ruu_2d"), Ruu_y=curve("Ruu_y", "Ruu_2d"),
            Rvv_x=curve("Rvv_x", "Rvv_2d"), Rvv_y=curve("Rvv_y", "Rvv_2d"), R_LL=LLm, R_NN=NNm,
            R_LL_std=np.std(LL, axis=0), R_NN_std=np.std(NN, axis=0), L_integral=self.integral_length_scale(r, LLm),
            lambda_taylor=self.taylor_microscale(r, LLm), fluctuation_type=self.fluctuation_type,
            number_of_snapshots=len(self.snapshot_paths), metadata=meta, **mean)
        return self.results

    def save(self, output_path=None) -> Path:
        if self.results is None:
            raise RuntimeError("Call compute() before save()")
        if output_path is None:
            base = self.output_dir if self.output_dir is not None else Path(".")
            output_path = base / "two_point_correlation_results.npz"
        output_path = Path(output_path)
        output_path.parent.mkdir(parents=True, exist_ok=True)
        r = self.results
        data = {k: getattr(r, k) for k in ("r", "rx", "ry", "Ruu_2d", "Rvv_2d", "Ruv_2d", "Rrhop_2d", "Rp_theta_2d",
                                           "Rrho_2d", "Rp_2d", "Romega_2d", "Rtheta_2d", "Ruu_x", "Ruu_y", "Rvv_x",
                                           "Rvv_y", "R_LL", "R_NN", "R_LL_std", "R_NN_std")}
        data.update(L_integral=np.asarray(r.L_integral), lambda_taylor=np.asarray(r.lambda_taylor),
                    fluctuation_type=np.asarray(r.fluctuation_type),
                    number_of_snapshots=np.asarray(r.number_of_snapshots))
        for k, v in r.metadata.items():
            data[f"metadata_{k}"] = np.asarray(v)
        np.savez_compressed(output_path, **data)
        return output_path

    @staticmethod
    def integral_length_scale(r, R_LL) -> float:
        """Trapezoid integral of R_LL up to its first zero crossing (linearly
        interpolated), or over the whole range with a warning."""
        if r.size < 2 or R_LL.size < 2:
            return float("nan")
        neg = np.flatnonzero(R_LL < 0.0)
        if neg.size == 0:
            warnings.warn("R_LL has no zero crossing; integrating to L/2", RuntimeWarning)
            return float(np.trapezoid(R_LL, r))
        first = int(neg[0])
        if first == 0:
            return 0.0
        rs, Rs = list(r[:first]), list(R_LL[:first])
        r0, r1, R0, R1 = r[first - 1], r[first], R_LL[first - 1], R_LL[first]
        if R1 != R0:
            rs.append(float(r0 - R0 * (r1 - r0) / (R1 - R0)))
            Rs.append(0.0)
        return float(np.trapezoid(np.asarray(Rs), np.asarray(rs)))

    @staticmethod
    def taylor_microscale(r, R_LL, max_points: int = 6) -> float:
        """sqrt(1 / (2a)) from a least-squares fit R_LL = 1 - a r^2 over the
        first max_points separations after r = 0."""
        if r.size < 4 or R_LL.size < 4:
            warnings.warn("Not enough points for Taylor microscale fit", RuntimeWarning)
            return float("nan")
        n = min(max_points, r.size - 1)
        xs = r[1: n + 1] ** 2
        ys = 1.0 - R_LL[1: n + 1]
        ok = np.isfinite(xs) & np.isfinite(ys) & (xs > 0.0)
        if np.count_nonzero(ok) < 3:
            warnings.warn("Invalid points for Taylor microscale fit", RuntimeWarning)
            return float("nan")
        slope = float(np.dot(xs[ok], ys[ok]) / np.dot(xs[ok], xs[ok]))
        if slope <= 0.0 or not np.isfinite(slope):
            warnings.warn("Taylor microscale fit has non-positive curvature", RuntimeWarning)
            return float("nan")
        return float(np.sqrt(1.0 / (2.0 * slope)))

    @staticmethod
    def _directional_curves(Ruu, Rvv, nx_half, ny_half):
        # planes are [ny, nx]: x separations along a row
        return {"Ruu_x": Ruu[0, :nx_half], "Ruu_y": Ruu[:ny_half, 0], "Rvv_x": Rvv[0, :nx_half],
                "Rvv_y": Rvv[:ny_half, 0]}

    @staticmethod
    def _normalize_curve(values):
        zero = float(values[0])
        if abs(zero) <= np.finfo(float).eps:
            return np.full_like(values, np.nan)
        return values / zero
