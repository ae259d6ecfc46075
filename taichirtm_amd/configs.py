"""Seeded synthetic models and acquisition geometries of the benchmark workloads.

Units: velocities in m/s, density in g/cm^3 and moduli mu = rho*vs^2,
lam = ((vp/vs)^2 - 2)*mu in g/cm^3*(m/s)^2 = kPa.  This is the SI model
scaled by 1/1000 in density and moduli; the scheme is linear, so velocities
come out 1000x the SI values and nothing else changes.  In SI units the
reference's float32 four-point harmonic mean mu_a*mu_b*mu_c*mu_d
(forward_modeling.py:271-276) overflows for rock moduli above ~4.3e9 Pa and
turns the shear stress into NaN; the kPa scale keeps every product finite.

Material conversion follows rtm.py:901-902 (float32): mu = (rho*vs**2),
lam = ((vp/vs)**2 - 2)*mu.  Sources and receivers sit on the row j = 1
(rtm.py:674-676).  RNG: numpy default_rng(seed).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np


def ricker(f0: float, nt: int, dt: float, t0: Optional[float] = None) -> np.ndarray:
    """r(t) = (1 - 2 (pi f0 (t - t0))^2) exp(-(pi f0 (t - t0))^2), t = it*dt, t0 = 1/f0."""
    t0 = 1.0 / f0 if t0 is None else t0
    t = np.arange(nt) * dt
    a = (np.pi * f0 * (t - t0)) ** 2
    return ((1.0 - 2.0 * a) * np.exp(-a)).astype(np.float32)


def moduli(vp, vs, rho):
    """float32 mu and lam as rtm.py:901-902 forms them."""
    vp = np.asarray(vp, np.float32)
    vs = np.asarray(vs, np.float32)
    rho = np.asarray(rho, np.float32)
    mu = (rho * vs ** 2).astype(np.float32)
    lam = (((vp / vs) ** 2 - 2) * mu).astype(np.float32)
    return mu, lam


def gardner(vp):
    """rho = 310 * vp**0.25 kg/m^3, returned in g/cm^3."""
    return (0.31 * np.asarray(vp, np.float64) ** 0.25).astype(np.float32)


def _smooth_noise(rng, shape, sigma):
    from scipy.ndimage import gaussian_filter
    g = gaussian_filter(rng.standard_normal(shape), sigma, mode="reflect")
    return g / g.std()


@dataclass
class Shot:
    src: tuple                       # (i, j)
    rec: np.ndarray                  # (nr, 2) int32
    nt: int
    component: str = "w"             # force component: "u", "v" or "w"


@dataclass
class Config:
    name: str
    nx: int
    nz: int
    dx: float
    dz: float
    dt: float
    nt: int
    fw: int
    f0: float
    vp: np.ndarray
    vs: np.ndarray
    rho: np.ndarray
    shots: List[Shot] = field(default_factory=list)
    isnap: int = 1

    @property
    def fs(self):
        return 1.0 / self.dt

    def moduli(self):
        return moduli(self.vp, self.vs, self.rho)

    def wavelets(self, shot: Shot):
        w = ricker(self.f0, shot.nt, self.dt)[None, :]
        z = np.zeros_like(w)
        return tuple(w if c == shot.component else z for c in ("u", "v", "w"))

    def kwargs(self, shot: Shot):
        """Constructor keywords shared by ForwardModeling and BackwardModeling."""
        mu, lam = self.moduli()
        return dict(nx=self.nx, nz=self.nz, dx=self.dx, dz=self.dz, nt=shot.nt, fs=self.fs,
                    mu_field=mu, lam_field=lam, rho_field=self.rho, absorbing_frame=self.fw,
                    src_loc_field=np.array([shot.src], np.int32),
                    recv_loc_field=np.ascontiguousarray(shot.rec, np.int32),
                    num_sources=1, num_receivers=len(shot.rec))

    def cell_updates(self):
        """Useful cell-updates of one RTM (forward + backward) over all shots."""
        return 2 * self.nx * self.nz * sum(s.nt for s in self.shots)


def _surface_receivers(i0, i1, j=1):
    i = np.arange(i0, i1, dtype=np.int32)
    return np.stack([i, np.full_like(i, j)], axis=1)


def psv_anomaly(nx=200, nz=200, nt=1500, dx=5.0, dt=4e-4, fw=20, radius=20, f0=25.0,
                src=(100, 5), rec_range=(5, 195)):
    """Plain P-SV, one shot: background Vp 2000, Vs 1155, rho 2.0 with a disk
    of radius 20 cells at the centre (Vp 2500, Vs 1443, rho 2.2); vertical
    force (w) 25 Hz Ricker at (100, 5); receivers on j = 1, i = 5..194."""
    vp = np.full((nx, nz), 2000.0, np.float32)
    vs = np.full((nx, nz), 1155.0, np.float32)
    rho = np.full((nx, nz), 2.0, np.float32)
    ii, jj = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    disk = (ii - nx // 2) ** 2 + (jj - nz // 2) ** 2 <= radius ** 2
    vp[disk], vs[disk], rho[disk] = 2500.0, 1443.0, 2.2
    shot = Shot(src=src, rec=_surface_receivers(*rec_range), nt=nt, component="w")
    return Config("psv_anomaly_%dx%dx%d" % (nx, nz, nt), nx, nz, dx, dx, dt, nt, fw, f0,
                  vp, vs, rho, [shot])


def sh_layered(nx=1200, nz=600, nt=4000, dx=5.0, dt=5e-4, fw=40, f0=20.0, seed=0,
               n_layers=8, perturb=0.03, corr=10.0):
    """Plain SH, one shot: 8 horizontal layers at seeded depths, Vs rising
    linearly 800 -> 3000 m/s with depth, times (1 + 3% smooth Gaussian noise,
    correlation length 10 cells); Vp = 1.8 Vs; Gardner rho; 20 Hz Ricker on v
    at the surface centre; a receiver on every surface cell."""
    rng = np.random.default_rng(seed)
    tops = np.sort(rng.choice(np.arange(1, nz - 1), size=n_layers - 1, replace=False))
    layer = np.searchsorted(tops, np.arange(nz), side="right")
    vs_layer = np.linspace(800.0, 3000.0, n_layers)
    vs = np.broadcast_to(vs_layer[layer][None, :], (nx, nz)).astype(np.float64)
    vs = vs * (1.0 + perturb * _smooth_noise(rng, (nx, nz), corr))
    vs = vs.astype(np.float32)
    vp = (1.8 * vs).astype(np.float32)
    rho = gardner(vp)
    shot = Shot(src=(nx // 2, 1), rec=_surface_receivers(0, nx), nt=nt, component="v")
    return Config("sh_layered_%dx%dx%d" % (nx, nz, nt), nx, nz, dx, dx, dt, nt, fw, f0,
                  vp, vs, rho, [shot])


def psv_faulted(nx=2000, nz=1000, n_shots=64, nt_range=(2000, 5000), nr_range=(300, 1200),
                dx=5.0, dt=4e-4, fw=40, f0=20.0, seed=0, n_layers=10):
    """Batched P-SV: layered Vp 1500-4500 with two dipping faults (throw 20-60
    cells), Vp/Vs uniform in 1.7-2.0 per layer, Gardner rho; shots evenly
    spaced on the surface with record lengths drawn from nt_range and
    receiver spreads from nr_range, centred on the source."""
    rng = np.random.default_rng(seed)
    tops = np.sort(rng.choice(np.arange(20, nz - 20), size=n_layers - 1, replace=False))
    vp_layer = np.linspace(1500.0, 4500.0, n_layers)
    ratio = rng.uniform(1.7, 2.0, n_layers)
    ii, jj = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    shift = np.zeros((nx, nz))
    for k in range(2):
        x0 = rng.uniform(0.25, 0.75) * nx
        dip = rng.uniform(0.5, 2.0) * (1 if k == 0 else -1)
        throw = rng.integers(20, 61)
        shift += np.where(ii - x0 > dip * jj, throw, 0)
    layer = np.searchsorted(tops, jj - shift, side="right")
    vp = vp_layer[layer].astype(np.float32)
    vs = (vp / ratio[layer]).astype(np.float32)
    rho = gardner(vp)
    shots = []
    xs = np.linspace(fw + 10, nx - fw - 10, n_shots).astype(int)
    for s in range(n_shots):
        nt = int(rng.integers(nt_range[0], nt_range[1] + 1))
        nr = int(rng.integers(nr_range[0], nr_range[1] + 1))
        lo = max(0, xs[s] - nr // 2)
        hi = min(nx, lo + nr)
        shots.append(Shot(src=(int(xs[s]), 1), rec=_surface_receivers(lo, hi), nt=nt, component="w"))
    return Config("psv_faulted_%dx%d_%dshots" % (nx, nz, n_shots), nx, nz, dx, dx, dt,
                  max(s.nt for s in shots), fw, f0, vp, vs, rho, shots)


def psv_fwi(nx=600, nz=300, nt=2500, dx=10.0, dt=8e-4, n_shots=16, fw=30, f0=15.0, seed=0,
            blob=0.08):
    """FWI gradient problem: smooth seeded background Vp 2000-3500, Vs = Vp/1.8,
    Gardner rho; the true model adds two Gaussian blobs of +/-8%; the start
    model is the background.  16 surface shots, receivers on every surface
    cell.  Returns (true, start) configs."""
    rng = np.random.default_rng(seed)
    jj = np.arange(nz)[None, :] / (nz - 1)
    bg = 2000.0 + 1500.0 * jj + 150.0 * _smooth_noise(rng, (nx, nz), 25.0)
    bg = np.clip(bg, 2000.0, 3500.0)
    ii, kk = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    pert = np.zeros((nx, nz))
    for sign, cx, cz in ((+1, 0.35, 0.5), (-1, 0.65, 0.55)):
        pert += sign * blob * np.exp(-(((ii - cx * nx) / (0.06 * nx)) ** 2
                                       + ((kk - cz * nz) / (0.1 * nz)) ** 2))
    xs = np.linspace(fw + 5, nx - fw - 5, n_shots).astype(int)
    rec = _surface_receivers(0, nx)
    out = []
    for vp in ((bg * (1 + pert)), bg):
        vp = vp.astype(np.float32)
        vs = (vp / 1.8).astype(np.float32)
        shots = [Shot(src=(int(x), 1), rec=rec, nt=nt, component="w") for x in xs]
        out.append(Config("psv_fwi_%dx%dx%d" % (nx, nz, nt), nx, nz, dx, dx, dt, nt, fw, f0,
                          vp, vs, gardner(vp), shots))
    return tuple(out)


def psv_salt(nx=4000, nz=2000, nt=8000, n_shots=32, nr_range=(1000, 3000), dx=5.0, dt=4e-4,
             fw=40, f0=15.0, seed=0, n_layers=12):
    """Large budgeted adjoint: layered background 1500-4000 m/s with a salt
    body (Vp 4500, Vs 2600, rho 2.15) from thresholded smoothed noise; shots
    with uneven receiver spreads centred on their sources.  Returns (true,
    start) configs: with the salt, and the layered background alone."""
    rng = np.random.default_rng(seed)
    tops = np.sort(rng.choice(np.arange(20, nz - 20), size=n_layers - 1, replace=False))
    vp_layer = np.linspace(1500.0, 4000.0, n_layers)
    layer = np.searchsorted(tops, np.arange(nz), side="right")
    vp = np.broadcast_to(vp_layer[layer][None, :], (nx, nz)).astype(np.float32).copy()
    vs = (vp / 1.8).astype(np.float32)
    rho = gardner(vp)
    noise = _smooth_noise(rng, (nx // 4, nz // 4), 12.0)
    noise = np.kron(noise, np.ones((4, 4)))[:nx, :nz]
    ii, jj = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    window = np.exp(-(((ii - nx / 2) / (0.25 * nx)) ** 2 + ((jj - 0.45 * nz) / (0.18 * nz)) ** 2))
    salt = noise * window > 0.6
    background = (vp.copy(), vs.copy(), rho.copy())
    vp[salt], vs[salt], rho[salt] = 4500.0, 2600.0, 2.15
    xs = np.linspace(fw + 10, nx - fw - 10, n_shots).astype(int)
    shots = []
    for x in xs:
        nr = int(rng.integers(nr_range[0], nr_range[1] + 1))
        lo = max(0, x - nr // 2)
        hi = min(nx, lo + nr)
        shots.append(Shot(src=(int(x), 1), rec=_surface_receivers(lo, hi), nt=nt, component="w"))
    name = "psv_salt_%dx%dx%d" % (nx, nz, nt)
    return (Config(name, nx, nz, dx, dx, dt, nt, fw, f0, vp, vs, rho, shots),
            Config(name, nx, nz, dx, dx, dt, nt, fw, f0, *background, list(shots)))
