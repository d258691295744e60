"""Seeded synthetic earth models and shot geometry of the benchmark workloads.

This is input preparation on the host (numpy); it is not on the propagation
path. The reference builds its models with SEP tools (W/create-model-lame.f90,
``Smooth rect1=20 rect2=20`` in R/Makefile:85-89); the workloads here define
their own seeded construction instead:

* Vp rises linearly with depth, plus ``nlayers`` flat interfaces at seeded
  depths, each adding a seeded +-``jump`` m/s step to everything below it,
  plus one Gaussian anomaly (sigma ``anom_sigma`` cells, +``anom_dv`` m/s) at a
  seeded position in the middle half of the model.
* Vs = Vp / 1.732; rho = 310 Vp^0.25 kg/m^3 (Gardner, 0.31 Vp^0.25 g/cc);
  lambda = rho (Vp^2 - 2 Vs^2), mu = rho Vs^2.
* The migration model is the Gaussian-smoothed (sigma 10 cells) Vp of the
  true model, with Vs and rho derived from the smoothed Vp again.
* The computational grid pads the physical grid by ``abc + 5`` cells on every
  side (edge-replicated material), so that the physical domain is exactly the
  region where the sponge factor is 1 and a source or receiver at z = 20 m is
  updated by every operator (the reference's outer 4-5 rows are never
  updated and 45 rows are inside its sponge, W/abc-mod.f90:35).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Model:
    vp: np.ndarray  # (nz, nx), m/s, Fortran order
    vs: np.ndarray
    rho: np.ndarray

    @property
    def lam(self):
        return np.asfortranarray(self.rho * (self.vp ** 2 - 2.0 * self.vs ** 2))

    @property
    def mu(self):
        return np.asfortranarray(self.rho * self.vs ** 2)


def elastic_from_vp(vp: np.ndarray) -> Model:
    vp = np.asfortranarray(vp, np.float64)
    vs = vp / 1.732
    rho = 310.0 * vp ** 0.25
    return Model(vp, np.asfortranarray(vs), np.asfortranarray(rho))


def layered_vp(nz, nx, seed, vp_top=1800.0, vp_bot=4500.0, nlayers=5, jump=200.0,
               anom_sigma=15.0, anom_dv=300.0, dip=None):
    """Seeded layered Vp model (nz, nx). ``dip`` (radians, optional) tilts the
    layer interfaces."""
    rng = np.random.default_rng(seed)
    z = np.arange(nz, dtype=np.float64)[:, None]
    x = np.arange(nx, dtype=np.float64)[None, :]
    vp = vp_top + (vp_bot - vp_top) * z / max(nz - 1, 1) + 0.0 * x
    depths = np.sort(rng.uniform(0.15 * nz, 0.95 * nz, nlayers))
    signs = rng.choice([-1.0, 1.0], nlayers)
    slope = np.tan(dip) if dip is not None else 0.0
    for d, s in zip(depths, signs):
        vp = vp + s * jump * (z >= d + slope * (x - nx / 2))
    cz = rng.uniform(0.35 * nz, 0.65 * nz)
    cx = rng.uniform(0.35 * nx, 0.65 * nx)
    vp = vp + anom_dv * np.exp(-((z - cz) ** 2 + (x - cx) ** 2) / (2.0 * anom_sigma ** 2))
    return np.asfortranarray(vp)


def smooth(vp, sigma=10.0):
    from scipy.ndimage import gaussian_filter

    return np.asfortranarray(gaussian_filter(np.asarray(vp, np.float64), sigma=sigma, mode="nearest"))


def pad_edge(a, pad):
    return np.asfortranarray(np.pad(a, pad, mode="edge"))


@dataclass
class Geometry:
    """One shot on a padded computational grid (0-based cells)."""

    nz: int
    nx: int
    pad: int
    xs: int
    zs: int
    zr: int
    rx0: int
    nrx: int
    x_origin: int = 0  # first physical column of this window in the full model
    extra: dict = field(default_factory=dict)


def single_shot_geometry(nz_phys, nx_phys, abc, dx, zsrc_m=20.0, xsrc=None, zrec_m=20.0):
    pad = abc + 5
    xs_phys = nx_phys // 2 if xsrc is None else xsrc
    zs_phys = int(round(zsrc_m / dx))
    zr_phys = int(round(zrec_m / dx))
    return Geometry(nz=nz_phys + 2 * pad, nx=nx_phys + 2 * pad, pad=pad, xs=pad + xs_phys, zs=pad + zs_phys,
                    zr=pad + zr_phys, rx0=pad, nrx=nx_phys)


def config0(seed=20260101, nz=300, nx=600, dx=10.0, dt=1e-3, nt=2000, fp=15.0, abc=40):
    """Plain 2D isotropic elastic RTM, one shot: true and migration models on
    the padded grid plus geometry."""
    vp = layered_vp(nz, nx, seed)
    vps = smooth(vp, 10.0)
    geo = single_shot_geometry(nz, nx, abc, dx)
    true = elastic_from_vp(pad_edge(vp, geo.pad))
    mig = elastic_from_vp(pad_edge(vps, geo.pad))
    return dict(true=true, mig=mig, geo=geo, dx=dx, dt=dt, nt=nt, fp=fp, abc=abc)


def batch_apertures(nshots, nx_model, seed, wmin=600, wmax=1200):
    """Config-1 shot layout: source x uniform along the model, aperture width
    drawn uniformly from [wmin, wmax] columns, window centred on the source and
    clamped inside the model. Returns a list of (xs, x0, w) in model columns."""
    rng = np.random.default_rng(seed)
    xs = np.linspace(0, nx_model - 1, nshots + 2)[1:-1].round().astype(int)
    ws = rng.integers(wmin, wmax + 1, nshots)
    out = []
    for x, w in zip(xs, ws):
        w = int(min(w, nx_model))
        x0 = int(np.clip(x - w // 2, 0, nx_model - w))
        out.append((int(x), x0, w))
    return out


def reference_dt_nt(dx, ttime):
    """The reference driver's time step and level count, W/rtm.f90:66-67, in
    single precision: dt = 0.4*dx/(sqrt(2)*2000), nt = int(Ttime/dt)."""
    f = np.float32
    dt = f(0.4) * f(dx) / (np.sqrt(f(2.0)) * f(2000.0))
    return float(dt), int(f(ttime) / dt)


def window_shot(vp_true, vp_mig, x0, w, xs_model, abc, dx, dt, zsrc_m=20.0, zrec_m=20.0):
    """One shot of a batched survey: the model columns [x0, x0+w) padded by
    abc+5 cells. Returns (spec dict for the batch API, true coef5, migration
    coef5) with coefficients from the library's model_op."""
    from . import lib

    pad = abc + 5
    nzp = vp_true.shape[0]
    coefs = []
    for vp in (vp_true, vp_mig):
        m = elastic_from_vp(pad_edge(vp[:, x0:x0 + w], pad))
        coefs.append(lib.model_op(m.rho, m.lam, m.mu, dx, dt))
    spec = dict(nz=nzp + 2 * pad, nx=w + 2 * pad, xs=pad + (xs_model - x0), zs=pad + int(round(zsrc_m / dx)),
                zr=pad + int(round(zrec_m / dx)), rx0=pad, nrx=w, img_z0=pad, img_x0=pad, img_nz=nzp, img_nx=w,
                stack_z0=0, stack_x0=x0)
    return spec, coefs[0], coefs[1]


def survey(nz, nx, nshots, seed, abc, dx, dt, wmin=600, wmax=1200, sigma=10.0):
    """Seeded batched survey (config-1 construction): layered model, smoothed
    migration model, shot windows. Returns (vp_true, vp_mig, list of
    (xs, x0, w))."""
    vp = layered_vp(nz, nx, seed)
    vps = smooth(vp, sigma)
    return vp, vps, batch_apertures(nshots, nx, seed + 1, min(wmin, nx), min(wmax, nx))
