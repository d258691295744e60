"""Lame-parameter model builder (host setup), following the reference's
W/create-model-lame.f90 step for step, and its command-line form.

Layers are painted in the reference's order, each overwriting the previous
(1-based depth indices as in the reference):

1. water layer rows 1..waterz (W/create-model-lame.f90:123-132);
2. low-velocity layer rows waterz+1..waterz+nlvl, mu growing by lvlvz1 per
   row, lambda = lvllambda + lvlvz2 per row or = mu when lvllambda is -1
   (Poisson solid) (:136-150);
3. background rows waterz+nlvl+1..nz: lambda = lambda1 + vz*(iz-waterz-nlvl),
   mu = mu1 + vz/2*(...) or lambda/2 when mu1 is -1, rho = rho1 (:153-164);
4. reflector rows z1..z1+thick (inclusive, so thick+1 rows): lambda2, mu2 or
   half the background mu when mu2 is -1, rho2 (:167-179);
5. Gaussian anomaly exp(-((ix-anomx)^2+(iz-anomz)^2)/(2 sigma^2)): a blend
   towards (anomlambda, anommu, anomrho) when anomstr is 0, else a relative
   scaling by 1 + anomstr*exp(...) (:182-197);
6. disc (circlelambda, circlemu, circlerho) where the squared distance to
   (circlez, circlex) is at most circleradius -- the reference compares the
   squared distance with the radius itself, which is kept (:200-211).

Arithmetic is float32 like the reference's ``real`` arrays. Unset values
keep the reference defaults (W/create-model-lame.f90:22-85).
"""
from __future__ import annotations

import sys

import numpy as np

from . import sep

DEFAULTS = dict(oz=0.0, ox=0.0, water=False, waterz=0, waterlambda=2.25e9, watermu=0.0, waterrho=1000.0,
                lvl=False, nlvl=0, lvlvz1=0.0, lvlvz2=0.0, lvllambda=-1.0, lvlmu=0.0, lvlrho=1000.0,
                lambda1=0.0, mu1=-1.0, rho1=1000.0, vz=0.0, ref=False, z1=0, thick=0, lambda2=-1.0, mu2=-1.0,
                rho2=-1.0, anom=False, anomz=0, anomx=0, anomsigma=0.0, anomstr=0.0, anomrho=0.0, anomlambda=0.0,
                anommu=0.0, circle=False, circlez=0, circlex=0, circleradius=0, circlelambda=-1.0, circlemu=-1.0,
                circlerho=-1.0)
_INT = {"nz", "nx", "waterz", "nlvl", "z1", "thick", "anomz", "anomx", "circlez", "circlex", "circleradius"}
_BOOL = {"water", "lvl", "ref", "anom", "circle"}


def params_from(pars):
    """Typed parameters from a parsed SEP parameter dict (defaults filled)."""
    p = {}
    for k in ("nz", "dz", "nx", "dx"):
        p[k] = sep.get(pars, k, int if k in _INT else float)
    for k, d in DEFAULTS.items():
        typ = bool if k in _BOOL else int if k in _INT else float
        p[k] = sep.get(pars, k, typ, d) if k in pars else d
    return p


def build(p):
    """(lambda, mu, rho) float32 arrays (nz, nx), Fortran order."""
    f = np.float32
    nz, nx = int(p["nz"]), int(p["nx"])
    lam = np.zeros((nz, nx), f, order="F")
    mu = np.zeros((nz, nx), f, order="F")
    rho = np.zeros((nz, nx), f, order="F")
    iz1 = np.arange(1, nz + 1)  # 1-based rows
    waterz = int(p["waterz"])  # the later loops use waterz / nlvl whether or not the layers are on
    if p["water"]:
        r = slice(0, min(waterz, nz))
        lam[r, :] = f(p["waterlambda"])
        mu[r, :] = f(p["watermu"])
        rho[r, :] = f(p["waterrho"])
    nlvl = int(p["nlvl"])
    if p["lvl"]:
        for iz in range(1 + waterz, waterz + nlvl + 1):
            if 1 <= iz <= nz:
                m = f(f(p["lvlmu"]) + f(iz - waterz - 1) * f(p["lvlvz1"]))
                mu[iz - 1, :] = m
                lam[iz - 1, :] = (f(f(p["lvllambda"]) + f(iz - waterz - 1) * f(p["lvlvz2"]))
                                  if p["lvllambda"] != -1.0 else m)
                rho[iz - 1, :] = f(p["lvlrho"])
    for iz in range(1 + waterz + nlvl, nz + 1):
        lv = f(f(p["lambda1"]) + f(p["vz"]) * f(iz - waterz - nlvl))
        lam[iz - 1, :] = lv
        mu[iz - 1, :] = (f(f(p["mu1"]) + f(0.5) * f(p["vz"]) * f(iz - waterz - nlvl)) if p["mu1"] != -1.0
                         else f(f(0.5) * lv))
        rho[iz - 1, :] = f(p["rho1"])
    if p["ref"]:
        for iz in range(int(p["z1"]), int(p["z1"]) + int(p["thick"]) + 1):
            if 1 <= iz <= nz:
                lam[iz - 1, :] = f(p["lambda2"])
                mu[iz - 1, :] = f(p["mu2"]) if p["mu2"] != -1.0 else f(0.5) * mu[iz - 1, :]
                rho[iz - 1, :] = f(p["rho2"])
    if p["anom"]:
        ix1 = np.arange(1, nx + 1)
        dz2 = (iz1[:, None] - int(p["anomz"])) ** 2
        dx2 = (ix1[None, :] - int(p["anomx"])) ** 2
        aux = ((dx2 + dz2).astype(f) / (f(2.0) * f(p["anomsigma"]) * f(p["anomsigma"]))).astype(f)
        e = np.exp(-aux).astype(f)
        if p["anomstr"] == 0.0:
            lam[:] = lam + (f(p["anomlambda"]) - lam) * e
            mu[:] = mu + (f(p["anommu"]) - mu) * e
            rho[:] = rho + (f(p["anomrho"]) - rho) * e
        else:
            s = (f(1.0) + f(p["anomstr"]) * e).astype(f)
            lam[:] = lam * s
            mu[:] = mu * s
            rho[:] = rho * s
    if p["circle"]:
        ix1 = np.arange(1, nx + 1)
        aux = (ix1[None, :] - int(p["circlex"])) ** 2 + (iz1[:, None] - int(p["circlez"])) ** 2
        inside = aux <= int(p["circleradius"])
        lam[inside] = f(p["circlelambda"])
        mu[inside] = f(p["circlemu"])
        rho[inside] = f(p["circlerho"])
    return lam, mu, rho


def write_model(p, lam, mu, rho, out_lambda, out_mu, out_rho):
    """The three SEP outputs of the reference program (stdout = lambda, tags
    mu= and rho=), with axes (z, x)."""
    ax = [(p["oz"], p["dz"], "z"), (p["ox"], p["dx"], "x")]
    for path, a in ((out_lambda, lam), (out_mu, mu), (out_rho, rho)):
        sep.write(path, a, ax, program="create-model-lame")


def main(argv=None):
    """``python -m elastic_amd.model_builder par=layer-lambda.p mu=mu.H
    rho=rho.H out=lambda.H`` (the reference writes lambda to stdout; a path
    is required here because the payload is written next to it)."""
    pars = sep.parse_args(sys.argv[1:] if argv is None else argv)
    p = params_from(pars)
    lam, mu, rho = build(p)
    write_model(p, lam, mu, rho, pars.get("out", "lambda.H"), pars["mu"], pars["rho"])
    return 0


if __name__ == "__main__":
    sys.exit(main())
