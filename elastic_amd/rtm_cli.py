"""Command-line driver with the reference rtm.x interface (W/rtm.f90).

    python -m elastic_amd.rtm_cli stdin=lambdaS.H mu=muS.H rho=rhoS.H \\
        Vx_rec=seis-Vx.H Vz_rec=... sigmaxx_rec=... sigmazz_rec=... sigmaxz_rec=... \\
        drho=drho.H dlambda=dlambda.H dmu=dmu.H par=rtm-layer.p adj=1 lame=1

Parameters and their defaults are those of W/rtm.f90:38-56: the model axes
come from the input header (n1/o1/d1 = z, n2/o2/d2 = x); ``xsource``,
``zsource``, ``zrec`` are 1-based samples; ``Ttime`` is the propagation time,
``fp`` the Ricker peak frequency (25), ``abc`` the sponge width (0); ``adj``
selects RTM (1: reads the five (nt, nx) receiver files, writes drho, dlambda,
dmu) or Born modelling (0: reads dlambda, dmu, drho, writes the receiver
files). The time step is the reference's fixed rule dt = 0.4 dx / (sqrt(2)
2000), nt = int(Ttime / dt) in single precision (W/rtm.f90:66-67).
The reference reads the first model from stdin; here ``stdin=`` names it
(a header piped on stdin is accepted too, its payload resolved from the
current directory). ``lame=0`` (vp/vs input) is not supported.
Outputs are written in the reference's layout and axes
(W/seis-mod-rtm.f90:21-142). Propagation and imaging run on the GPU
(rtm_init_c_ / rtm_op_c_ of librtm_hip.so).
"""
from __future__ import annotations

import os
import sys
import tempfile
import time

import numpy as np

from . import lib, sep
from .synth import reference_dt_nt


def _read_stdin_header(pars):
    if "stdin" in pars:
        return sep.read(pars["stdin"])
    if sys.stdin is not None and not sys.stdin.isatty():
        text = sys.stdin.read()
        with tempfile.NamedTemporaryFile("w", suffix=".H", dir=os.getcwd(), delete=False) as t:
            t.write(text)
        try:
            return sep.read(t.name)
        finally:
            os.unlink(t.name)
    raise SystemExit("rtm_cli: no input model (stdin= or a header on stdin)")


def run(pars, log=sys.stderr):
    t0 = time.time()
    m1, h = _read_stdin_header(pars)
    nz, nx = int(float(h["n1"])), int(float(h["n2"]))
    dz, dx = float(h.get("d1", 1.0)), float(h.get("d2", 1.0))
    xsource = sep.get(pars, "xsource", int)
    zsource = sep.get(pars, "zsource", int)
    ttime = sep.get(pars, "Ttime", float)
    fp = sep.get(pars, "fp", float, 25.0)
    abc = sep.get(pars, "abc", int, 0)
    zrec = sep.get(pars, "zrec", int, 1)
    free_surface = sep.get(pars, "free_surface", bool, False)
    adj = sep.get(pars, "adj", bool)
    lame = sep.get(pars, "lame", bool)
    if not lame:
        raise SystemExit("rtm_cli: lame=0 (vp/vs models) is not supported; convert to Lame parameters")
    dt, nt = reference_dt_nt(dx, ttime)
    lam = np.asfortranarray(m1.reshape(nz, nx, order="F"), np.float32)
    mu, _ = sep.read(pars["mu"])
    rho, _ = sep.read(pars["rho"])
    mu = np.asfortranarray(mu.reshape(nz, nx, order="F"), np.float32)
    rho = np.asfortranarray(rho.reshape(nz, nx, order="F"), np.float32)
    print(f"rtm_cli: nz={nz} nx={nx} dx={dx:g} dt={dt:g} nt={nt} lame={lame} adj={adj}", file=log)
    tags = ["Vx_rec", "Vz_rec", "sigmaxx_rec", "sigmazz_rec", "sigmaxz_rec"]
    if adj:
        data = []
        for t in tags:
            a, _ = sep.read(pars[t])
            data.append(np.asfortranarray(a.reshape(nt, nx, order="F"), np.float32))
        dm = [np.zeros((nz, nx), np.float32, order="F") for _ in range(3)]
    else:
        dm = []
        for t in ("drho", "dlambda", "dmu"):
            a, _ = sep.read(pars[t])
            dm.append(np.asfortranarray(a.reshape(nz, nx, order="F"), np.float32))
        data = [np.zeros((nt, nx), np.float32, order="F") for _ in tags]
    op = lib.RtmOperator(nz, dz, nx, dx, dt, xsource, zsource, zrec, fp, nt, abc, free_surface, rho, lam, mu, True)
    op.op(adj, *dm, *data)
    op.close()
    if adj:
        ax = [(0.0, dz, ""), (0.0, dx, "")]
        for t, a in zip(("drho", "dlambda", "dmu"), dm):
            sep.write(pars[t], a, ax, program="rtm_cli")
    else:
        ax = [(0.0, dt, ""), (0.0, dx, "")]
        for t, a in zip(tags, data):
            sep.write(pars[t], a, ax, program="rtm_cli")
    print(f"rtm_cli: done in {time.time() - t0:.2f} s", file=log)
    return dm if adj else data


def main(argv=None):
    run(sep.parse_args(sys.argv[1:] if argv is None else argv))
    return 0


if __name__ == "__main__":
    sys.exit(main())
