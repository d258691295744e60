"""Source-wavefield movie writer with the reference's snap_wfld outputs
(W/snap-mod-rtm.f90:154-283).

The reference keeps nsnap = 100 frames at levels k*dsnap (dsnap = nt/nsnap,
k = 1..nsnap) of each of the five fields, windowed to rows abc+6.. and
columns abc+1..nx-abc (1-based), and writes them as SEP cubes Vx0.H, Vz0.H,
sigmaxx0.H, sigmazz0.H, sigmaxz0.H with axes (z: dz, x: dx, time: dt*dsnap).
Its row window is one row longer than the array it allocates
(nz-2*abc-6 rows); here the window has exactly nz-2*abc-6 rows, starting at
the same row. The frames come from one forward run on the GPU
(rtm_forward_movie_hip); only the frames are copied to the host.
"""
from __future__ import annotations

import os

import numpy as np

from . import lib, sep

FIELDS = ("Vx0", "Vz0", "sigmaxx0", "sigmazz0", "sigmaxz0")


def window(nz, nx, abc):
    """(z0, x0, wnz, wnx), 0-based, of the reference's movie window."""
    return abc + 5, abc, nz - 2 * abc - 6, nx - 2 * abc


def snap_wfld(s: lib.RtmShot, coef5, wavelet, outdir, dz=None, nsnap=100):
    """Write the five movie cubes into outdir; returns {name: path}."""
    z0, x0, wnz, wnx = window(s.nz, s.nx, max(s.abc, 0))
    movie, dsnap = lib.forward_movie(s, coef5, wavelet, nsnap, z0, x0, wnz, wnx)
    os.makedirs(outdir, exist_ok=True)
    dz = s.dx if dz is None else dz
    out = {}
    for f, name in enumerate(FIELDS):
        cube = np.asfortranarray(np.moveaxis(movie[:, f], 0, -1))  # (wnz, wnx, nsnap)
        path = os.path.join(outdir, name + ".H")
        sep.write(path, cube, [(0.0, dz, ""), (0.0, s.dx, ""), (0.0, s.dt * dsnap, "")], program="snap_wfld")
        out[name] = path
    return out
