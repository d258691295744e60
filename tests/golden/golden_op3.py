"""Golden vectors of the reference adjoint rtm_op3_c_ (loader and checker).

``rtm_op3_ref.npz`` is written by ``make_golden_op3.py`` from the reference's
own C adjoint (R/Src_cuda_722-1/rtmlib-modC.c compiled into oracle/_ref). The
input is the rtm_op3_c_ entry state: levels nt-1 and nt-2 complete, lower
levels holding only the injected receiver row (W/rtmlib-mod.f90:99-146).
Stored: the compact input, the coefficients, every output level at three
checkpoints and the SHA-256 of the complete FP32 output.
"""
from __future__ import annotations

import hashlib
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(HERE, "rtm_op3_ref.npz")
CHECK_LEVELS = (0, 1, 29)


def expand(top, rows, nz, nx, nt, zrec):
    hist = []
    for f in range(5):
        h = np.zeros((nz, nx, nt), np.float32, order="F")
        h[:, :, nt - 2:] = top[f]
        h[zrec, :, : nt - 2] = rows[f]
        hist.append(h)
    return hist


def load():
    z = np.load(PATH, allow_pickle=False)
    nz, nx, nt, zrec = (int(v) for v in z["dims"])
    hist = expand(z["top"], z["rows"], nz, nx, nt, zrec)
    coefs = [np.asfortranarray(c) for c in z["coefs"]]
    return dict(hist=hist, coefs=coefs, levels=z["levels"], sha=str(z["sha256"]), nz=nz, nx=nx, nt=nt, zrec=zrec)


def digest(out):
    m = hashlib.sha256()
    for a in out:
        m.update(np.asfortranarray(a, np.float32).tobytes(order="F"))
    return m.hexdigest()


def check(g, out, exact=True, tol=1e-5):
    """exact: bit equality (FP32 reference order); else rel-L2 <= tol per field
    at the stored levels."""
    for f in range(5):
        for j, lev in enumerate(CHECK_LEVELS):
            got = np.asarray(out[f][:, :, lev])
            exp = g["levels"][f, j]
            if exact:
                assert np.array_equal(got.astype(np.float32), exp), (f, lev)
            else:
                err = np.linalg.norm((got - exp).ravel()) / np.linalg.norm(exp.ravel())
                assert err <= tol, (f, lev, err)
    if exact:
        assert digest(out) == g["sha"]
