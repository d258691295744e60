"""Write tests/golden/rtm_op3_ref.npz from the reference's own C adjoint.

Needs oracle/_ref/librtm_ref.so (``make -C oracle``, built from the reference
source R/Src_cuda_722-1/rtmlib-modC.c where it lies). Run from the repo root:
``python -m tests.golden.make_golden_op3``.
"""
import numpy as np

import oracle as O
from tests.golden import golden_op3 as G


def main():
    nz, nx, nt, zrec = 14, 16, 60, 3  # nt >= 59: the reference prints it%(nt/59)
    rng = np.random.default_rng(20150713)
    top = rng.uniform(-1, 1, (5, nz, nx, 2)).astype(np.float32)
    rows = rng.uniform(-1, 1, (5, nx, nt - 2)).astype(np.float32)
    m1 = rng.uniform(0.4e-7, 0.6e-7, (2, nz, nx)).astype(np.float32)
    aux = rng.uniform(0.5e6, 2.0e6, (3, nz, nx)).astype(np.float32)
    coefs = np.concatenate([m1, aux])
    hist = G.expand(top, rows, nz, nx, nt, zrec)
    out = O.ref_rtm_op3_c(hist, [np.asfortranarray(c) for c in coefs], zrec)
    levels = np.stack([np.stack([out[f][:, :, lev] for lev in G.CHECK_LEVELS]) for f in range(5)])
    np.savez_compressed(G.PATH, dims=np.array([nz, nx, nt, zrec]), top=top, rows=rows, coefs=coefs,
                        levels=levels.astype(np.float32), sha256=np.array(G.digest(out)))
    print("wrote", G.PATH)


if __name__ == "__main__":
    main()
