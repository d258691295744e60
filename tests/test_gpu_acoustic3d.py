"""GPU parity of the 3D acoustic subdomain propagator and imaging against the
FP64 oracle (grid spans several z/x tiles and y-chunks; 10-cell sponge)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

N1, N2, N3, NT, ABC = 80, 44, 140, 90, 10
SRC, ZR = (20, 22, 70), 12


def _model(seed=3):
    rng = np.random.default_rng(seed)
    z = np.arange(N1)[:, None, None]
    v = 0.30 + 0.08 * z / N1 + 0.0 * np.zeros((1, N2, N3))
    v = v + 0.03 * (z > 45 + 0.1 * (np.arange(N3)[None, None, :] - 70))
    c = np.asfortranarray(v ** 2)
    cs = np.asfortranarray(np.full_like(c, c.mean()) + 0.01 * rng.uniform(-1, 1, c.shape))
    return c, cs


def _amp():
    t = np.arange(1, NT + 1, dtype=np.float64)
    a = (np.pi * (t - 25.0) / 18.0) ** 2
    return (1 - 2 * a) * np.exp(-a)


def test_source_pass_data_vs_fp64(hip):
    c, _ = _model()
    amp = _amp()
    g = hip.Acoustic3D(c, amp, SRC, ZR, abc=ABC, snap_every=5, snap_prec="fp32")
    g.model_data()
    rec = g.data()
    ref, _ = O.ac3_run(c, amp, SRC, ZR, abc=ABC, nthreads=8)
    assert np.abs(ref).max() > 0
    assert O.rel_l2(rec, ref) <= 1e-5
    g.close()


def test_observed_difference_vs_fp64(hip):
    c, cs = _model()
    amp = _amp()
    g = hip.Acoustic3D(cs, amp, SRC, ZR, abc=ABC, snap_every=5, snap_prec="fp32")
    g.model_data(c_true=c)
    rec = g.data()
    r1, _ = O.ac3_run(c, amp, SRC, ZR, abc=ABC, nthreads=8)
    r2, _ = O.ac3_run(cs, amp, SRC, ZR, abc=ABC, nthreads=8)
    assert O.rel_l2(rec, r1 - r2) <= 1e-4
    g.close()


@pytest.mark.parametrize("prec,tol", [("fp32", 1e-4), ("bf16", 1e-3)])
def test_migrate_image_vs_fp64(hip, prec, tol):
    c, cs = _model()
    amp = _amp()
    g = hip.Acoustic3D(cs, amp, SRC, ZR, abc=ABC, snap_every=5, snap_prec=prec)
    g.model_data(c_true=c)
    rec = g.data()
    g.migrate()
    img = g.image()
    _, snap = O.ac3_run(cs, amp, SRC, ZR, abc=ABC, j=5, want_snap=True, nthreads=8)
    ref = O.ac3_run(cs, amp, SRC, ZR, abc=ABC, mode=1, rec=rec.astype(np.float64), j=5, snap=snap, nthreads=8)
    ref = np.asfortranarray(ref.transpose(2, 1, 0))
    err = O.rel_l2(img, ref)
    print(prec, err)
    assert np.abs(ref).max() > 0
    assert err <= tol
    g.close()
