"""The allocation-free drop-in adjoint (rtm_op3_hip_noalloc) against the
allocating one (rtm_op3_hip): bit-for-bit the same histories on two calls in
a row, the second on a smaller grid (the shared device array keeps the first
call's contents), for the tile and the marching kernel forms; sizes beyond
its scratch are refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inputs(seed, nz, nx, nt, zrec=3):
    rng = np.random.default_rng(seed)
    hist = []
    for _ in range(5):
        h = np.zeros((nz, nx, nt), np.float32, order="F")
        h[:, :, nt - 2:] = rng.uniform(-1, 1, (nz, nx, 2))
        h[zrec, :, : nt - 2] = rng.uniform(-1, 1, (nx, nt - 2))
        hist.append(h)
    m1 = [rng.uniform(0.4e-7, 0.6e-7, (nz, nx)).astype(np.float32) for _ in range(2)]
    aux = [rng.uniform(0.5e6, 2.0e6, (nz, nx)).astype(np.float32) for _ in range(3)]
    return hist, [m1[0], m1[1], aux[0], aux[1], aux[2]]


@pytest.mark.parametrize("form", ["tiles", "marching"])
def test_noalloc_equals_allocating_drop_in_twice(hip, monkeypatch, form):
    monkeypatch.setenv("RTM_TBLOCK", "2" if form == "marching" else "0")
    if form == "marching":
        monkeypatch.setenv("RTM_MARCH_W", "16")
    for seed, (nz, nx, nt) in ((1, (390, 690, 40)), (2, (70, 93, 17))):
        print(f"noalloc {form}: {nz}x{nx}x{nt}", flush=True)
        hist, coefs = _inputs(seed, nz, nx, nt)
        a = hip.rtm_op3(hist, coefs)
        b = hip.rtm_op3_noalloc(hist, coefs)
        for f in range(5):
            assert np.array_equal(a[f], b[f]), (form, nz, nx, f)
            assert np.any(a[f][:, :, 0] != hist[f][:, :, 0])


def test_noalloc_refuses_grids_beyond_its_scratch(hip):
    nz, nx = 1300, 1300  # 20 * nz * nx floats > 128 MiB
    hist = [np.zeros((nz, nx, 3), np.float32, order="F") for _ in range(5)]
    coefs = [np.zeros((nz, nx), np.float32, order="F") for _ in range(5)]
    with pytest.raises(hip.RtmError) as e:
        hip.rtm_op3_noalloc(hist, coefs)
    assert e.value.code == -1 and "scratch" in str(e.value)
