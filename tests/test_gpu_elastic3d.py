"""GPU parity of the 9-field 3D elastic adjoint step (include/elastic3d.h)
against the FP64 oracle (oracle/elastic3d_oracle.cpp).

The reference's own test runs one step on 200^3 random data and compares its
CPU loop with its CUDA kernel (K3/rtmlib-modC.c:190-242, :697-705); here the
same shape and data distribution go through the reference's host interface
(rtm3d_gpu_func) and are compared with FP64 at the wavefield tolerance. The
kernel sums the products that share a gather before gathering, so it is not
bitwise equal to the reference's FP32 term order; the FP32 reference-order
oracle's own distance from FP64 is reported beside it.
"""
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-6  # per-field rel-L2 against FP64 for one step


def _data(ny, nz, nx, seed):
    fin, fin1, fout = (O.el3_random(ny, nz, nx, seed + s) for s in range(3))
    coef = O.el3_random(ny, nz, nx, seed + 3, k=6)
    return fin, fin1, fout, coef


def _field_errors(got, ref):
    return [O.rel_l2(got[k], ref[k]) for k in range(9)]


def test_reference_shape_200_cubed_host_interface(hip):
    ny = nz = nx = 200
    print("elastic3d host interface 200^3 seed 100", flush=True)
    fin, fin1, fout, coef = _data(ny, nz, nx, 100)
    got, t = hip.el3_gpu_func(fin, fin1, fout, coef)
    ref = O.el3_step(fin, fin1, fout, coef, nthreads=8)
    err = _field_errors(got, ref)
    print("rel-L2 per field", ["%.2e" % e for e in err], "times", t, flush=True)
    assert max(err) <= TOL
    # the padding of out is not written
    for sl in [np.s_[:, :5], np.s_[:, -5:], np.s_[:, :, :5], np.s_[:, :, -5:], np.s_[..., :5], np.s_[..., -5:]]:
        assert np.array_equal(got[sl], fout[sl])
    assert np.all(t > 0)


@pytest.mark.parametrize("width", ["4", "8", "16"])
@pytest.mark.parametrize("shape", [(7, 5, 3), (13, 70, 11), (20, 129, 17), (9, 64, 8)])
def test_odd_shapes_device_interface(hip, shape, width, monkeypatch):
    monkeypatch.setenv("RTM_EL3_W", width)
    ny, nz, nx = shape
    print("elastic3d device interface", shape, "width", width, flush=True)
    fin, fin1, fout, coef = _data(ny, nz, nx, 7)
    dev = [hip.DeviceArray(a) for a in (fin, fin1, fout)]
    dc = hip.DeviceArray(coef)
    roles = hip.el3_steps_dev(dev, dc, 1)
    assert roles[0] is dev[0] and roles[1] is dev[1] and roles[2] is dev[2]
    got = roles[2].get()
    ref = O.el3_step(fin, fin1, fout, coef, nthreads=8)
    assert max(_field_errors(got, ref)) <= TOL
    assert np.array_equal(roles[0].get(), fin) and np.array_equal(roles[1].get(), fin1)


def test_three_steps_with_the_reference_rotation(hip):
    ny, nz, nx = 12, 40, 20
    print("elastic3d three steps", flush=True)
    fin, fin1, fout, coef = _data(ny, nz, nx, 31)
    dev = [hip.DeviceArray(a) for a in (fin, fin1, fout)]
    dc = hip.DeviceArray(coef)
    roles = hip.el3_steps_dev(dev, dc, 3)
    # (out, in, in1) <- (in1, out, in) twice: (b0, b1, b2) -> (b2, b0, b1) -> (b1, b2, b0)
    assert roles[0] is dev[1] and roles[1] is dev[2] and roles[2] is dev[0]
    ref = O.el3_steps(fin, fin1, fout, coef, 3, nthreads=8)
    for g, r in zip(roles, ref):
        assert O.rel_l2(g.get(), r) <= 1e-5


def test_dot_product_against_the_fp64_system(hip):
    ny, nz, nx = 16, 70, 18
    rng = np.random.default_rng(5)
    shp = (ny + 10, nx + 10, nz + 10)
    coef = rng.uniform(-1, 1, size=(6,) + shp).astype(np.float32)
    u = np.zeros((9,) + shp, np.float32)
    w = np.zeros((9,) + shp, np.float32)
    u[:, 5:-5, 5:-5, 5:-5] = rng.uniform(-1, 1, size=(9, ny, nx, nz))
    w[:, 5:-5, 5:-5, 5:-5] = rng.uniform(-1, 1, size=(9, ny, nx, nz))
    z = np.zeros_like(u)
    dev = [hip.DeviceArray(a) for a in (u, z, z)]
    gu = hip.el3_steps_dev(dev, hip.DeviceArray(coef), 1)[2].get().astype(np.float64)
    fw = O.el3_fwd(w.astype(np.float64), coef.astype(np.float64), nthreads=8)
    lhs = float(np.sum(gu[:, 5:-5, 5:-5, 5:-5] * w[:, 5:-5, 5:-5, 5:-5]))
    rhs = float(np.sum(u[:, 5:-5, 5:-5, 5:-5].astype(np.float64) * fw[:, 5:-5, 5:-5, 5:-5]))
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


def test_chunking_and_tile_order_do_not_change_results(hip):
    ny, nz, nx = 40, 100, 30
    fin, fin1, fout, coef = _data(ny, nz, nx, 61)
    dc = hip.DeviceArray(coef)
    outs = []
    saved = {k: os.environ.get(k) for k in ("RTM_EL3_YCHUNK", "RTM_XCD_REMAP", "RTM_EL3_W")}
    try:
        for ych, xcd, ew in [(None, None, None), ("1", "1", "16"), ("7", "0", "8"), ("40", "1", "16"),
                             ("9", "1", "4"), ("1000", "0", "16")]:
            for k, v in (("RTM_EL3_YCHUNK", ych), ("RTM_XCD_REMAP", xcd), ("RTM_EL3_W", ew)):
                if v is None:
                    os.environ.pop(k, None)
                else:
                    os.environ[k] = v
            dev = [hip.DeviceArray(a) for a in (fin, fin1, fout)]
            outs.append(hip.el3_steps_dev(dev, dc, 1)[2].get())
            print("ychunk", ych, "xcd", xcd, "width", ew, "planes per chunk", hip.el3_timings()[3], flush=True)
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


def test_rejects_bad_arguments(hip):
    a = [hip.DeviceArray(np.zeros((9, 11, 11, 11), np.float32)) for _ in range(3)]
    c = hip.DeviceArray(np.zeros((6, 11, 11, 11), np.float32))
    with pytest.raises(hip.RtmError):
        hip.el3_steps_dev(a, c, -1)
    lib = hip.load()
    assert lib.rtm3d_gpu_func(1, 1, 1, *([None] * 35)) != 0
