"""Pins of the 3D elastic adjoint-step oracle (oracle/elastic3d_oracle.cpp).

The reference holds no stored outputs for this step (its own test compares
its CPU loop with its CUDA kernel, K3/rtmlib-modC.c:697-705), and its CPU
loop only links together with its CUDA half, so the oracle is pinned by:
  * the dot-product test of the step's gather against the forward operator
    written as the velocity-stress system (el3_fwd);
  * polynomial exactness of that system's staggered derivatives against
    closed-form derivatives at i + 1/2 and i - 1/2;
  * the symmetry of the step under every permutation of the axes;
  * single taps read off the reference's source lines;
  * FP32 (reference operation order) against FP64, and the multi-step
    pointer rotation of K3/rtmlib-modC.c:597-690.
"""
import numpy as np
import pytest

import oracle as O

NY, NZ, NX = 9, 11, 10
IDX = {n: i for i, n in enumerate(O.EL3_FIELDS)}
CDX = {n: i for i, n in enumerate(O.EL3_COEFS)}


def _rand(seed, k=9, rim=True):
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1, 1, size=(k, NY + 10, NX + 10, NZ + 10))
    if not rim:
        b = np.zeros_like(a)
        b[:, 5:-5, 5:-5, 5:-5] = a[:, 5:-5, 5:-5, 5:-5]
        a = b
    return a


def _gather(u, coef, dtype=np.float64):
    z = np.zeros_like(u)
    return O.el3_step(u, z, z, coef, dtype=dtype)


def _interior(a):
    return a[..., 5:-5, 5:-5, 5:-5]


def test_dot_product_gather_vs_velocity_stress_system():
    """<G u, w> = <u, F w>: u, w supported on the interior, coefficients
    random everywhere."""
    coef = _rand(1, k=6)
    u = _rand(2, rim=False)
    w = _rand(3, rim=False)
    gu = _gather(u, coef)
    fw = O.el3_fwd(w, coef)
    lhs = float(np.sum(_interior(gu) * _interior(w)))
    rhs = float(np.sum(_interior(u) * _interior(fw)))
    assert abs(lhs) > 1.0
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_dot_product_fails_for_a_planted_transpose():
    """The dot test separates the gather from a wrongly paired system: swap
    m2 and m2m3 in the forward operator only."""
    coef = _rand(1, k=6)
    u = _rand(2, rim=False)
    w = _rand(3, rim=False)
    gu = _gather(u, coef)
    bad = coef.copy()
    bad[[CDX["m2"], CDX["m2m3"]]] = bad[[CDX["m2m3"], CDX["m2"]]]
    fw = O.el3_fwd(w, bad)
    lhs = float(np.sum(_interior(gu) * _interior(w)))
    rhs = float(np.sum(_interior(u) * _interior(fw)))
    assert abs(lhs - rhs) > 1e-3 * abs(lhs)


def _poly(deg, t):
    return (t / 8.0) ** deg - 0.5 * (t / 8.0) ** max(deg - 1, 0)


def _dpoly(deg, t):
    d = deg * (t / 8.0) ** (deg - 1) / 8.0 if deg >= 1 else 0.0 * t
    e = max(deg - 1, 0)
    d2 = e * (t / 8.0) ** (e - 1) / 8.0 if e >= 1 else 0.0 * t
    return d - 0.5 * d2


# axis -> array axis of a (k, y, x, z) block
AXES = {"y": 1, "x": 2, "z": 3}
# the velocity-stress system: output <- [(coef, field, axis, +1 for D+ / -1 for D-)]
SYSTEM = {
    "Vx": [("m1_x", "sxx", "x", -1), ("m1_x", "sxy", "y", -1), ("m1_x", "sxz", "z", -1)],
    "Vy": [("m1_y", "sxy", "x", -1), ("m1_y", "syy", "y", -1), ("m1_y", "syz", "z", -1)],
    "Vz": [("m1_z", "sxz", "x", -1), ("m1_z", "syz", "y", -1), ("m1_z", "szz", "z", -1)],
    "sxx": [("m2m3", "Vx", "x", +1), ("m2", "Vy", "y", +1), ("m2", "Vz", "z", +1)],
    "syy": [("m2", "Vx", "x", +1), ("m2m3", "Vy", "y", +1), ("m2", "Vz", "z", +1)],
    "szz": [("m2", "Vx", "x", +1), ("m2", "Vy", "y", +1), ("m2m3", "Vz", "z", +1)],
    "sxy": [("m3", "Vx", "y", -1), ("m3", "Vy", "x", -1)],
    "sxz": [("m3", "Vx", "z", -1), ("m3", "Vz", "x", -1)],
    "syz": [("m3", "Vy", "z", -1), ("m3", "Vz", "y", -1)],
}
CVAL = {"m1_y": 2.0, "m1_x": 3.0, "m1_z": 5.0, "m2": 7.0, "m3": 11.0, "m2m3": 13.0}


@pytest.mark.parametrize("field", O.EL3_FIELDS)
@pytest.mark.parametrize("axis", ["z", "x", "y"])
@pytest.mark.parametrize("deg", [1, 4, 9])
def test_system_exact_on_polynomials(field, axis, deg):
    """A staggered 10th-order derivative is exact for degree <= 9; D+ is the
    derivative at i + 1/2, D- at i - 1/2 (index units)."""
    coef = np.zeros((6, NY + 10, NX + 10, NZ + 10))
    for n, v in CVAL.items():
        coef[CDX[n]] = v
    x = np.zeros((9, NY + 10, NX + 10, NZ + 10))
    t = np.arange(x.shape[AXES[axis]], dtype=np.float64) - 5.0
    shape = [1, 1, 1]
    shape[AXES[axis] - 1] = -1
    x[IDX[field]] = _poly(deg, t).reshape(shape)
    y = O.el3_fwd(x, coef)
    for out, terms in SYSTEM.items():
        exp = np.zeros((NY + 10, NX + 10, NZ + 10))
        for cn, fn, ax, sgn in terms:
            if fn == field and ax == axis:
                exp = exp + CVAL[cn] * _dpoly(deg, t + 0.5 * sgn).reshape(shape)
        exp = np.broadcast_to(exp, x.shape[1:])
        got = y[IDX[out]]
        assert np.allclose(got[5:-5, 5:-5, 5:-5], exp[5:-5, 5:-5, 5:-5], rtol=0, atol=1e-9), (out, field, axis)


# relabelling of fields / coefficients under a swap of two axes
def _swap_names(a, b):
    ren = {}
    for n in O.EL3_FIELDS + O.EL3_COEFS:
        s = n.replace(a, "#").replace(b, a).replace("#", b)
        if s.startswith("s") and len(s) == 3:
            s = "s" + "".join(sorted(s[1:], key="xyz".index))
        ren[n] = s
    return ren


def test_step_symmetric_under_axis_swaps_cube():
    n = 10
    rng = np.random.default_rng(9)
    shp = (n + 10,) * 3
    coef = rng.uniform(-1, 1, size=(6,) + shp)
    u, v, o = (rng.uniform(-1, 1, size=(9,) + shp) for _ in range(3))
    base = O.el3_step(u, v, o, coef)
    for a, b in [("x", "y"), ("x", "z"), ("y", "z")]:
        ren = _swap_names(a, b)
        swap = (AXES[a], AXES[b])

        def perm(arr, names):
            idx = {nm: i for i, nm in enumerate(names)}
            out = np.empty_like(arr)
            for nm in names:
                out[idx[ren[nm]]] = np.swapaxes(arr[idx[nm]][None], *swap)[0]
            return out

        got = O.el3_step(perm(u, O.EL3_FIELDS), perm(v, O.EL3_FIELDS), perm(o, O.EL3_FIELDS),
                         perm(coef, O.EL3_COEFS))
        assert np.allclose(got, perm(base, O.EL3_FIELDS), rtol=0, atol=1e-12), (a, b)


C1 = np.float64(35.0 / 294912.0)
C5 = np.float64(19845.0 / 16384.0)


@pytest.mark.parametrize("out,coef,field,offset,weight", [
    # K3/rtmlib-modC.c:290  Vx += m2m3(ix-5) c1 sxx(ix-5)
    ("Vx", "m2m3", "sxx", ("x", -5), C1),
    # K3/rtmlib-modC.c:326  Vx += m3(iy-4) c1 sxy(iy-4)
    ("Vx", "m3", "sxy", ("y", -4), C1),
    # K3/rtmlib-modC.c:347  Vx -= m3(iz+5) c1 sxz(iz+5)
    ("Vx", "m3", "sxz", ("z", +5), -C1),
    # K3/rtmlib-modC.c:409  Vy -= m3(ix+5) c1 sxy(ix+5)
    ("Vy", "m3", "sxy", ("x", +5), -C1),
    # K3/rtmlib-modC.c:483  sxx -= m1_x(ix+1) c5 Vx(ix+1)
    ("sxx", "m1_x", "Vx", ("x", +1), -C5),
    # K3/rtmlib-modC.c:519  sxy += m1_y(ix-4) c1 Vy(ix-4)
    ("sxy", "m1_y", "Vy", ("x", -4), C1),
    # K3/rtmlib-modC.c:587  syz -= m1_z(iy+5) c1 Vz(iy+5)
    ("syz", "m1_z", "Vz", ("y", +5), -C1),
])
def test_single_taps_from_the_reference_source(out, coef, field, offset, weight):
    """An impulse in one input field, one coefficient array at 1, all others
    zero: the output cell at the tap's distance gets exactly the weight."""
    cf = np.zeros((6, NY + 10, NX + 10, NZ + 10))
    cf[CDX[coef]] = 1.0
    u = np.zeros((9, NY + 10, NX + 10, NZ + 10))
    pos = [5 + 4, 5 + 5, 5 + 5]  # output cell, interior
    ax, off = offset
    src = list(pos)
    src[AXES[ax] - 1] += off  # output i reads i + off
    u[IDX[field], src[0], src[1], src[2]] = 1.0
    g = _gather(u, cf)
    assert g[IDX[out], pos[0], pos[1], pos[2]] == weight


def test_fp32_reference_order_close_to_fp64():
    u, v, o = (O.el3_random(NY, NZ, NX, s) for s in (11, 12, 13))
    coef = O.el3_random(NY, NZ, NX, 14, k=6)
    g64 = O.el3_step(u, v, o, coef)
    g32 = O.el3_step(u, v, o, coef, dtype=np.float32)
    assert 0 < O.rel_l2(g32, g64) <= 1e-6
    # rim untouched
    assert np.array_equal(g32[:, :5], o[:, :5]) and np.array_equal(g32[..., -5:], o[..., -5:])


def test_rotation_matches_the_reference_pointer_swaps():
    """(out, in, in1) <- (in1, out, in) after every step but the last: the
    next step's in is the level just written, its in1 the previous in, and
    its out the buffer that held in1 (added to, as the reference does)."""
    b = [O.el3_random(NY, NZ, NX, s) for s in (21, 22, 23)]
    coef = O.el3_random(NY, NZ, NX, 24, k=6)
    fin, fin1, fout = O.el3_steps(b[0], b[1], b[2], coef, 3)
    s1 = O.el3_step(b[0], b[1], b[2], coef)          # level L
    s2 = O.el3_step(s1, b[0], b[1], coef)            # level L-1 into the old in1
    s3 = O.el3_step(s2, s1, b[0], coef)              # level L-2 into the old in
    assert np.array_equal(fout, s3)
    assert np.array_equal(fin, s2) and np.array_equal(fin1, s1)
    one = O.el3_steps(b[0], b[1], b[2], coef, 1)
    assert np.array_equal(one[2], s1) and np.array_equal(one[0], b[0])
