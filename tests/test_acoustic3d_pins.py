"""Pins of the 3D acoustic oracle (oracle/acoustic3d_oracle.cpp): exact
10th-order second-derivative weights, polynomial exactness of the Laplacian,
and the homogeneous-medium Green's function of a band-limited point source."""
from fractions import Fraction

import numpy as np
import pytest

import oracle as O


def test_second_derivative_weights_exact():
    # w0 + 2 sum w_k = 0; 2 sum w_k k^2 = 2; 2 sum w_k k^(2m) = 0, m = 2..5
    n = 5
    A = [[Fraction(2 * k ** (2 * m)) for k in range(1, n + 1)] for m in range(1, n + 1)]
    b = [Fraction(2 if m == 1 else 0) for m in range(1, n + 1)]
    M = [r[:] + [bb] for r, bb in zip(A, b)]
    for c in range(n):
        p = next(r for r in range(c, n) if M[r][c] != 0)
        M[c], M[p] = M[p], M[c]
        M[c] = [v / M[c][c] for v in M[c]]
        for r in range(n):
            if r != c:
                f = M[r][c]
                M[r] = [x - f * y for x, y in zip(M[r], M[c])]
    wk = [M[k][n] for k in range(n)]
    w0 = -2 * sum(wk)
    expect = [Fraction(-5269, 1800), Fraction(5, 3), Fraction(-5, 21), Fraction(5, 126), Fraction(-5, 1008),
              Fraction(1, 3150)]
    assert [w0] + wk == expect
    assert np.array_equal(O.ac3_weights(), np.array([float(v) for v in expect]))


@pytest.mark.parametrize("deg", range(0, 12))
def test_laplacian_polynomial_exact(deg):
    n1, n2, n3 = 16, 17, 18
    z = (np.arange(n1) - 8.0) / 16
    x = (np.arange(n2) - 8.0) / 16
    y = (np.arange(n3) - 9.0) / 16
    Z, X, Y = np.meshgrid(z, x, y, indexing="ij")
    p = Z ** deg + 0.5 * X ** max(deg - 1, 0) - Y ** deg
    h = 1.0 / 16
    d2 = lambda u, k: k * (k - 1) * u ** (k - 2) if k >= 2 else 0 * u
    exact = (d2(Z, deg) + 0.5 * d2(X, max(deg - 1, 0)) - d2(Y, deg)) * h * h
    lap = O.ac3_laplacian(p)
    sl = (slice(5, n1 - 5), slice(5, n2 - 5), slice(5, n3 - 5))
    assert np.allclose(lap[sl], exact[sl], rtol=0, atol=1e-13)
    assert not lap[:5].any() and not lap[:, :, -5:].any()


def test_homogeneous_greens_function():
    """p(r, t) = amp(t + 1 - r / sqrt(c)) / (4 pi c r) in index units for the
    leapfrog p_next = 2p - p_prev + c lap p with the amplitude added to level t
    (which acts as forcing of the update centred on t - 1). c = 0.16 is below
    the 3D 10th-order stability limit 4 / 20.5."""
    n, c, nt = 61, 0.16, 98
    t = np.arange(1, nt + 1, dtype=np.float64)
    f, t0 = 1.0 / 30.0, 35.0  # 30 steps per period: time-stepping dispersion ~2.5 %
    a = (np.pi * f * (t - t0)) ** 2
    amp = (1 - 2 * a) * np.exp(-a)
    rec, _ = O.ac3_run(np.full((n, n, n), c), amp, (30, 30, 30), zr=30 - 12, abc=0)
    trace = rec[:, 30, 30]  # receiver 12 cells above the source; no reflection before nt
    r = 12.0
    tt = t + 1 - r / np.sqrt(c)
    aa = (np.pi * f * (tt - t0)) ** 2
    ana = (1 - 2 * aa) * np.exp(-aa) / (4 * np.pi * c * r)
    assert abs(trace.max() / ana.max() - 1) < 0.02
    err = np.linalg.norm(trace - ana) / np.linalg.norm(ana)
    assert err < 0.05, err
