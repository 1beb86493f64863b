"""Pins of the push-forward restatement (moving/pushforward.rs)."""
from __future__ import annotations

import numpy as np
import pytest

from oracle.pushforward import PushForwardOracle, axis_weights, m4, peskin4


@pytest.mark.parametrize("kernel,h", [("cubic", 1), ("cubic", 2), ("peskin4", 1), ("peskin4", 2)])
def test_kernels_partition_unity_and_reproduce_linear_functions(kernel, h):
    for x in (3.0, 3.25, 3.5, 3.999, 7.13):
        first, w, dw = axis_weights(x, h, kernel)
        c = np.arange(first, first + 4 * h, dtype=float)
        assert abs(w.sum() - 1.0) <= 1e-14
        assert abs(c @ w - x) <= 1e-13
        assert abs(dw.sum()) <= 1e-13 and abs(c @ dw - 1.0) <= 1e-12


def test_width_one_closed_form_is_m4():
    for x in (2.0, 2.3, 2.5, 2.99):
        first, w, dw = axis_weights(x, 1, "cubic")
        for t in range(4):
            assert abs(w[t] - m4(first + t - x)) <= 1e-15
        if x != 2.0:  # the stencil's first cell moves at integer coordinates
            h = 1e-7
            fd = (axis_weights(x + h, 1, "cubic")[1] - axis_weights(x - h, 1, "cubic")[1]) / (2 * h)
            assert np.abs(fd - dw).max() <= 1e-7


def test_peskin_even_odd_sums():
    for r in (0.0, 0.2, 0.5, 0.8):
        assert abs(peskin4(r) + peskin4(r - 2.0) - 0.5) <= 1e-15
        assert abs(peskin4(r - 1.0) + peskin4(r + 1.0) - 0.5) <= 1e-15


@pytest.mark.parametrize("kernel,h", [("cubic", 1), ("peskin4", 2)])
def test_scatter_moments(kernel, h):
    shape = (14, 12, 16)
    rng = np.random.default_rng(0)
    xi = rng.uniform(4.5, 7.5, (40, 3))
    a = rng.random(40)
    v = rng.standard_normal((40, 3))
    o = PushForwardOracle(shape, (False, False, False), h, kernel)
    d, m = o.scatter(xi, a, v)
    assert abs(d.sum() - a.sum()) <= 1e-13
    assert np.abs(m.sum(0) - a @ v).max() <= 1e-13
    idx = np.arange(o.n)
    co = np.stack([idx // (12 * 16), (idx // 16) % 12, idx % 16], 1).astype(float)
    assert np.abs(d @ co - a @ xi).max() <= 1e-11


def test_support_outside_a_closed_axis_is_refused():
    o = PushForwardOracle((8, 8, 8))
    with pytest.raises(ValueError, match="outside"):
        o.scatter([[0.5, 4.0, 4.0]], [1.0])
    PushForwardOracle((8, 8, 8), (True, False, False)).scatter([[0.5, 4.0, 4.0]], [1.0])


@pytest.mark.parametrize("kernel,h,periodic", [("cubic", 1, (False, False, False)), ("peskin4", 1, (True, True, False)),
                                               ("cubic", 2, (True, False, True))])
def test_vjp_matches_central_differences(kernel, h, periodic):
    shape = (12, 14, 14)
    rng = np.random.default_rng(1)
    xi = rng.uniform(4.6, 6.9, (6, 3))
    a = rng.random(6)
    v = rng.standard_normal((6, 3))
    o = PushForwardOracle(shape, periodic, h, kernel)
    db = rng.standard_normal(o.n)
    mb = rng.standard_normal((o.n, 3))

    def phi(x, aa, vv):
        d, m = o.scatter(x, aa, vv)
        return db @ d + (mb * m).sum()

    xb, ab, vb = o.vjp(xi, a, v, db, mb)
    e = 1e-6
    for q in range(6):
        for c in range(3):
            xp, xm = xi.copy(), xi.copy()
            xp[q, c] += e
            xm[q, c] -= e
            assert abs((phi(xp, a, v) - phi(xm, a, v)) / (2 * e) - xb[q, c]) <= 1e-6 * max(1.0, abs(xb[q, c]))
            vp, vm = v.copy(), v.copy()
            vp[q, c] += e
            vm[q, c] -= e
            assert abs((phi(xi, a, vp) - phi(xi, a, vm)) / (2 * e) - vb[q, c]) <= 1e-7
        ap, am = a.copy(), a.copy()
        ap[q] += e
        am[q] -= e
        assert abs((phi(xi, ap, v) - phi(xi, am, v)) / (2 * e) - ab[q]) <= 1e-7
