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


def _images(x, shape, mirrored, h):
    """The point and those of its reflections in the symmetry faces (about -1/2 and
    n - 1/2) whose support reaches the lattice."""
    per_axis = []
    for e in range(3):
        opts = [x[e]]
        if mirrored[e][0]:
            opts.append(-1.0 - x[e])
        if mirrored[e][1]:
            opts.append(2.0 * shape[e] - 1.0 - x[e])
        per_axis.append([y for y in opts if -2.0 * h < y < shape[e] - 1.0 + 2.0 * h])
    return [(i, j, k) for i in per_axis[0] for j in per_axis[1] for k in per_axis[2]]


@pytest.mark.parametrize("kernel,h", [("cubic", 1), ("cubic", 2), ("peskin4", 1), ("peskin4", 2)])
def test_symmetry_face_is_the_point_and_its_mirror_image(kernel, h):
    """fold_axis reflects the support about the face; the kernels are even, so the folded
    scatter is the unfolded scatter of the point and its images, restricted to the lattice."""
    shape = (4 * h + 6, 4 * h + 4, 4 * h + 7)
    mirrored = ((True, True), (False, False), (True, False))
    rng = np.random.default_rng(5)
    n = 24
    xi = np.empty((n, 3))
    xi[:, 0] = np.where(rng.random(n) < 0.5, rng.uniform(-0.5, 2.0 * h, n),
                        rng.uniform(shape[0] - 1.0 - 2.0 * h, shape[0] - 0.5, n))
    xi[:, 1] = rng.uniform(2.0 * h, shape[1] - 2.0 * h - 1.0, n)
    xi[:, 2] = rng.uniform(-0.5, 2.5 * h, n)
    a, v = rng.random(n), rng.standard_normal((n, 3))
    o = PushForwardOracle(shape, (False, False, False), h, kernel, mirrored)
    d, m = o.scatter(xi, a, v)
    pad = 5 * h + 3
    big = tuple(s + 2 * pad for s in shape)
    ob = PushForwardOracle(big, (False, False, False), h, kernel)
    pts, aa, vv = [], [], []
    for q in range(n):
        for img in _images(xi[q], shape, mirrored, h):
            pts.append(np.asarray(img) + pad)
            aa.append(a[q])
            vv.append(v[q])
    db, mb = ob.scatter(np.array(pts), np.array(aa), np.array(vv))
    crop = (slice(pad, pad + shape[0]), slice(pad, pad + shape[1]), slice(pad, pad + shape[2]))
    assert np.abs(d.reshape(shape) - db.reshape(big)[crop]).max() <= 1e-14
    assert np.abs(m.reshape(shape + (3,)) - mb.reshape(big + (3,))[crop]).max() <= 1e-14
    assert abs(d.sum() - a.sum()) <= 1e-13


@pytest.mark.parametrize("kernel,h", [("cubic", 1), ("peskin4", 2)])
def test_planar_axis_is_the_sum_over_a_periodic_axis(kernel, h):
    rng = np.random.default_rng(6)
    n = 20
    xi = np.column_stack([rng.uniform(-30.0, 30.0, n), rng.uniform(2.0 * h, 10.0 - 2.0 * h - 1.0, n),
                          rng.uniform(0.0, 12.0, n)])
    a, v = rng.random(n), rng.standard_normal((n, 3))
    o = PushForwardOracle((1, 10, 12), (True, False, True), h, kernel)
    ob = PushForwardOracle((4 * h, 10, 12), (True, False, True), h, kernel)
    d, m = o.scatter(xi, a, v)
    db, mb = ob.scatter(xi, a, v)
    assert np.abs(d - db.reshape(4 * h, -1).sum(0)).max() <= 1e-14
    assert np.abs(m - mb.reshape(4 * h, -1, 3).sum(0)).max() <= 1e-14
    dbar, mbar = rng.standard_normal(o.n), rng.standard_normal((o.n, 3))
    xb, ab, vb = o.vjp(xi, a, v, dbar, mbar)
    xw, aw, vw = ob.vjp(xi, a, v, np.tile(dbar, 4 * h), np.tile(mbar, (4 * h, 1)))
    assert np.abs(xb[:, 1:] - xw[:, 1:]).max() <= 1e-13
    assert np.all(xb[:, 0] == 0.0) and np.abs(xw[:, 0]).max() <= 1e-13
    assert np.abs(ab - aw).max() <= 1e-13 and np.abs(vb - vw).max() <= 1e-13


@pytest.mark.parametrize("args,msg", [
    (dict(shape=(3, 8, 8), periodic=(True, False, False)), "shorter"),
    (dict(shape=(8, 1, 8)), "planar"),
    (dict(shape=(8, 8, 8), periodic=(True, False, False), mirrored=((True, False), (False, False), (False, False))),
     "mirror"),
    (dict(shape=(8, 8, 7), width=2, mirrored=((False, False), (False, False), (False, True))), "mirror"),
])
def test_frame_checks(args, msg):
    with pytest.raises(ValueError, match=msg):
        PushForwardOracle(**args)


def test_point_beyond_a_symmetry_face_is_refused():
    o = PushForwardOracle((8, 8, 8), mirrored=((True, False), (False, False), (False, False)))
    o.scatter([[-0.5, 4.0, 4.0]], [1.0])
    with pytest.raises(ValueError, match="beyond a symmetry face"):
        o.scatter([[-0.51, 4.0, 4.0]], [1.0])


@pytest.mark.parametrize("kernel,h", [("cubic", 1), ("peskin4", 2)])
def test_vjp_at_symmetry_faces_matches_central_differences(kernel, h):
    shape = (4 * h + 4, 4 * h + 5, 4 * h + 4)
    mirrored = ((True, True), (True, False), (False, False))
    xi = np.array([[0.3, 0.7, 2.0 * h + 0.4], [shape[0] - 1.2, 1.45, 2.0 * h + 1.3], [1.6, 2.0 * h + 0.6, 2.0 * h + 0.2]])
    rng = np.random.default_rng(7)
    a, v = rng.random(3), rng.standard_normal((3, 3))
    o = PushForwardOracle(shape, (False, False, False), h, kernel, mirrored)
    db, mb = rng.standard_normal(o.n), rng.standard_normal((o.n, 3))

    def phi(x):
        d, m = o.scatter(x, a, v)
        return db @ d + (mb * m).sum()

    xb, _, _ = o.vjp(xi, a, v, db, mb)
    e = 1e-6
    for q in range(3):
        for c in range(3):
            xp, xm = xi.copy(), xi.copy()
            xp[q, c] += e
            xm[q, c] -= e
            assert abs((phi(xp) - phi(xm)) / (2 * e) - xb[q, c]) <= 1e-6 * max(1.0, abs(xb[q, c]))
