"""GPU parity of the push-forward of body points into the occupancy
(moving/pushforward.rs: scatter and endpoint_vjp) and of the chain points -> occupancy
-> occupancy-coupled gradient -> point cotangents.

Bars: d and m within 1e-15 (absolute, weights <= 1) of the oracle's point-order sums
(the device sums each cell's points in sorted-bin order) and the same bits on every
call; the transpose within 1e-13;
the chain's FP64 point gradient within 1e-10 of the same chain over the oracle; at
128^3 with 200k points the total occupancy and momentum equal sum a and sum a v.
"""
from __future__ import annotations

import numpy as np
import pytest

from implexity_amd import lbm
from implexity_amd.lbm import Lattice, Objective
from oracle import oracle as orc
from oracle.pushforward import PushForwardOracle
from tests.lbm_cases import channel_ports, make_pair

pytestmark = pytest.mark.gpu

CASES = [
    ((14, 12, 16), (False, False, False), 1, "cubic"),
    ((14, 12, 16), (True, False, True), 2, "cubic"),
    ((12, 14, 10), (True, True, True), 1, "peskin4"),
    ((16, 14, 14), (False, True, False), 2, "peskin4"),
]


def _points(shape, periodic, h, seed, n=40):
    rng = np.random.default_rng(seed)
    lo = np.array([0.0 if p else 2.0 * h for p in periodic])
    hi = np.array([float(s) - (0.0 if p else 2.0 * h + 1.0) for s, p in zip(shape, periodic)])
    xi = lo + (hi - lo) * rng.random((n, 3))
    return xi, rng.random(n), rng.standard_normal((n, 3)) * 0.01


@pytest.mark.parametrize("shape,periodic,h,kernel", CASES)
def test_pushforward_matches_oracle(shape, periodic, h, kernel):
    xi, a, v = _points(shape, periodic, h, 1)
    d_want, m_want = PushForwardOracle(shape, periodic, h, kernel).scatter(xi, a, v)
    d, m = lbm.pushforward(shape, xi, a, v, periodic, h, kernel)
    assert np.abs(d - d_want).max() <= 1e-15
    assert np.abs(m - m_want).max() <= 1e-16


@pytest.mark.parametrize("shape,periodic,h,kernel", CASES)
def test_pushforward_vjp_matches_oracle(shape, periodic, h, kernel):
    xi, a, v = _points(shape, periodic, h, 2)
    n = int(np.prod(shape))
    rng = np.random.default_rng(3)
    db, mb = rng.standard_normal(n), rng.standard_normal((n, 3))
    xw, aw, vw = PushForwardOracle(shape, periodic, h, kernel).vjp(xi, a, v, db, mb)
    xb, ab, vb = lbm.pushforward_vjp(shape, xi, a, v, db, mb, periodic, h, kernel)
    for got, want in ((xb, xw), (ab, aw), (vb, vw)):
        assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()


@pytest.mark.parametrize("shape,periodic,h,kernel", [((1, 9, 10), (True, False, True), 1, "cubic"),
                                                     ((8, 12, 1), (True, True, True), 2, "peskin4")])
def test_pushforward_clustered_and_wrapping_is_deterministic(shape, periodic, h, kernel):
    """Many points per start cell, a planar (one-cell periodic) axis, a periodic axis as
    long as the support, and coordinates outside [0, n) on periodic axes: the oracle's
    sums, and the same bits on a second call."""
    rng = np.random.default_rng(8)
    n = 300
    centre = np.array([s / 2.0 for s in shape])
    xi = centre + rng.standard_normal((n, 3)) * 0.3
    for e, (s_e, p) in enumerate(zip(shape, periodic)):
        if p:
            xi[: n // 3, e] += s_e * rng.integers(-3, 4, n // 3)
        else:
            xi[:, e] = np.clip(xi[:, e], 2.0 * h, s_e - 2.0 * h - 1.0 - 1e-9)
    a, v = rng.random(n), rng.standard_normal((n, 3))
    d_want, m_want = PushForwardOracle(shape, periodic, h, kernel).scatter(xi, a, v)
    d1, m1 = lbm.pushforward(shape, xi, a, v, periodic, h, kernel)
    d2, m2 = lbm.pushforward(shape, xi, a, v, periodic, h, kernel)
    assert np.abs(d1 - d_want).max() <= 1e-13 * np.abs(d_want).max()
    assert np.abs(m1 - m_want).max() <= 1e-13 * np.abs(m_want).max()
    assert np.array_equal(d1.view(np.uint64), d2.view(np.uint64))
    assert np.array_equal(m1.view(np.uint64), m2.view(np.uint64))


MIRROR_CASES = [
    ((12, 10, 14), (False, False, False), ((True, True), (False, False), (True, False)), 1, "cubic"),
    ((12, 10, 14), (False, True, False), ((False, True), (False, False), (True, True)), 2, "cubic"),
    ((9, 12, 10), (False, False, True), ((True, False), (True, True), (False, False)), 1, "peskin4"),
    ((16, 1, 12), (False, True, False), ((True, True), (False, False), (False, True)), 2, "peskin4"),
]


def _face_points(shape, periodic, mirrored, h, seed, n=60):
    """Points crowding the symmetry faces, inside the admissible range elsewhere."""
    rng = np.random.default_rng(seed)
    xi = np.empty((n, 3))
    for e in range(3):
        lo, hi = mirrored[e]
        if periodic[e]:
            xi[:, e] = rng.uniform(-2.0 * shape[e], 3.0 * shape[e], n)
            continue
        a = -0.5 if lo else 2.0 * h - 1.0 + 1e-6
        b = shape[e] - 0.5 if hi else shape[e] - 2.0 * h - 1e-6
        near = np.where(rng.random(n) < 0.5, rng.uniform(a, a + 2.0 * h, n), rng.uniform(b - 2.0 * h, b, n))
        xi[:, e] = np.clip(near, a, b)
    return xi, rng.random(n), rng.standard_normal((n, 3))


@pytest.mark.parametrize("shape,periodic,mirrored,h,kernel", MIRROR_CASES)
def test_pushforward_symmetry_faces_match_oracle(shape, periodic, mirrored, h, kernel):
    xi, a, v = _face_points(shape, periodic, mirrored, h, 10)
    o = PushForwardOracle(shape, periodic, h, kernel, mirrored)
    d_want, m_want = o.scatter(xi, a, v)
    d, m = lbm.pushforward(shape, xi, a, v, periodic, h, kernel, mirrored)
    assert np.abs(d - d_want).max() <= 1e-14 * max(1.0, np.abs(d_want).max())
    assert np.abs(m - m_want).max() <= 1e-14 * max(1.0, np.abs(m_want).max())
    assert abs(d.sum() - a.sum()) <= 1e-12 * a.sum()
    nn = int(np.prod(shape))
    rng = np.random.default_rng(11)
    db, mb = rng.standard_normal(nn), rng.standard_normal((nn, 3))
    xw, aw, vw = o.vjp(xi, a, v, db, mb)
    xb, ab, vb = lbm.pushforward_vjp(shape, xi, a, v, db, mb, periodic, h, kernel, mirrored)
    for got, want in ((xb, xw), (ab, aw), (vb, vw)):
        assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()


def test_pushforward_totals_at_128_cubed():
    shape = (128, 128, 128)
    rng = np.random.default_rng(4)
    xi = 3.0 + 121.0 * rng.random((200_000, 3))
    a = rng.random(200_000) * 1e-3
    v = rng.standard_normal((200_000, 3)) * 0.01
    d, m = lbm.pushforward(shape, xi, a, v, (False, False, False), 1, "cubic")
    assert abs(d.sum() - a.sum()) <= 1e-11 * a.sum()
    assert np.abs(m.sum(0) - a @ v).max() <= 1e-11 * np.abs(a @ v).max() + 1e-15


def _sphere(center, radius, spacing):
    g = np.arange(-radius, radius + 1e-9, spacing)
    pts = np.array([(x, y, z) for x in g for y in g for z in g if x * x + y * y + z * z <= radius * radius])
    return pts + np.asarray(center), np.full(len(pts), spacing ** 3)


@pytest.mark.parametrize("law,kind", [("psm", "pressure_drop"), ("brinkman_occupancy", "pressure_drop"),
                                      ("psm", "solid_force"), ("brinkman_occupancy", "solid_force")])
def test_point_gradient_through_the_coupled_flow(law, kind):
    """points -> (d, m) -> occupancy-coupled channel -> pressure drop or force on the body ->
    (d_bar, m_bar) -> point cotangents, on the device against the same chain over the
    oracle."""
    shape, steps = (14, 10, 10), 16
    xi, a = _sphere((6.6, 4.7, 5.2), 2.2, 0.7)
    v = np.tile([0.01, 0.0, -0.004], (len(a), 1))
    pf = dict(periodic=(False, False, False), width=1, kernel="cubic")
    d, m = lbm.pushforward(shape, xi, a, v, **pf)
    cfg, prob = make_pair(shape, (False, False, False), precision="f64", ports=channel_ports(shape, 0.03),
                          coupling=(law, 0.1), solid_momentum=m, drag=(0.08, 0.1))
    t_end = steps if kind == "pressure_drop" else steps - 1
    ob = Objective(kind=kind, axis=0, plane_a=1, plane_b=shape[0] - 2, t_begin=8, t_end=t_end, weight=1.0 / 9,
                   vector_weights=(1.0, 0.0, 0.0, 0, 0, 0))
    with Lattice(cfg) as lat:
        lat.set_design(d)
        lat.set_solid_momentum(m)
        lat.initial_populations()
        _, db, _, _ = lat.gradient(steps, ob)
        mb = lat.solid_momentum_bar()
    xb, ab, vb = lbm.pushforward_vjp(shape, xi, a, v, db, mb, **pf)

    po = PushForwardOracle(shape, (False, False, False), 1, "cubic")
    d_o, m_o = po.scatter(xi, a, v)
    o = orc.Oracle(prob)
    oob = orc.Objective(kind=kind, window=(8, t_end), weight=1.0 / 9, axis=0, plane_a=1, plane_b=shape[0] - 2,
                        vector_weights=(1.0, 0.0, 0.0, 0, 0, 0))
    mbar = np.zeros((o.n, 3))
    _, db_o, _, _ = orc.scan_adjoint(o, d_o, steps, oob, m_bar=mbar)
    xw, aw, vw = po.vjp(xi, a, v, db_o, mbar)
    for got, want in ((xb, xw), (ab, aw), (vb, vw)):
        assert np.abs(got - want).max() <= 1e-10 * np.abs(want).max()
