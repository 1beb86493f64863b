"""Pins of the oracle's subcycled occupancy (SubcycledField::subcycle / subcycle_adjoint on
the moving field, crates/implexity-physics-lbm/src/moving/field.rs:1499-1541 accumulate,
:1870-1919 subcycle, :2055-2392 reverse, :2482-2848 reverse_substep):

* one substep blends at theta = 1/2: it equals a plain step with the endpoints' mean
  occupancy and momentum bit for bit, and splits the impulse evenly (g0 = g1);
* with equal endpoints the impulse shares add up to the momentum the coupled cells
  exchange, sum_x d (g0 + g1) . w = -sum_k w . dp_k, the solid-force functional of the
  states the substeps collide (itself pinned in tests/test_oracle_psm.py);
* the transpose passes central differences in the start state, both occupancies and
  both solid momenta, for <state_bar, f_m> + <g_bar, g>.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as orc
from tests.lbm_cases import random_alpha

LAWS = ["psm", "psm_superposition", "brinkman_occupancy"]


def _setup(law, shape=(6, 5, 4), Q=19, coll="trt", seed=3, equal=False):
    n = int(np.prod(shape))
    rng = np.random.default_rng(seed)
    d0 = random_alpha(shape, seed) * 10.0
    d0[rng.random(n) < 0.4] = 0.0
    d0[rng.random(n) < 0.1] = 1.2
    m0 = rng.standard_normal((n, 3)) * 0.02 * d0[:, None]
    if equal:
        d1, m1 = d0.copy(), m0.copy()
    else:
        d1 = np.roll(d0.reshape(shape), 1, axis=0).reshape(-1)
        m1 = np.roll(m0.reshape(shape + (3,)), 1, axis=0).reshape(n, 3) + rng.standard_normal((n, 3)) * 1e-3
    prob = orc.Problem(shape=shape, periodic=(True, False, True), Q=Q, kind=coll, omega=1.6, accel=(1e-4, 0, 0),
                       coupling=law, psm_kappa=0.1, psm_c_alpha=0.08, psm_drag_shape=0.1, solid_momentum=m0,
                       occupancy_end=(d1, m1, 0, 1))
    return prob, d0, m0, d1, m1


@pytest.mark.parametrize("law", LAWS)
def test_single_substep_is_the_midpoint_step(law):
    prob, d0, m0, d1, m1 = _setup(law)
    o = orc.Oracle(prob)
    f = o.run(d0, 3)
    states, g = o.subcycle(d0, f, 1)
    mid = orc.Problem(**{**vars(prob), "solid_momentum": (m0 + m1) * 0.5, "occupancy_end": None})
    om = orc.Oracle(mid)
    dm = (d0 + d1) * 0.5
    assert np.array_equal(states[1], om.interval(dm, f))
    assert np.array_equal(g[:, 0], g[:, 1])
    # g0 = (-dp / d) / 2 with dp the exchange of the plain step: sum d g0 . w = -(1/2) w . dp
    w = np.array([0.3, -1.1, 0.7])
    occupied = dm > 0
    lhs = float(np.sum(dm[occupied] * (g[occupied, 0] @ w)))
    assert abs(lhs + 0.5 * om.solid_force(dm, f, w)) <= 1e-14 * max(1.0, abs(lhs))


@pytest.mark.parametrize("law", LAWS)
def test_equal_endpoints_impulse_sums_to_exchange(law):
    prob, d0, m0, _, _ = _setup(law, equal=True)
    o = orc.Oracle(prob)
    f = o.run(d0, 2)
    m = 5
    states, g = o.subcycle(d0, f, m)
    single = orc.Oracle(orc.Problem(**{**vars(prob), "occupancy_end": None}))
    w = np.array([1.0, 0.4, -0.6])
    want = -sum(single.solid_force(d0, states[k], w) for k in range(m))
    got = float(np.sum(d0[:, None] * (g[:, 0] + g[:, 1]) @ w))
    assert abs(got - want) <= 1e-12 * abs(want)
    # the midpoint weights sum to m / 2 per endpoint: a steady exchange splits evenly
    assert abs(sum((k - 0.5) / m for k in range(1, m + 1)) - m / 2) < 1e-15


@pytest.mark.parametrize("law,Q,coll", [("psm", 19, "trt"), ("psm_superposition", 19, "bgk"),
                                        ("brinkman_occupancy", 19, "trt"), ("psm", 27, "trt")])
def test_subcycle_adjoint_central_differences(law, Q, coll):
    prob, d0, m0, d1, m1 = _setup(law, Q=Q, coll=coll, seed=11)
    n = int(np.prod(prob.shape))
    o = orc.Oracle(prob)
    f = o.run(d0, 2)
    m = 4
    rng = np.random.default_rng(5)
    sb = rng.standard_normal(n * Q)
    gb = rng.standard_normal((n, 2, 3))

    def J(fp, a0, a1, b0, b1):
        p = orc.Problem(**{**vars(prob), "solid_momentum": b0, "occupancy_end": (a1, b1, 0, 1)})
        oo = orc.Oracle(p)
        st, g = oo.subcycle(a0, fp, m)
        return float(sb @ st[-1] + np.sum(gb * g))

    states, _ = o.subcycle(d0, f, m)
    pb, d0b, m0b, d1b, m1b = o.subcycle_adjoint(d0, states, m, sb, gb)
    h = 1e-6
    occ0, occ1 = d0 > 0, d1 > 0
    dirs = [
        (rng.standard_normal(n * Q) * 1e-3, 0, pb),
        (rng.standard_normal(n) * occ0 * 0.05, 1, d0b),
        (rng.standard_normal(n) * occ1 * 0.05, 2, d1b),
        (rng.standard_normal((n, 3)) * 1e-3, 3, m0b),
        (rng.standard_normal((n, 3)) * 1e-3, 4, m1b),
    ]
    base = [f, d0, d1, m0, m1]
    for v, which, bar in dirs:
        plus = [x.copy() for x in base]
        minus = [x.copy() for x in base]
        plus[which] = plus[which] + h * v
        minus[which] = minus[which] - h * v
        fd = (J(plus[0], plus[1], plus[2], plus[3], plus[4]) - J(minus[0], minus[1], minus[2], minus[3],
                                                                    minus[4])) / (2 * h)
        ad = float(np.sum(bar * v))
        assert abs(fd - ad) <= 2e-7 * max(abs(fd), 1e-3), (which, fd, ad)
