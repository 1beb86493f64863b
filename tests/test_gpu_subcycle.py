"""GPU parity of the subcycled occupancy (lbm_subcycle / lbm_subcycle_adjoint: the moving
field's SubcycledField::subcycle and subcycle_adjoint, crates/implexity-physics-lbm/src/
moving/field.rs:1870-1919 and :2055-2392) through the C ABI against the oracle.

Bars: FP64 end states and impulse shares g bit-identical to the oracle for the Psm,
PsmSuperposition and moving-Brinkman laws on D3Q19 and D3Q27; on D3Q19 the transpose
(start-state, both occupancies', both momenta's cotangents) within 1e-11 of the oracle,
also when the substeps are recomputed under a two-slot inner schedule; the handle keeps
the start state; admission refuses as Field::check_state; FP32 states within 2e-5 of the
deviation scale.
"""
from __future__ import annotations

import numpy as np
import pytest

from implexity_amd.lbm import ContractError, ConvergenceError, Lattice, optimal_forward_steps
from oracle import oracle as orc
from tests.lbm_cases import channel_ports, make_pair, random_alpha

pytestmark = pytest.mark.gpu

CASES = [
    # shape, periodic, Q, collision, law, with_ports, substeps
    ((7, 6, 5), (True, False, False), 19, "trt", "psm", False, 4),
    ((9, 6, 6), (False, False, False), 19, "trt", "psm_superposition", True, 3),
    ((6, 5, 7), (True, True, True), 19, "bgk", "brinkman_occupancy", False, 5),
    ((6, 5, 4), (True, False, True), 27, "trt", "psm", False, 4),
    ((6, 4, 5), (True, True, False), 27, "bgk", "brinkman_occupancy", False, 2),
]


def _case(shape, periodic, Q, coll, law, with_ports, precision="f64", seed=9):
    n = int(np.prod(shape))
    rng = np.random.default_rng(seed)
    d0 = random_alpha(shape, seed) * 10.0
    d0[rng.random(n) < 0.4] = 0.0
    d0[rng.random(n) < 0.1] = 1.2
    if with_ports:  # keep the port planes and their neighbours free of solid
        d0.reshape(shape)[:2] = 0.0
        d0.reshape(shape)[-2:] = 0.0
    m0 = rng.standard_normal((n, 3)) * 0.02 * d0[:, None]
    d1 = np.roll(d0.reshape(shape), 1, axis=1).reshape(-1)
    m1 = np.roll(m0.reshape(shape + (3,)), 1, axis=1).reshape(n, 3) + rng.standard_normal((n, 3)) * 1e-3 * (d1 > 0)[:, None]
    ports = channel_ports(shape, 0.03) if with_ports else ()
    cfg, prob = make_pair(shape, periodic, Q=Q, collision=coll, precision=precision, omega=1.6,
                          accel=(0.0, 0.0, 0.0) if with_ports else (1e-4, 0.0, 0.0), ports=ports,
                          coupling=(law, 0.1), solid_momentum=m0, drag=(0.08, 0.1))
    prob.occupancy_end = (d1, m1, 0, 1)
    return cfg, prob, d0, m0, d1, m1


def _lattice(cfg, d0, m0, d1, m1):
    lat = Lattice(cfg)
    lat.set_design(d0)
    lat.set_solid_momentum(m0)
    lat.set_occupancy_end(d1, m1, 0, 1)
    lat.initial_populations()
    return lat


@pytest.mark.parametrize("shape,periodic,Q,coll,law,with_ports,m", CASES)
def test_subcycle_f64_bit_exact(shape, periodic, Q, coll, law, with_ports, m):
    cfg, prob, d0, m0, d1, m1 = _case(shape, periodic, Q, coll, law, with_ports)
    o = orc.Oracle(prob)
    f = o.run(d0, 3)
    # two macro steps in a row: the second starts its blend at the step count it is at
    s1, g1_want = o.subcycle(d0, f, m, t0=3)
    s2, g2_want = o.subcycle(d0, s1[-1], m, t0=3 + m)
    with _lattice(cfg, d0, m0, d1, m1) as lat:
        lat.interval(3)
        g1, st1 = lat.subcycle(m)
        mid = lat.state()
        g2, _ = lat.subcycle(m)
        end = lat.state()
    assert np.array_equal(mid, s1[-1]), np.abs(mid - s1[-1]).max()
    assert np.array_equal(end, s2[-1]), np.abs(end - s2[-1]).max()
    assert np.array_equal(g1, g1_want), np.abs(g1 - g1_want).max()
    assert np.array_equal(g2, g2_want)
    assert st1.forward_steps == m and np.abs(g1).max() > 0


# The D3Q27 transpose took a GPU memory fault (FP64, TRT Psm law) whose cause is not
# established; lbm_subcycle_adjoint refuses D3Q27 until it is, and these cases stay off.
Q27_REASON = "D3Q27 subcycle transpose refused: GPU memory fault with no established cause"
ADJ_CASES = [c if c[2] == 19 else pytest.param(*c, marks=pytest.mark.skip(reason=Q27_REASON)) for c in CASES]


@pytest.mark.parametrize("shape,periodic,Q,coll,law,with_ports,m", ADJ_CASES)
@pytest.mark.parametrize("slots", [0, 2])
def test_subcycle_adjoint_f64_matches_oracle(shape, periodic, Q, coll, law, with_ports, m, slots):
    cfg, prob, d0, m0, d1, m1 = _case(shape, periodic, Q, coll, law, with_ports)
    n = int(np.prod(shape))
    o = orc.Oracle(prob)
    f = o.run(d0, 3)
    states, _ = o.subcycle(d0, f, m, t0=3)
    rng = np.random.default_rng(4)
    sb = rng.standard_normal(n * Q)
    gb = rng.standard_normal((n, 2, 3))
    want = o.subcycle_adjoint(d0, states, m, sb, gb, t0=3)
    if slots:
        with Lattice(cfg) as lat:
            cfg.checkpoint_budget_bytes = lat.state_bytes * slots
    with _lattice(cfg, d0, m0, d1, m1) as lat:
        lat.interval(3)
        start = lat.state()
        got = lat.subcycle_adjoint(m, sb, gb)
        held = lat.state()
        st = got[-1]
    assert np.array_equal(held, start)
    if slots:  # the inner schedule's recomputation (Field::prepare's BinomialSchedule)
        assert st.forward_steps == optimal_forward_steps(m, slots)
    for g, w, name in zip(got[:5], want, ("previous", "d0", "m0", "d1", "m1")):
        g, w = np.asarray(g).reshape(-1), np.asarray(w).reshape(-1)
        assert np.abs(g - w).max() <= 1e-11 * np.abs(w).max(), (name, np.abs(g - w).max(), np.abs(w).max())


def test_subcycle_adjoint_without_impulse_cotangents():
    """g_bar = 0 (NULL): the transpose of the blended substeps alone."""
    cfg, prob, d0, m0, d1, m1 = _case(*CASES[0][:6])
    n = int(np.prod(CASES[0][0]))
    o = orc.Oracle(prob)
    f = o.run(d0, 2)
    states, _ = o.subcycle(d0, f, 3, t0=2)
    sb = np.random.default_rng(8).standard_normal(n * 19)
    want = o.subcycle_adjoint(d0, states, 3, sb, None, t0=2)
    with _lattice(cfg, d0, m0, d1, m1) as lat:
        lat.interval(2)
        got = lat.subcycle_adjoint(3, sb)
    for g, w in zip(got[:5], want):
        g, w = np.asarray(g).reshape(-1), np.asarray(w).reshape(-1)
        assert np.abs(g - w).max() <= 1e-11 * max(np.abs(w).max(), 1e-300)


def test_subcycle_admission_refuses_above_the_mach_limit():
    cfg, prob, d0, m0, d1, m1 = _case(*CASES[0][:6])
    cfg.mach_limit = 1e-3
    with _lattice(cfg, d0, m0, d1, m1) as lat:
        with pytest.raises(ConvergenceError, match=r"moving LBM admission refused \(mach\)"):
            lat.subcycle(4)
    cfg.mach_limit = 0.2
    with _lattice(cfg, d0, m0, d1, m1) as lat:
        _, st = lat.subcycle(4)
    assert 0 < st.max_mach < 0.2 and st.min_population > 0


def test_subcycle_adjoint_refuses_d3q27():
    """Refused on the host before any device work."""
    cfg, prob, d0, m0, d1, m1 = _case(*CASES[3][:6])
    with _lattice(cfg, d0, m0, d1, m1) as lat:
        with pytest.raises(ContractError, match="D3Q19"):
            lat.subcycle_adjoint(2, np.zeros(lat.n * 27))


def test_subcycle_needs_the_second_endpoint():
    cfg, prob, d0, m0, d1, m1 = _case(*CASES[0][:6])
    with Lattice(cfg) as lat:
        lat.set_design(d0)
        lat.set_solid_momentum(m0)
        with pytest.raises(ContractError, match="second occupancy endpoint"):
            lat.subcycle(3)


@pytest.mark.parametrize("law", ["psm", "brinkman_occupancy"])
def test_subcycle_f32(law):
    shape = (7, 6, 5)
    cfg, prob, d0, m0, d1, m1 = _case(shape, (True, False, False), 19, "trt", law, False, precision="f32")
    o = orc.Oracle(prob)
    f = o.initial()
    states, g_want = o.subcycle(d0, f, 6)
    with _lattice(cfg, d0, m0, d1, m1) as lat:
        g, _ = lat.subcycle(6)
        got = lat.state()
    _, w, _ = orc.lattice(19)
    dev = states[-1].reshape(-1, 19) - w
    assert np.abs(got.reshape(-1, 19) - states[-1].reshape(-1, 19)).max() <= 2e-5 * np.abs(dev).max()
    assert np.abs(g - g_want).max() <= 1e-4 * np.abs(g_want).max()
