"""Pins of the oracle's admission (LbmProblem::admit, solver.rs:707-793), run without a
GPU: closed forms on states whose Mach number, mass and smallest population are known,
and the refusal categories, each chosen so that a plausible slip (the wrong velocity,
solid cells counted, the exchange left out, a drift not normalised) fails one of them.
"""
from __future__ import annotations

import numpy as np

from oracle import oracle as orc
from tests.lbm_cases import channel_ports, make_pair, random_alpha


def _equilibrium(Q, n, rho, u):
    c, w, _ = orc.lattice(Q)
    cu = c @ u
    cell = rho * w * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * u @ u)
    return np.tile(cell, n)


def _periodic(Q=19, solid=None, accel=(0.0, 0.0, 0.0)):
    return orc.Oracle(orc.Problem(shape=(4, 3, 5), periodic=(True, True, True), Q=Q, kind="trt", omega=1.3,
                                  solid=solid, accel=accel))


def test_rest_state_closed_form():
    """Equilibrium at rest: Mach 0, no drift, smallest population the smallest weight."""
    o = _periodic()
    alpha = np.full(o.n, 0.02)
    states, changes = orc.trajectory(o, alpha, 5, f0=_equilibrium(19, o.n, 1.0, np.zeros(3)))
    refusal, mach, drift, fmin = orc.admit(o, alpha, states, changes, 0.1)
    assert refusal is None
    assert mach < 1e-15 and drift < 1e-15
    assert fmin == 1.0 / 36.0


def test_uniform_flow_mach_is_brinkman_speed():
    """A uniform equilibrium with alpha = 0 and no force has Brinkman velocity u0 exactly,
    so Mach = sqrt(3) |u0|; with alpha > 0 the Brinkman velocity (j/rho + a/2)/(1 + alpha/2)
    sets it (a slip to j/rho fails the second case)."""
    u0 = np.array([0.04, -0.01, 0.02])
    for alpha_v in (0.0, 0.2):
        o = _periodic()
        alpha = np.full(o.n, alpha_v)
        f0 = _equilibrium(19, o.n, 1.0, u0)
        refusal, mach, _, _ = orc.admit(o, alpha, [f0], [], 0.2)
        want = np.sqrt(3.0) * np.linalg.norm(u0) / (1 + alpha_v / 2)
        assert abs(mach - want) < 1e-15
        assert refusal is None
        assert orc.admit(o, alpha, [f0], [], 0.9 * want)[0] == "rejected"
        assert orc.admit(o, alpha, [f0], [], 1.1 * want)[0] is None


def test_negative_population_only_counts_in_fluid_cells():
    solid = np.zeros((4, 3, 5), bool)
    solid[1, 1, 2] = True
    o = _periodic(solid=solid)
    alpha = np.zeros(o.n)
    f = _equilibrium(19, o.n, 1.0, np.zeros(3))
    x_solid = (1 * 3 + 1) * 5 + 2
    g = f.copy()
    g[x_solid * 19 + 4] = -1e-3  # in the solid cell: not a fluid population
    assert orc.admit(o, alpha, [f, g], [0.0], 0.1)[0] is None
    g = f.copy()
    g[7 * 19 + 4] = -1e-3  # in a fluid cell
    refusal, _, _, fmin = orc.admit(o, alpha, [f, g], [0.0], 0.1)
    assert refusal == "populations" and fmin == -1e-3
    g = f.copy()
    g[x_solid * 19 + 4] = np.nan  # non-finite anywhere refuses
    assert orc.admit(o, alpha, [f, g], [0.0], 0.1)[0] == "populations"


def test_drift_relative_to_initial_mass():
    """A state whose fluid mass is 1 + 3e-6 times the initial mass drifts 3e-6 (refused at
    the reference's 1e-6, admitted at 1e-5)."""
    o = _periodic()
    alpha = np.zeros(o.n)
    f = _equilibrium(19, o.n, 1.2, np.zeros(3))
    g = f * (1 + 3e-6)
    refusal, _, drift, _ = orc.admit(o, alpha, [f, f, g], [0.0, 0.0], 0.1)
    assert abs(drift - 3e-6) < 1e-15 and refusal == "rejected"
    assert orc.admit(o, alpha, [f, f, g], [0.0, 0.0], 0.1, drift_limit=1e-5)[0] is None


def test_port_exchange_is_booked():
    """Through an inlet/outlet channel the fluid mass changes, and the port exchange
    recorded with every step accounts for it to rounding; leaving the exchange out
    (or booking it one step late) leaves a drift orders of magnitude larger."""
    shape = (7, 5, 6)
    _, prob = make_pair(shape, Q=19, ports=channel_ports(shape, 0.03))
    o = orc.Oracle(prob)
    alpha = random_alpha(shape, 4)
    states, changes = orc.trajectory(o, alpha, 12)
    refusal, _, drift, _ = orc.admit(o, alpha, states, changes, 0.2)
    assert refusal is None and drift < 1e-14
    _, _, drift_none, _ = orc.admit(o, alpha, states, [0.0] * len(changes), 0.2)
    _, _, drift_late, _ = orc.admit(o, alpha, states, [0.0] + changes[:-1], 0.2)
    assert drift_none > 1e3 * max(drift, 1e-16) and drift_late > 1e3 * max(drift, 1e-16)


def test_closed_box_conserves_mass():
    """Halfway bounce-back walls and a body force: the fluid mass is conserved to rounding."""
    o = orc.Oracle(orc.Problem(shape=(6, 5, 4), periodic=(False, False, False), Q=19, kind="trt", omega=1.1,
                               accel=(1e-4, -5e-5, 2e-5)))
    alpha = random_alpha((6, 5, 4), 2)
    states, changes = orc.trajectory(o, alpha, 15)
    refusal, mach, drift, fmin = orc.admit(o, alpha, states, changes, 0.2)
    assert refusal is None and drift < 1e-14 and 0.0 < mach < 0.01 and fmin > 0.0
