"""Pins of the coupled D3Q7 temperature oracle (oracle/thermal.py, lbm_oracle.c).

The thermal model has no counterpart in the reference, so each part is pinned by what
the mathematics fixes: lattice moments, exact heat balance, the closed-form decay and
drift of a Fourier mode, the boundary values the ports impose, and central differences
or the dot-product identity for every transpose.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as orc
from oracle.thermal import CoupledObjective, Thermal, ThermalOracle
from tests.lbm_cases import channel_ports, make_pair, random_alpha


def test_d3q7_tables():
    c, w, opp = orc.lattice(7)
    assert c[0].tolist() == [0, 0, 0]
    # rest first, then the axis vectors in lexicographic order
    assert [tuple(v) for v in c[1:]] == sorted(tuple(v) for v in c[1:])
    assert np.allclose(np.abs(c[1:]).sum(1), 1)
    assert w.sum() == 1.0
    assert np.allclose(w @ c, 0.0)
    assert np.allclose(np.einsum("q,qa,qb->ab", w, c, c), 0.25 * np.eye(3))
    for q in range(7):
        assert (c[opp[q]] == -c[q]).all()
        assert opp[q] == (0 if q == 0 else 7 - q)


def _tor(shape=(6, 5, 4), periodic=(True, False, False), ports=(), accel=(1e-4, 0.0, 0.0), th=None):
    _, prob = make_pair(shape, periodic, Q=19, collision="trt", omega=1.25, accel=accel, ports=ports)
    return ThermalOracle(orc.Oracle(prob), th or Thermal(kappa=0.02, ratio=10.0, source=1e-3, inlet=0.3))


def test_cell_heat_balance_and_equilibrium():
    t = _tor()
    rng = np.random.default_rng(1)
    g = rng.random(7) * 0.1
    u = rng.standard_normal(3) * 0.05
    out = t.cell(g, u, 1.3, 2e-3)
    assert abs(out.sum() - (g.sum() + 2e-3)) < 1e-15
    # the equilibrium is a fixed point without a source; its first moment is T u
    c, w, _ = orc.lattice(7)
    T = 0.7
    eq = w * T * (1.0 + 4.0 * (c @ u))
    assert np.allclose(t.cell(eq, u, 1.3, 0.0), eq, rtol=0, atol=1e-15)
    assert np.allclose(c.T @ eq, T * u, rtol=0, atol=1e-15)


def test_cell_vjp_central_differences():
    t = _tor()
    rng = np.random.default_rng(2)
    g = rng.random(7) * 0.1
    u = rng.standard_normal(3) * 0.05
    om, S = 1.3, 2e-3
    gb = rng.standard_normal(7)
    g_bar, u_bar, om_bar, S_bar = t.cell_vjp(g, u, om, gb)
    h = 1e-6

    def J(g_, u_, om_, S_):
        return float(gb @ t.cell(g_, u_, om_, S_))
    for i in range(7):
        e = np.zeros(7)
        e[i] = h
        assert abs((J(g + e, u, om, S) - J(g - e, u, om, S)) / (2 * h) - g_bar[i]) < 1e-8
    for d in range(3):
        e = np.zeros(3)
        e[d] = h
        assert abs((J(g, u + e, om, S) - J(g, u - e, om, S)) / (2 * h) - u_bar[d]) < 1e-8
    assert abs((J(g, u, om + h, S) - J(g, u, om - h, S)) / (2 * h) - om_bar) < 1e-8
    assert abs((J(g, u, om, S + h) - J(g, u, om, S - h)) / (2 * h) - S_bar) < 1e-8


@pytest.mark.parametrize("accel", [(0.0, 0.0, 0.0), (2e-4, 1e-5, 0.0)])
def test_heat_balance_closed_domain(accel):
    """Walls all round (adiabatic), any flow: total heat grows by sum S per step."""
    shape = (6, 5, 4)
    t = _tor(shape, periodic=(False, False, False), accel=accel)
    alpha = random_alpha(shape, 3)
    phi = np.random.default_rng(4).random(t.n)
    f, g = t.orc.initial(), t.initial()
    g[:] = np.random.default_rng(5).random(g.size) * 0.01
    total = g.sum()
    for _ in range(15):
        f, g = t.step(alpha, phi, f, g)
        total += t.th.source * phi.sum()
        assert abs(g.sum() - total) <= 1e-13 * abs(total)


def _mode_run(u0, steps, kappa, N=64):
    """Periodic N x 2 x 2 lattice, uniform flow u0 along x (equilibrium populations, no
    force, no drag), T = 1 + 0.01 cos(2 pi x / N): returns the complex amplitude of the
    mode after each of the given step counts, and the wavenumber."""
    shape = (N, 2, 2)
    _, prob = make_pair(shape, (True, True, True), Q=19, collision="trt", omega=1.25, accel=(0.0, 0.0, 0.0))
    t = ThermalOracle(orc.Oracle(prob), Thermal(kappa=kappa, ratio=1.0, source=0.0, inlet=0.0))
    c19, w19, _ = orc.lattice(19)
    cu = c19 @ np.array([u0, 0.0, 0.0])
    feq = w19 * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * u0 * u0)
    f = np.tile(feq, t.n)
    x = np.arange(t.n) // 4
    T0 = 1.0 + 0.01 * np.cos(2 * np.pi * x / N)
    c7, w7, _ = orc.lattice(7)
    g = (w7[None, :] * T0[:, None] * (1.0 + 4.0 * (c7 @ np.array([u0, 0.0, 0.0])))[None, :]).reshape(-1)
    alpha = np.zeros(t.n)
    phi = np.zeros(t.n)
    out = []
    done = 0
    for s in steps:
        for _ in range(s - done):
            f, g = t.step(alpha, phi, f, g)
        done = s
        T = t.temperature(g).reshape(N, 4)
        assert np.allclose(T, T[:, :1])
        out.append(np.fft.fft(T[:, 0])[1])
    return out, 2 * np.pi / N


# The rate and phase are measured between two times (100 and 400 steps), so that the
# start-up of the non-equilibrium part (the initial state is not the slow eigenvector)
# cancels; what remains are the O(k^2) lattice corrections (k = 2 pi / 64).
@pytest.mark.parametrize("kappa", [0.01, 0.05])
def test_diffusion_mode_decays_at_kappa_k2(kappa):
    """A resting fluid: the Fourier mode decays as exp(-kappa k^2 t) with
    kappa = (tau_T - 1/2) / 4 (c_s^2 = 1/4)."""
    (a1, a2), k = _mode_run(0.0, (100, 400), kappa)
    rate = np.log(abs(a1) / abs(a2)) / 300
    assert abs(rate / (kappa * k * k) - 1.0) < 5e-3, rate / (kappa * k * k)
    assert abs(np.angle(a2 / a1)) < 1e-12


def test_advected_mode_moves_at_flow_speed():
    """Uniform flow u0: the mode drifts by k u0 t in phase and still decays at kappa k^2."""
    u0, kappa = 0.05, 0.02
    (a1, a2), k = _mode_run(u0, (100, 400), kappa)
    r = a2 / a1
    speed = -np.angle(r) / (k * 300)
    assert abs(speed / u0 - 1.0) < 2e-3, speed / u0
    rate = -np.log(abs(r)) / 300
    assert abs(rate / (kappa * k * k) - 1.0) < 2e-2, rate / (kappa * k * k)


def test_ports_impose_inlet_temperature_and_zero_gradient():
    shape = (7, 5, 4)
    t = _tor(shape, periodic=(False, False, False), ports=channel_ports(shape, 0.03), accel=(0.0, 0.0, 0.0))
    alpha = random_alpha(shape, 6)
    phi = np.random.default_rng(7).random(t.n)
    f, g = t.orc.initial(), t.initial()
    for _ in range(5):
        f, g = t.step(alpha, phi, f, g)
        T = t.temperature(g)
        o = t.orc
        vel = o.port_kind == 0
        assert np.allclose(T[o.port_cell[vel]], t.th.inlet, rtol=0, atol=1e-15)
        assert np.array_equal(g.reshape(-1, 7)[o.port_cell[~vel]], g.reshape(-1, 7)[o.port_nb[~vel]])


def test_thermal_interval_vjp_dot_product():
    """<lam, J v> by central differences of the coupled thermal step equals <J^T lam, v>
    for random directions in f, g, alpha and phi."""
    shape = (7, 5, 4)
    t = _tor(shape, periodic=(False, False, False), ports=channel_ports(shape, 0.03), accel=(0.0, 0.0, 0.0))
    rng = np.random.default_rng(8)
    alpha = random_alpha(shape, 9)
    phi = rng.random(t.n)
    f, g = t.run(alpha, phi, 4)
    g = g + rng.random(g.size) * 0.01
    lam = rng.standard_normal(t.n * 7)
    vf = rng.standard_normal(f.size) * 1e-3
    vg = rng.standard_normal(g.size)
    va = rng.standard_normal(t.n) * 0.01
    vp = rng.standard_normal(t.n)
    h = 1e-6

    def step(s):
        return t.interval(alpha + s * va, phi + s * vp, f + s * vf, g + s * vg)
    lhs = lam @ ((step(h) - step(-h)) / (2 * h))
    gb, fb, ab, pb = t.interval_vjp(alpha, phi, f, g, lam)
    rhs = gb @ vg + fb @ vf + ab @ va + pb @ vp
    assert abs(lhs - rhs) <= 1e-7 * abs(rhs), (lhs, rhs)


def test_coupled_gradient_central_differences():
    """dJ/dalpha and dJ/dphi of the coupled scan (mean solid T + 0.1 pressure drop over
    the last steps) against central differences on a few voxels."""
    shape = (8, 5, 4)
    t = _tor(shape, periodic=(False, False, False), ports=channel_ports(shape, 0.03), accel=(0.0, 0.0, 0.0))
    rng = np.random.default_rng(10)
    alpha = random_alpha(shape, 11)
    phi = rng.random(t.n)
    steps = 12
    ob = CoupledObjective(window=(7, steps), weight=1.0 / 6, terminal_weight=0.5, w_T=1.0, w_dp=0.1)
    J, ab, pb, _, _ = t.scan_adjoint(ob, alpha, phi, steps)
    assert abs(J - t.objective_value(ob, alpha, phi, steps)) <= 1e-14 * abs(J)
    h = 1e-6
    for x in np.argsort(-np.abs(pb))[:4]:
        p2, p3 = phi.copy(), phi.copy()
        p2[x] += h
        p3[x] -= h
        fd = (t.objective_value(ob, alpha, p2, steps) - t.objective_value(ob, alpha, p3, steps)) / (2 * h)
        assert abs(fd - pb[x]) <= 1e-6 * abs(pb[x]), (x, fd, pb[x])
    for x in np.argsort(-np.abs(ab))[:4]:
        a2, a3 = alpha.copy(), alpha.copy()
        a2[x] += h
        a3[x] -= h
        fd = (t.objective_value(ob, a2, phi, steps) - t.objective_value(ob, a3, phi, steps)) / (2 * h)
        assert abs(fd - ab[x]) <= 1e-6 * abs(ab[x]), (x, fd, ab[x])
