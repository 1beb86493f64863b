"""Pins of the Smagorinsky closure in the oracle (moving/psm.rs:161-197, 415-473).

* an equilibrium state has no non-equilibrium flux: the closure leaves omega unchanged;
* on a BGK cell without force the relaxation rate the collision applies equals Hou et
  al.'s closed form tau_e = (tau0 + sqrt(tau0^2 + 18 sqrt(2) C^2 |Pi| / rho)) / 2, with
  |Pi| the Frobenius norm of sum_i (f_i - eq_i) c_i c_i computed here with numpy;
* mass is conserved; the cell VJP matches central differences; the step VJP satisfies the
  dot-product identity.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as orc
from tests.lbm_cases import channel_ports, make_pair, random_alpha


def _oracle(kind, les=(0.1, 0.0), accel=(0.0, 0.0, 0.0), shape=(4, 4, 4), ports=(), periodic=(True, True, True)):
    _, prob = make_pair(shape, periodic, Q=19, collision=kind, omega=1.7, accel=accel, ports=ports, les=les)
    return orc.Oracle(prob)


def _eq(rho, u):
    c, w, _ = orc.lattice(19)
    cu = c @ u
    return rho * w * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * (u @ u))


def test_equilibrium_state_keeps_omega():
    o_les, o_plain = _oracle("trt"), _oracle("trt", les=(0.0, 0.0))
    f = _eq(1.02, np.array([0.03, -0.01, 0.02]))
    a = np.allclose(o_les.cell_collide(f, 0.0), o_plain.cell_collide(f, 0.0), rtol=0, atol=1e-16)
    assert a


@pytest.mark.parametrize("C,floor", [(0.1, 0.0), (0.17, 1e-4)])
def test_bgk_rate_matches_hou_closure(C, floor):
    o = _oracle("bgk", les=(C, floor))
    rng = np.random.default_rng(3)
    c, w, _ = orc.lattice(19)
    f = _eq(1.0, np.array([0.02, 0.01, -0.015])) + rng.standard_normal(19) * w * 0.02
    out = o.cell_collide(f, 0.0)
    rho = f.sum()
    u = (c.T @ f) / rho
    n = f - _eq(rho, u)
    pi = np.einsum("i,ia,ib->ab", n, c, c)
    P = np.sqrt((pi * pi).sum() + floor * floor) - floor
    tau0 = 1 / 1.7
    tau_e = (tau0 + np.sqrt(tau0 * tau0 + 18 * np.sqrt(2) * C * C * P / rho)) / 2
    k = np.argmax(np.abs(n))
    omega_applied = -(out[k] - f[k]) / n[k]
    assert abs(omega_applied * tau_e - 1) < 1e-12
    assert tau_e > tau0 * (1 + 1e-3)  # the closure adds eddy viscosity
    assert abs(out.sum() - f.sum()) < 1e-15


@pytest.mark.parametrize("kind", ["bgk", "trt"])
def test_cell_vjp_central_differences_with_les(kind):
    o = _oracle(kind, les=(0.15, 1e-5), accel=(1e-4, -2e-5, 3e-5))
    rng = np.random.default_rng(5)
    c, w, _ = orc.lattice(19)
    f = _eq(1.01, np.array([0.03, 0.01, -0.02])) + rng.standard_normal(19) * w * 0.03
    alpha = 0.05
    g = rng.standard_normal(19)
    fb, ab = o.cell_vjp(f, alpha, g)
    h = 1e-6
    for i in range(19):
        e = np.zeros(19)
        e[i] = h
        fd = (g @ o.cell_collide(f + e, alpha) - g @ o.cell_collide(f - e, alpha)) / (2 * h)
        assert abs(fd - fb[i]) <= 1e-7 * max(1.0, abs(fb[i])), (i, fd, fb[i])
    fd = (g @ o.cell_collide(f, alpha + h) - g @ o.cell_collide(f, alpha - h)) / (2 * h)
    assert abs(fd - ab) <= 1e-7 * max(1.0, abs(ab))


def test_step_vjp_dot_product_with_les():
    shape = (7, 5, 4)
    o = _oracle("trt", les=(0.12, 1e-6), shape=shape, ports=channel_ports(shape, 0.05), periodic=(False, False, False))
    rng = np.random.default_rng(7)
    alpha = random_alpha(shape, 8)
    f = o.run(alpha, 6)
    lam = rng.standard_normal(f.size)
    v = rng.standard_normal(f.size) * 1e-3
    h = 1e-6
    lhs = lam @ ((o.interval(alpha, f + h * v) - o.interval(alpha, f - h * v)) / (2 * h))
    fb, _ = o.interval_vjp(alpha, f, lam)
    assert abs(lhs - fb @ v) <= 1e-7 * abs(fb @ v)


# --- WALE (moving/wale.rs) ---------------------------------------------------------------
def _wale_oracle(shape=(4, 4, 4), periodic=(True, True, True), C=0.5, floor=1e-12, ports=(), solid=None):
    _, prob = make_pair(shape, periodic, Q=19, collision="trt", omega=1.7, ports=ports, solid=solid, wale=(C, floor))
    return orc.Oracle(prob)


def _wale_nicoud_ducros(g, C, floor, tau0):
    """Nicoud & Ducros (1999): nu = C^2 (Sd:Sd)^(3/2) / ((S:S)^(5/2) + (Sd:Sd)^(5/4)),
    Sd the traceless symmetric part of g^2, S the symmetric part of g; omega = 1/(tau0 + 3 nu)."""
    g2 = g @ g
    sd = (g2 + g2.T) / 2 - np.trace(g2) / 3 * np.eye(3)
    s = (g + g.T) / 2
    a, b = (sd * sd).sum(), (s * s).sum()
    nu = C * C * a ** 1.5 / (b ** 2.5 + a ** 1.25 + floor)
    return 1 / (tau0 + 3 * nu)


def test_wale_rate_closed_forms():
    o = _wale_oracle()
    tau0 = 1 / 1.7
    shear = np.zeros((3, 3))
    shear[0, 1] = 0.01  # u_x = gamma y: g^2 = 0, so WALE adds no viscosity in pure shear
    assert o.wale_omega(shear, tau0) == 1 / tau0
    rng = np.random.default_rng(1)
    for _ in range(5):
        g = rng.standard_normal((3, 3)) * 0.01
        assert abs(o.wale_omega(g, tau0) / _wale_nicoud_ducros(g, 0.5, 1e-12, tau0) - 1) < 1e-13


def test_wale_omega_vjp_central_differences():
    o = _wale_oracle()
    tau0 = 1 / 1.7
    g = np.random.default_rng(2).standard_normal((3, 3)) * 0.01
    gb = o.wale_omega_vjp(g, tau0, 0.7)
    h = 1e-7
    for a in range(3):
        for b in range(3):
            e = np.zeros((3, 3))
            e[a, b] = h
            fd = 0.7 * (o.wale_omega(g + e, tau0) - o.wale_omega(g - e, tau0)) / (2 * h)
            assert abs(fd - gb[a, b]) <= 1e-6 * np.abs(gb).max(), (a, b, fd, gb[a, b])


def test_wale_stencil_on_a_linear_field():
    """u = M x (equilibrium populations): central differences inside and one-sided ones at
    the closed faces and next to solid cells are exact, so a fluid cell gets omega(M) --
    with column b of M zeroed where neither neighbour along b is a fluid cell inside the
    lattice (Stencil::new's Difference::Zero)."""
    shape = (5, 4, 6)
    solid = np.zeros(shape, bool)
    solid[2, 1, 3] = True
    o = _wale_oracle(shape, periodic=(False, False, False), solid=solid)
    M = np.array([[0.002, 0.004, -0.001], [0.003, -0.001, 0.002], [0.001, 0.002, -0.001]])
    idx = np.stack(np.meshgrid(*[np.arange(s) for s in shape], indexing="ij"), -1).reshape(-1, 3)
    u = idx @ M.T
    c, w, _ = orc.lattice(19)
    cu = u @ c.T
    f = (w[None] * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * (u * u).sum(1, keepdims=True))).reshape(-1)
    rates = o.wale_rates(f)
    fluid = ~solid.reshape(-1)
    n_one_sided = 0
    for x in np.flatnonzero(fluid):
        p0 = idx[x]
        g = M.copy()
        for b in range(3):
            ok = []
            for d in (1, -1):
                q = p0.copy()
                q[b] += d
                ok.append(0 <= q[b] < shape[b] and not solid[tuple(q)])
            if not any(ok):
                g[:, b] = 0.0
            n_one_sided += ok.count(False) == 1
        assert abs(rates[x] / _wale_nicoud_ducros(g, 0.5, 1e-12, 1 / 1.7) - 1) < 1e-11, x
    assert n_one_sided > 50
    assert np.all(rates[~fluid] == 1.7)


def test_wale_step_vjp_dot_product():
    shape = (7, 5, 5)
    solid = np.zeros(shape, bool)
    solid[3, 2, 2] = True
    o = _wale_oracle(shape, periodic=(False, False, False), C=0.5, ports=channel_ports(shape, 0.05), solid=solid)
    rng = np.random.default_rng(3)
    alpha = random_alpha(shape, 4)
    f = o.run(alpha, 6)
    lam = rng.standard_normal(f.size)
    lam.reshape(-1, 19)[solid.reshape(-1)] = 0.0
    v = rng.standard_normal(f.size) * 1e-3
    v.reshape(-1, 19)[solid.reshape(-1)] = 0.0
    h = 1e-6
    lhs = lam @ ((o.interval(alpha, f + h * v) - o.interval(alpha, f - h * v)) / (2 * h))
    fb, _ = o.interval_vjp(alpha, f, lam)
    assert abs(lhs - fb @ v) <= 1e-7 * abs(fb @ v), (lhs, fb @ v)
