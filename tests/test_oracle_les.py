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
