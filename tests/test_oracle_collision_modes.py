"""Pins of the oracle's moment-space and cumulant collisions away from their default
rates (run without a GPU).

The MRT and regularized operators relax y = f - f_eq + s/2 group by group
(moving/collision.rs:349-394): in the W-weighted space of lattice functions, the
relaxation is a spectral operator whose eigenspaces are the moment groups.  The tests
build those eigenspaces independently of the oracle's Gram-Schmidt basis -- from
parity (odd and even lattice functions) and W-orthogonality to the conserved and
second-order polynomials -- excite one mode at a time on top of the equilibrium of a
forced Brinkman cell, and check that the collision removes exactly rate * y of it, with
the rate of the mode's group.  A swapped group assignment, a wrong rate or a
non-orthogonal basis fails.

The cumulant operator (collision.rs:636-752) is checked through the 27 central moments
of its output about the collision velocity u, computed here by direct summation: mass,
first moments F/2, second moments relaxed with omega (deviatoric) and bulk_rate (trace),
and every higher moment rho times the Wick (Gaussian) closure of sigma = M2/rho.  At
omega = 1, bulk = 1 and no force the output is the D3Q27 product equilibrium
rho prod_a w(c_a; u_a).

The dissipation objective is pinned against the porous term of the reference-restated
`vector` (solver.rs:847-848,889-890) and a closed form.
"""
from __future__ import annotations

import itertools

import numpy as np
import pytest
from scipy.linalg import null_space

from oracle import oracle as orc

RHO, U = 1.03, np.array([0.03, -0.02, 0.01])
ACCEL, ALPHA = np.array([2e-4, -1e-4, 3e-4]), 0.04


def _cell(Q, kind, **kw):
    return orc.Oracle(orc.Problem(shape=(3, 3, 3), periodic=(True, True, True), Q=Q, kind=kind,
                                  accel=tuple(ACCEL), **kw))


def _eq_src(Q, rho, u, F):
    c, w, _ = orc.lattice(Q)
    cu, cF = c @ u, c @ F
    eq = rho * w * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * u @ u)
    src = w * (3 * cF + 9 * cu * cF - 3 * (u @ F))
    return eq, src


def _group_spaces(Q):
    """Orthonormal (in the W inner product) bases of the moment groups' subspaces of lattice
    functions p, built from parity and W-orthogonality alone."""
    c, w, opp = orc.lattice(Q)
    W = np.diag(w)
    cx, cy, cz = c.T.astype(float)
    one = np.ones(Q)
    cc = cx ** 2 + cy ** 2 + cz ** 2
    bulk = [cc - 1.0]
    shear = [2 * cx ** 2 - cy ** 2 - cz ** 2, cy ** 2 - cz ** 2, cx * cy, cx * cz, cy * cz]
    eye = np.eye(Q)
    odd_funcs = np.array([eye[i] - eye[opp[i]] for i in range(Q) if i < opp[i]]).T
    even_funcs = np.array([eye[0]] + [eye[i] + eye[opp[i]] for i in range(Q) if i < opp[i]]).T

    def complement(funcs, against):
        A = np.array(against).T  # columns: functions to be W-orthogonal to
        coef = null_space(A.T @ W @ funcs)
        return funcs @ coef

    odd = complement(odd_funcs, [cx, cy, cz])
    even_high = complement(even_funcs, [one, cc] + shear)

    def orthonormal(vs):
        V = np.array(vs).T if isinstance(vs, list) else vs
        L = np.linalg.cholesky(V.T @ W @ V)
        return V @ np.linalg.inv(L).T

    return {"bulk": orthonormal(bulk), "shear": orthonormal(shear), "odd": orthonormal(odd),
            "even_high": orthonormal(even_high)}


@pytest.mark.parametrize("Q,dims", [(19, (1, 5, 6, 3)), (27, (1, 5, 10, 7))])
def test_group_space_dimensions(Q, dims):
    sp = _group_spaces(Q)
    assert tuple(sp[g].shape[1] for g in ("bulk", "shear", "odd", "even_high")) == dims
    assert 4 + sum(dims) == Q  # with mass and momentum, the groups span every lattice function


def _mode_response(o, Q, p, eps=1e-3):
    """Excites y = eps w p on the equilibrium of the forced Brinkman cell and returns
    (y, f + src - collide(f)), which is K y for the relaxation operator K."""
    _, w, _ = orc.lattice(Q)
    F = RHO * (ACCEL - ALPHA * U)
    eq, src = _eq_src(Q, RHO, U, F)
    y = eps * w * p
    f = eq - 0.5 * src + y
    return y, f + src - o.cell_collide(f, ALPHA)


@pytest.mark.parametrize("Q", [19, 27])
@pytest.mark.parametrize("bulk,even", [(1.1, 0.7), (0.6, 1.45)])
def test_mrt_relaxes_each_group_at_its_rate(Q, bulk, even):
    omega, magic = 1.6, 0.21
    rates = {"bulk": bulk, "shear": omega, "odd": orc.odd_rate(omega, magic), "even_high": even}
    o = _cell(Q, "mrt", omega=omega, magic=magic, bulk_rate=bulk, even_rate=even)
    rng = np.random.default_rng(11)
    for group, V in _group_spaces(Q).items():
        for _ in range(3):
            p = V @ rng.standard_normal(V.shape[1])
            y, Ky = _mode_response(o, Q, p)
            np.testing.assert_allclose(Ky, rates[group] * y, rtol=0, atol=1e-15,
                                       err_msg=group)


@pytest.mark.parametrize("Q", [19, 27])
def test_regularized_keeps_only_bulk_and_shear(Q):
    """Regularized at omega != 1: bulk and shear relax with omega, every other mode with 1,
    so the non-hydrodynamic part of f' - f_eq - s/2 vanishes (collision.rs:376-394)."""
    omega = 1.6
    o = _cell(Q, "regularized", omega=omega)
    rng = np.random.default_rng(12)
    for group, V in _group_spaces(Q).items():
        rate = omega if group in ("bulk", "shear") else 1.0
        p = V @ rng.standard_normal(V.shape[1])
        y, Ky = _mode_response(o, Q, p)
        np.testing.assert_allclose(Ky, rate * y, rtol=0, atol=1e-15, err_msg=group)
    # a mixed input: the collision output, less equilibrium and half the source, keeps
    # only the bulk and shear parts, scaled by 1 - omega
    _, w, _ = orc.lattice(Q)
    sp = _group_spaces(Q)
    parts = {g: sp[g] @ rng.standard_normal(sp[g].shape[1]) for g in sp}
    p = sum(parts.values())
    F = RHO * (ACCEL - ALPHA * U)
    eq, src = _eq_src(Q, RHO, U, F)
    f = eq - 0.5 * src + 1e-3 * w * p
    left = o.cell_collide(f, ALPHA) - eq - 0.5 * src
    want = (1 - omega) * 1e-3 * w * (parts["bulk"] + parts["shear"])
    np.testing.assert_allclose(left, want, rtol=0, atol=5e-16)


def _wick(idx, sigma):
    """Sum over the perfect matchings of the index list of the product of sigma entries."""
    if not idx:
        return 1.0
    if len(idx) % 2:
        return 0.0
    a, rest = idx[0], idx[1:]
    return sum(sigma[a][rest[k]] * _wick(rest[:k] + rest[k + 1:], sigma) for k in range(len(rest)))


def _central_moments(f, c, u):
    d = c - u
    return {e: float(np.sum(f * d[:, 0] ** e[0] * d[:, 1] ** e[1] * d[:, 2] ** e[2]))
            for e in itertools.product(range(3), repeat=3)}


@pytest.mark.parametrize("omega,bulk,accel", [(1.6, 1.0, True), (1.6, 0.8, True), (0.9, 1.3, False),
                                              (1.0, 1.0, True)])
def test_cumulant_central_moments(omega, bulk, accel):
    c, w, _ = orc.lattice(27)
    a = ACCEL if accel else np.zeros(3)
    o = orc.Oracle(orc.Problem(shape=(3, 3, 3), periodic=(True, True, True), Q=27, kind="cumulant", omega=omega,
                               bulk_rate=bulk, accel=tuple(a)))
    rng = np.random.default_rng(13)
    for _ in range(3):
        f = w * (1 + 0.15 * rng.standard_normal(27))
        rho = f.sum()
        u = (c.T @ f / rho + a / 2) / (1 + ALPHA / 2)
        F = rho * (a - ALPHA * u)
        out = o.cell_collide(f, ALPHA)
        k_in, k_out = _central_moments(f, c, u), _central_moments(out, c, u)
        # zeroth and first order
        assert abs(k_out[(0, 0, 0)] - rho) < 1e-15
        for ax, e in enumerate([(1, 0, 0), (0, 1, 0), (0, 0, 1)]):
            assert abs(k_out[e] - F[ax] / 2) < 1e-16
        # second order: deviatoric parts scaled by 1 - omega, trace relaxed toward rho
        xx, yy, zz = (2, 0, 0), (0, 2, 0), (0, 0, 2)
        tr_in = k_in[xx] + k_in[yy] + k_in[zz]
        tr_out = k_out[xx] + k_out[yy] + k_out[zz]
        assert abs(tr_out - (tr_in + (rho - tr_in) * bulk)) < 1e-15
        for p, q in [(xx, yy), (xx, zz)]:
            assert abs((k_out[p] - k_out[q]) - (1 - omega) * (k_in[p] - k_in[q])) < 1e-15
        for e in [(1, 1, 0), (1, 0, 1), (0, 1, 1)]:
            assert abs(k_out[e] - (1 - omega) * k_in[e]) < 1e-16
        # third order and above: rho times the Wick closure of sigma = M2 / rho
        sigma = np.array([[k_out[xx], k_out[(1, 1, 0)], k_out[(1, 0, 1)]],
                          [k_out[(1, 1, 0)], k_out[yy], k_out[(0, 1, 1)]],
                          [k_out[(1, 0, 1)], k_out[(0, 1, 1)], k_out[zz]]]) / rho
        for e, v in k_out.items():
            if sum(e) >= 3:
                idx = [0] * e[0] + [1] * e[1] + [2] * e[2]
                assert abs(v - rho * _wick(idx, sigma)) < 2e-16, (e, v, rho * _wick(idx, sigma))


def test_cumulant_unit_rates_are_the_product_equilibrium():
    """omega = 1, bulk = 1, no force: the output is rho prod_a w(c_a; u_a) with the 1-D
    weights w(0) = 2/3 - u^2, w(+-1) = (1/3 + u^2 +- u)/2, whatever the input."""
    c, w, _ = orc.lattice(27)
    o = orc.Oracle(orc.Problem(shape=(3, 3, 3), periodic=(True, True, True), Q=27, kind="cumulant", omega=1.0,
                               bulk_rate=1.0))
    rng = np.random.default_rng(14)
    for _ in range(4):
        f = w * (1 + 0.2 * rng.standard_normal(27))
        rho = f.sum()
        u = c.T @ f / rho

        def w1(ci, ua):
            return 2.0 / 3.0 - ua * ua if ci == 0 else (1.0 / 3.0 + ua * ua + ci * ua) / 2.0

        want = np.array([rho * w1(ci[0], u[0]) * w1(ci[1], u[1]) * w1(ci[2], u[2]) for ci in c])
        np.testing.assert_allclose(o.cell_collide(f, 0.0), want, rtol=0, atol=4e-16)  # 2-4 ulp of w_0


def test_dissipation_is_vectors_porous_term():
    """The dissipation objective sum rho alpha |u|^2 over fluid cells equals the porous
    term of vector (solver.rs:889-890) at unit scales, and the closed form N rho alpha |u|^2
    on a uniform equilibrium."""
    shape = (4, 3, 5)
    solid = np.zeros(shape, bool)
    solid[2, 1, 3] = True
    o = orc.Oracle(orc.Problem(shape=shape, periodic=(True, True, True), Q=19, kind="trt", omega=1.2, solid=solid,
                               accel=tuple(ACCEL)))
    rng = np.random.default_rng(15)
    alpha = 0.05 * rng.random(o.n)
    state = o.run(alpha, 3, f0=np.tile(_eq_src(19, 1.0, U, np.zeros(3))[0], o.n))
    v = o.vector(None, alpha, state, [1.0, 1.0, 1.0, 1.0])
    assert abs(o.dissipation(alpha, state) - v[4]) < 1e-15 * abs(v[4])
    # closed form: uniform equilibrium, constant alpha, no force: u = (j/rho)/(1 + alpha/2)
    o2 = orc.Oracle(orc.Problem(shape=shape, periodic=(True, True, True), Q=19, kind="trt", omega=1.2))
    a0 = 0.07
    f0 = np.tile(_eq_src(19, RHO, U, np.zeros(3))[0], o2.n)
    ub = U / (1 + a0 / 2)
    assert abs(o2.dissipation(np.full(o2.n, a0), f0) - o2.n * RHO * a0 * (ub @ ub)) < 1e-15
