"""Pins of the oracle's stationary D3Q19 port rules (crates/implexity-physics-lbm/src/
ports.rs): the moment reconstruction `reconstruct` of "legacy_moment" ports (:343-406),
its right inverse (:325-341), the neighbour non-equilibrium pressure rule
reconstruct_neighbor (:484-497) and the cell Jacobian of apply_ports_vjp (:532-621).

The pins are properties the rules are built to have, not their formulas:
* the right inverse is a right inverse (rows . ri = I) of (1, c_t0, c_t1) over the
  unknown set;
* the moment rule returns density rho and momentum rho v0 (velocity ports: rho from the
  known populations, ports.rs:368-370; pressure ports: the normal velocity from the
  given density), leaves the known populations untouched bit for bit and maps the
  equilibrium of (rho, v0) to itself;
* the neighbour rule returns density rho_b and momentum rho_b (authored tangential
  velocity, neighbour normal velocity) whatever the neighbour's non-equilibrium part,
  and the neighbour's equilibrium to the boundary equilibrium;
* the Jacobian equals central differences, and for the (affine) moment rules is the
  same at every point;
* the step with these ports passes central differences of its VJP.

The regularized rules of "regularized_known_stress" ports (reconstruct_regularized,
:455-482, with the stress fit of known_stress_operator, :409-440) are pinned by
* the operator: numpy's SVD pseudo-inverse of the weighted known rows (the reference's
  route, with its rank-6 and condition-number checks) within 1e-14, and bit for bit the
  correctly rounded value of the exact rational operator, which is identified by the two
  identities that define a weighted least-squares left inverse (op B = I and
  B^T W^-1 (I - B op) = 0) holding exactly in rationals;
* a regularized state eq(rho, u) + basis . s with the port's own boundary values is a
  fixed point, and the rule is idempotent on any cell;
* the incoming (unknown) populations do not reach the result;
* the result has the density and momentum of the moment reconstruction (the stress
  basis carries neither);
* central differences of the Jacobian (taken at the source, not affine) and of the step
  VJP.
"""
from __future__ import annotations

from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as orc

C = np.array([
    (0, 0, 0), (-1, -1, 0), (-1, 0, -1), (-1, 0, 0), (-1, 0, 1), (-1, 1, 0), (0, -1, -1), (0, -1, 0), (0, -1, 1),
    (0, 0, -1), (0, 0, 1), (0, 1, -1), (0, 1, 0), (0, 1, 1), (1, -1, 0), (1, 0, -1), (1, 0, 0), (1, 0, 1),
    (1, 1, 0)], np.float64)
W = np.array([1 / 3 if not c.any() else (1 / 18 if np.abs(c).sum() == 1 else 1 / 36) for c in C])
O = orc.Oracle
BASE = orc.ST_PORT_BASE


def equilibrium(rho, u):
    """Second-order D3Q19 equilibrium, textbook form."""
    u = np.asarray(u, np.float64)
    cu = C @ u
    return rho * W * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * (u @ u))


def moments(f):
    return f.sum(), C.T @ f


def random_cell(rng, rho=1.02, u=(0.03, -0.02, 0.015), neq=2e-3):
    """Equilibrium plus a non-equilibrium part with zero mass and momentum."""
    f = equilibrium(rho, u)
    g = rng.standard_normal(19) * neq
    # project out mass and momentum: g -= W (a + 3 c.b) with a, b from the moments of g
    a = g.sum()
    b = C.T @ g
    g = g - W * (a + 3 * (C @ b))
    assert abs(g.sum()) < 1e-15 and np.abs(C.T @ g).max() < 1e-15
    return f + g


def face_axis_sign(face):
    return face // 2, 1 if face % 2 == 0 else -1


@pytest.mark.parametrize("face", range(6))
def test_right_inverse_is_a_right_inverse(face):
    axis, sign = face_axis_sign(face)
    unknown = [q for q in range(19) if C[q, axis] == sign]
    tang = [a for a in range(3) if a != axis]
    ri = O.stationary_right_inverse(face)
    assert ri.shape == (5, 3)
    rows = np.stack([np.ones(5), C[unknown, tang[0]], C[unknown, tang[1]]])
    assert np.array_equal(rows @ ri, np.eye(3))


@pytest.mark.parametrize("face", range(6))
@pytest.mark.parametrize("pressure", [False, True])
def test_moment_rule_moments_and_known_populations(face, pressure):
    rng = np.random.default_rng(face + 10 * pressure)
    axis, sign = face_axis_sign(face)
    f = random_cell(rng)
    v0 = np.array([0.021, -0.013, 0.008])
    tang = [a for a in range(3) if a != axis]
    if pressure:
        rho = 1.013
        kind, value = BASE + 12 + face, (rho, v0[tang[0]], v0[tang[1]])
    else:
        kind, value = BASE + 6 + face, tuple(v0)
    out = O.stationary_port(f, kind, value)
    unknown = C[:, axis] == sign
    assert np.array_equal(out[~unknown], f[~unknown])
    zero, outgoing = C[:, axis] == 0, C[:, axis] == -sign
    compat = f[zero].sum() + 2 * f[outgoing].sum()
    if pressure:
        v = v0.copy()
        v[axis] = (1 - compat / rho) * sign
    else:
        rho, v = compat / (1 - sign * v0[axis]), v0
    r, j = moments(out)
    assert abs(r - rho) < 2e-15
    assert np.abs(j - rho * v).max() < 2e-15


@pytest.mark.parametrize("face", range(6))
@pytest.mark.parametrize("pressure", [False, True])
def test_moment_rule_keeps_equilibrium(face, pressure):
    axis, _ = face_axis_sign(face)
    rho, v0 = 0.996, np.array([0.02, 0.011, -0.017])
    f = equilibrium(rho, v0)
    tang = [a for a in range(3) if a != axis]
    if pressure:
        kind, value = BASE + 12 + face, (rho, v0[tang[0]], v0[tang[1]])
    else:
        kind, value = BASE + 6 + face, tuple(v0)
    out = O.stationary_port(f, kind, value)
    assert np.abs(out - f).max() < 1e-16


@pytest.mark.parametrize("face", range(6))
def test_neighbor_rule_moments(face):
    rng = np.random.default_rng(40 + face)
    axis, _ = face_axis_sign(face)
    un = np.array([0.025, -0.018, 0.012])
    nb = random_cell(rng, rho=1.01, u=un)
    rho_b, tang_u = 0.995, (0.004, -0.006)
    out = O.stationary_port(nb, BASE + face, (rho_b,) + tang_u)
    rn, jn = moments(nb)
    target = np.empty(3)
    t = iter(tang_u)
    for d in range(3):
        target[d] = jn[d] / rn if d == axis else next(t)
    r, j = moments(out)
    assert abs(r - rho_b) < 2e-15
    assert np.abs(j - rho_b * target).max() < 2e-15
    # the neighbour's equilibrium maps to the boundary equilibrium
    eq_n = equilibrium(rn, jn / rn)
    out_eq = O.stationary_port(eq_n, BASE + face, (rho_b,) + tang_u)
    assert np.abs(out_eq - equilibrium(rho_b, target)).max() < 1e-16


def test_neighbor_rule_differs_from_moving_pressure_rule():
    """The moving solver's pressure rule takes the neighbour's whole velocity; the
    stationary one keeps the authored tangential velocity."""
    rng = np.random.default_rng(3)
    nb = random_cell(rng, u=(0.02, 0.03, -0.01))
    o = orc.Oracle(orc.Problem(shape=(4, 4, 4), periodic=(False, True, True), Q=19))
    moving = o.port_populations(nb, 1, (1.0, 0.0, 0.0))
    stationary = O.stationary_port(nb, BASE + 0, (1.0, 0.0, 0.0))
    _, jm = moments(moving)
    _, js = moments(stationary)
    assert abs(js[1]) < 1e-16 and abs(jm[1]) > 1e-3


@pytest.mark.parametrize("kind_off,face", [(0, 1), (0, 2), (6, 0), (6, 5), (12, 3), (12, 4), (18, 1), (18, 2),
                                            (24, 0), (24, 5)])
def test_jacobian_matches_central_differences(kind_off, face):
    rng = np.random.default_rng(kind_off + face)
    f = random_cell(rng)
    kind = BASE + kind_off + face
    value = (0.01, -0.02, 0.015) if kind_off in (6, 18) else (1.004, 0.003, -0.002)
    jac = O.stationary_port_jacobian(f, kind, value)
    h = 1e-6
    fd = np.empty((19, 19))
    for k in range(19):
        e = np.zeros(19)
        e[k] = h
        fd[:, k] = (O.stationary_port(f + e, kind, value) - O.stationary_port(f - e, kind, value)) / (2 * h)
    assert np.abs(jac - fd).max() < 1e-8
    if kind_off in (6, 12):  # affine: the same Jacobian everywhere
        assert np.array_equal(jac, O.stationary_port_jacobian(f * 1.3 + 0.01, kind, value))
    else:
        assert not np.array_equal(jac, O.stationary_port_jacobian(f * 1.3 + 0.01, kind, value))


def _problem(ports, accel=(0.0, 0.0, 0.0), shape=(6, 5, 4)):
    return orc.Problem(shape=shape, periodic=(False, False, True), Q=19, kind="bgk_guo", omega=1.25,
                       accel=accel, ports=ports)


def test_step_vjp_central_differences_with_stationary_ports():
    shape = (6, 5, 4)
    mask = np.zeros(shape, bool)
    mask[0, 1:-1, :] = True  # the face without its edges on the closed y walls
    ports = [orc.Port("xmin", "velocity_moment", (0.02, 0.0, 0.0), mask=mask),
             orc.Port("xmax", "pressure_neighbor", (0.998, 0.0, 0.0))]
    o = orc.Oracle(_problem(ports))
    rng = np.random.default_rng(7)
    alpha = rng.uniform(0, 0.05, o.n)
    f = o.run(alpha, 3)
    lam = rng.standard_normal(o.n * 19)
    v = rng.standard_normal(o.n * 19) * 1e-3
    fb, ab = o.interval_vjp(alpha, f, lam)
    h = 1e-6
    # the whole field, then a direction on the port cells and their neighbours only
    near = np.zeros(shape, bool)
    near[:2], near[-2:] = True, True
    for vv in (v, v * np.repeat(near.reshape(-1), 19)):
        fd = (o.interval(alpha, f + h * vv) - o.interval(alpha, f - h * vv)) / (2 * h)
        assert abs(lam @ fd - fb @ vv) <= 1e-8 * np.abs(lam).sum() * np.abs(fd).max()


def test_moment_port_cells_and_values():
    """Row building: masked cells in ascending order, the cell as its own source, v0 =
    u - a/2 in the values."""
    shape = (6, 5, 4)
    mask = np.zeros(shape, bool)
    mask[5, 2, 1:3] = True
    ports = [orc.Port("xmax", "pressure_moment", (1.01, 0.004, -0.003), mask=mask)]
    o = orc.Oracle(_problem(ports, accel=(1e-4, 2e-4, -6e-4)))
    cells, nbs, kinds, values = o.port_cells()
    expect = [o.index(5, 2, 1), o.index(5, 2, 2)]
    assert cells == expect and nbs == expect
    assert kinds == [BASE + 12 + 1] * 2
    assert values[0] == [1.01, 0.004 - 0.5 * 2e-4, -0.003 + 0.5 * 6e-4]


# --- regularized_known_stress ---------------------------------------------------------

def stress_basis():
    """(19, 6): 4.5 w (Q00, Q11, Q22, 2 Q01, 2 Q02, 2 Q12), Q = c c - I/3."""
    out = np.empty((19, 6))
    for q in range(19):
        Qm = np.outer(C[q], C[q]) - np.eye(3) / 3
        out[q] = 4.5 * W[q] * np.array([Qm[0, 0], Qm[1, 1], Qm[2, 2], 2 * Qm[0, 1], 2 * Qm[0, 2], 2 * Qm[1, 2]])
    return out


def frac_basis():
    w = [Fraction(1, 3) if not C[q].any() else (Fraction(1, 18) if np.abs(C[q]).sum() == 1 else Fraction(1, 36))
         for q in range(19)]
    out = []
    for q in range(19):
        c = [int(v) for v in C[q]]
        Qm = [[Fraction(c[a] * c[b]) - (Fraction(1, 3) if a == b else 0) for b in range(3)] for a in range(3)]
        out.append([Fraction(9, 2) * w[q] * e for e in (Qm[0][0], Qm[1][1], Qm[2][2], 2 * Qm[0][1], 2 * Qm[0][2],
                                                         2 * Qm[1][2])])
    return out, w


@pytest.mark.parametrize("face", range(6))
def test_stress_operator_matches_the_svd_pseudo_inverse(face):
    axis, sign = face_axis_sign(face)
    known = [q for q in range(19) if C[q, axis] != sign]
    got_known, op = O.stationary_stress_operator(face)
    assert list(got_known) == known
    B = stress_basis()
    assert np.array_equal(O.stationary_stress_basis(), B)
    weighted = B[known] / np.sqrt(W[known])[:, None]
    assert np.linalg.matrix_rank(weighted) == 6
    assert np.linalg.cond(weighted) <= 1e12
    ref = np.linalg.pinv(weighted) / np.sqrt(W[known])[None, :]
    assert np.abs(op - ref).max() <= 1e-14 * np.abs(ref).max()


@pytest.mark.parametrize("face", range(6))
def test_stress_operator_is_the_correctly_rounded_exact_operator(face):
    """The exact operator has small denominators; recover it from the doubles and check
    that it is the weighted least-squares left inverse exactly, then that every double is
    its nearest (float(Fraction) rounds correctly)."""
    known, op = O.stationary_stress_operator(face)
    fb, fw = frac_basis()
    B = [fb[q] for q in known]
    winv = [1 / fw[q] for q in known]
    opf = [[Fraction(float(v)).limit_denominator(100000) for v in row] for row in op]
    for r in range(6):
        for s_ in range(6):  # op B = I
            assert sum(opf[r][k] * B[k][s_] for k in range(14)) == (1 if r == s_ else 0)
    # B^T W^-1 (I - B op) = 0
    resid = [[(1 if i == j else 0) - sum(B[i][s_] * opf[s_][j] for s_ in range(6)) for j in range(14)]
             for i in range(14)]
    for a in range(6):
        for j in range(14):
            assert sum(B[i][a] * winv[i] * resid[i][j] for i in range(14)) == 0
    assert all(float(opf[r][k]) == op[r, k] for r in range(6) for k in range(14))


def _regularized(face, pressure, v):
    axis, _ = face_axis_sign(face)
    tang = [a for a in range(3) if a != axis]
    if pressure:
        return BASE + 24 + face, lambda rho: (rho, v[tang[0]], v[tang[1]])
    return BASE + 18 + face, lambda rho: tuple(v)


@pytest.mark.parametrize("face", range(6))
@pytest.mark.parametrize("pressure", [False, True])
def test_regularized_fixed_point_and_idempotence(face, pressure):
    rng = np.random.default_rng(60 + face + 6 * pressure)
    rho, u = 1.012, np.array([0.023, -0.017, 0.011])
    s = rng.standard_normal(6) * 1e-3
    f = equilibrium(rho, u) + stress_basis() @ s
    kind, value = _regularized(face, pressure, u)
    out = O.stationary_port(f, kind, value(rho))
    assert np.abs(out - f).max() < 2e-16
    g = random_cell(rng)
    once = O.stationary_port(g, kind, value(0.997))
    twice = O.stationary_port(once, kind, value(0.997))
    assert np.abs(twice - once).max() < 2e-16
    assert np.abs(once - g).max() > 1e-4


@pytest.mark.parametrize("face", range(6))
@pytest.mark.parametrize("pressure", [False, True])
def test_regularized_ignores_incoming_populations(face, pressure):
    rng = np.random.default_rng(80 + face + 6 * pressure)
    axis, sign = face_axis_sign(face)
    f = random_cell(rng)
    kind, value = _regularized(face, pressure, np.array([0.02, 0.01, -0.015]))
    out = O.stationary_port(f, kind, value(1.006))
    g = f.copy()
    unknown = C[:, axis] == sign
    g[unknown] += rng.standard_normal(unknown.sum()) * 1e-2
    assert np.array_equal(O.stationary_port(g, kind, value(1.006)), out)
    g = f.copy()
    g[~unknown] += rng.standard_normal((~unknown).sum()) * 1e-6
    assert not np.array_equal(O.stationary_port(g, kind, value(1.006)), out)


@pytest.mark.parametrize("face", range(6))
@pytest.mark.parametrize("pressure", [False, True])
def test_regularized_moments_are_the_moment_rules(face, pressure):
    rng = np.random.default_rng(100 + face + 6 * pressure)
    f = random_cell(rng, neq=5e-3)
    v = np.array([0.015, -0.02, 0.01])
    kind, value = _regularized(face, pressure, v)
    out = O.stationary_port(f, kind, value(1.004))
    target = O.stationary_port(f, kind - 12, value(1.004))
    r, j = moments(out)
    rt, jt = moments(target)
    assert abs(r - rt) < 2e-15 and np.abs(j - jt).max() < 2e-15
    assert np.abs(out - target).max() > 1e-4  # the stress is refitted, not copied


def test_regularized_port_rows_and_acceleration():
    shape = (6, 5, 4)
    mask = np.zeros(shape, bool)
    mask[0, 1:4, 1:3] = True
    ports = [orc.Port("xmin", "velocity_regularized", (0.02, 0.001, 0.0), mask=mask),
             orc.Port("xmax", "pressure_regularized", (0.998, 0.003, 0.0))]
    o = orc.Oracle(_problem(ports))
    cells, nbs, kinds, values = o.port_cells()
    assert cells == nbs
    assert kinds[:6] == [BASE + 18] * 6 and kinds[6:] == [BASE + 24 + 1] * (len(kinds) - 6)
    assert values[0] == [0.02, 0.001, 0.0] and values[6] == [0.998, 0.003, 0.0]
    with pytest.raises(ValueError, match="zero acceleration"):
        orc.Oracle(_problem(ports, accel=(0.0, 1e-5, 0.0))).port_cells()


def test_step_vjp_central_differences_with_regularized_ports():
    shape = (6, 5, 4)
    mask = np.zeros(shape, bool)
    mask[0, 1:-1, :] = True
    ports = [orc.Port("xmin", "velocity_regularized", (0.02, 0.0, 0.004), mask=mask),
             orc.Port("xmax", "pressure_regularized", (0.998, 0.002, 0.0))]
    o = orc.Oracle(_problem(ports))
    rng = np.random.default_rng(17)
    alpha = rng.uniform(0, 0.05, o.n)
    f = o.run(alpha, 3)
    lam = rng.standard_normal(o.n * 19)
    v = rng.standard_normal(o.n * 19) * 1e-3
    fb, _ = o.interval_vjp(alpha, f, lam)
    h = 1e-6
    near = np.zeros(shape, bool)
    near[:2], near[-2:] = True, True
    for vv in (v, v * np.repeat(near.reshape(-1), 19)):
        fd = (o.interval(alpha, f + h * vv) - o.interval(alpha, f - h * vv)) / (2 * h)
        assert abs(lam @ fd - fb @ vv) <= 1e-8 * np.abs(lam).sum() * np.abs(fd).max()
