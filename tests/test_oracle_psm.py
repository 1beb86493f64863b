"""Oracle pins of the occupancy coupling (moving/psm.rs: saturate / beta :96-152, the
Psm and PsmSuperposition branches of cell_update :228-266 and of cell_update_vjp
:331-510) on the stationary step."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as orc
from tests.lbm_cases import random_alpha

C19, W19, _ = orc.lattice(19)


C_ALPHA, DRAG_SHAPE = 0.08, 0.1


def _oracle(law, kind="trt", kappa=0.1, shape=(4, 4, 4), accel=(2e-4, -1e-4, 0.0), **kw):
    return orc.Oracle(orc.Problem(shape=shape, periodic=(True, True, True), kind=kind, omega=1.3, accel=accel,
                                  coupling=law, psm_kappa=kappa, psm_c_alpha=C_ALPHA, psm_drag_shape=DRAG_SHAPE,
                                  **kw))


def _state(seed, rho=1.01, amp=2e-3):
    return W19 * rho + amp * np.random.default_rng(seed).standard_normal(19) * W19


def test_beta_is_the_noble_torczynski_weight():
    """Below the saturation band B = d t / ((1 - d) + t) with t = tau - 1/2 (Noble &
    Torczynski 1998); the returned partials of B/d match central differences in both
    branches."""
    kappa = 0.1
    for d in (0.05, 0.4, 0.85, 0.93, 0.97, 1.05, 1.3):
        for t in (0.05, 0.3, 1.2):
            bd, bdd, bdt = orc.psm_beta(d, t, kappa)
            if d <= 1 - kappa:
                assert abs(bd * d - d * t / ((1 - d) + t)) <= 1e-15
            h = 1e-6
            fd_d = (orc.psm_beta(d + h, t, kappa)[0] - orc.psm_beta(d - h, t, kappa)[0]) / (2 * h)
            fd_t = (orc.psm_beta(d, t + h, kappa)[0] - orc.psm_beta(d, t - h, kappa)[0]) / (2 * h)
            assert abs(fd_d - bdd) <= 1e-7 * max(1.0, abs(bdd)), (d, t)
            assert abs(fd_t - bdt) <= 1e-7 * max(1.0, abs(bdt)), (d, t)
    # saturated: B = 1
    assert abs(orc.psm_beta(1.25, 0.3, kappa)[0] * 1.25 - 1.0) <= 1e-15


def test_saturation_joins_smoothly():
    kappa = 0.1
    lo, hi = 1 - kappa, 1 + kappa
    assert orc.psm_saturate(lo, kappa) == lo and orc.psm_saturate(hi, kappa) == 1.0
    xs = np.linspace(lo - 0.05, hi + 0.05, 401)
    v = np.array([orc.psm_saturate(x, kappa) for x in xs])
    assert np.all(np.diff(v) >= 0) and v.max() <= 1.0
    h = 1e-7
    for x in (lo + 1e-9, hi - 1e-9, 0.95, 1.02):
        fd = (orc.psm_saturate(x + h, kappa) - orc.psm_saturate(x - h, kappa)) / (2 * h)
        left = (orc.psm_saturate(x, kappa) - orc.psm_saturate(x - h, kappa)) / h
        right = (orc.psm_saturate(x + h, kappa) - orc.psm_saturate(x, kappa)) / h
        assert abs(left - right) <= 1e-5 and abs(fd - (left + right) / 2) <= 1e-5
    # C1 at the joins: slope 1 at lo, 0 at hi
    assert abs((orc.psm_saturate(lo + h, kappa) - lo) / h - 1.0) <= 1e-5
    assert abs((1.0 - orc.psm_saturate(hi - h, kappa)) / h) <= 1e-5


@pytest.mark.parametrize("law", ["psm", "psm_superposition"])
@pytest.mark.parametrize("kind", ["trt", "bgk"])
@pytest.mark.parametrize("d", [0.3, 0.95, 1.4])
def test_mass_and_momentum_exchange(law, kind, d):
    """Both laws keep the cell's mass; the exchange is dp = B (rho u_s - j + rho a/2)
    (Psm) or rho B (u_s - u) (PsmSuperposition) with u_s = m / d and u = j/rho + a/2, and
    equals the change of momentum it causes."""
    o = _oracle(law, kind)
    a = np.array(o.prob.accel)
    f = _state(1)
    us = np.array([0.012, -0.005, 0.003])
    m = d * us
    post, dp = o.psm_cell(f, d, m)
    fluid = o.cell_collide(f, 0.0)
    rho, j = f.sum(), f @ C19
    u = j / rho + a / 2
    t = 1 / 1.3 - 0.5
    B = orc.psm_beta(d, t, o.prob.psm_kappa)[0] * d
    assert abs(post.sum() - rho) <= 1e-15
    if law == "psm":
        want = B * (rho * us - j + rho * a / 2)
        change = post @ C19 - (f @ C19 + (1 - B) * (fluid - f) @ C19)
    else:
        want = rho * B * (us - u)
        change = post @ C19 - fluid @ C19
    assert np.abs(dp - want).max() <= 1e-16
    assert np.abs(dp - change).max() <= 1e-16
    if d >= 1.1:  # full solid: the cell moves with the solid
        assert np.abs(post @ C19 - rho * (us + a / 2)).max() <= 1e-16


@pytest.mark.parametrize("law", ["psm", "psm_superposition"])
def test_vanishing_occupancy_is_the_fluid_collision(law):
    o = _oracle(law)
    f = _state(2)
    fluid = o.cell_collide(f, 0.0)
    for d in (1e-6, 1e-9):
        post, _ = o.psm_cell(f, d, d * np.array([0.01, 0.0, 0.0]))
        assert np.abs(post - fluid).max() <= 10 * d * np.abs(f).max()


def test_resistance_closed_form_and_partials():
    """alpha = c e / ((q + 1) - e) with e the saturated occupancy, gamma = alpha / d
    (c / ((q + 1) - d) below the band); the partials against central differences."""
    kappa, c, q = 0.1, C_ALPHA, DRAG_SHAPE
    for d in (0.05, 0.5, 0.93, 0.97, 1.05, 1.4):
        a, g, part = orc.psm_resistance(d, c, q, kappa)
        e = orc.psm_saturate(d, kappa)
        assert abs(a - c * e / ((q + 1) - e)) <= 1e-16
        assert abs(g - a / d) <= 1e-15 * max(1.0, g)
        h = 1e-6
        ap, gp, _ = orc.psm_resistance(d + h, c, q, kappa)
        am, gm, _ = orc.psm_resistance(d - h, c, q, kappa)
        assert abs((ap - am) / (2 * h) - part[0]) <= 1e-7 and abs((gp - gm) / (2 * h) - part[2]) <= 1e-7
        ap, gp, _ = orc.psm_resistance(d, c + h, q, kappa)
        am, gm, _ = orc.psm_resistance(d, c - h, q, kappa)
        assert abs((ap - am) / (2 * h) - part[1]) <= 1e-7 and abs((gp - gm) / (2 * h) - part[3]) <= 1e-7


@pytest.mark.parametrize("kind", ["trt", "bgk"])
@pytest.mark.parametrize("d", [0.3, 0.95, 1.4])
def test_brinkman_law_at_rest_is_the_stationary_operator(kind, d):
    """With the solid at rest the Brinkman law is the stationary Brinkman collision with
    alpha = resistance(d), bit for bit; a moving solid adds rho gamma m to the force, and
    the exchange dp = rho z is the change of momentum beyond j + rho a."""
    o = _oracle("brinkman_occupancy", kind)
    a = np.array(o.prob.accel)
    f = _state(5)
    alpha, gamma, _ = orc.psm_resistance(d, C_ALPHA, DRAG_SHAPE, 0.1)
    post, dp = o.psm_cell(f, d, np.zeros(3))
    assert np.array_equal(post, o.cell_collide(f, alpha))
    m = d * np.array([0.01, -0.02, 0.005])
    post, dp = o.psm_cell(f, d, m)
    rho, j = f.sum(), f @ C19
    assert abs(post.sum() - rho) <= 1e-15
    assert np.abs(post @ C19 - (j + rho * a + dp)).max() <= 1e-16
    u = (j / rho + a / 2 + gamma * m / 2) / (1 + alpha / 2)
    assert np.abs(dp - rho * (gamma * m - alpha * u)).max() <= 1e-16


def test_brinkman_law_strong_drag_moves_with_the_solid():
    """As the drag grows the cell's forced velocity u = (j + F/2)/rho tends to the solid's,
    m / d.  For BGK at omega = 1 the post-collision momentum is rho u + F/2 = 2 rho u - j,
    so u is read back as (post momentum + j) / (2 rho)."""
    d = 0.6
    us = np.array([0.02, -0.01, 0.004])
    f = _state(6)
    rho = f.sum()
    o = orc.Oracle(orc.Problem(shape=(4, 4, 4), periodic=(True, True, True), kind="bgk", omega=1.0,
                               accel=(0.0, 0.0, 0.0), coupling="brinkman_occupancy", psm_kappa=0.1,
                               psm_c_alpha=1e9, psm_drag_shape=DRAG_SHAPE))
    post, _ = o.psm_cell(f, d, d * us)
    u = (post @ C19 + f @ C19) / (2 * rho)
    assert np.abs(u - us).max() <= 1e-8


@pytest.mark.parametrize("law", ["psm", "psm_superposition", "brinkman_occupancy"])
@pytest.mark.parametrize("kind", ["trt", "bgk"])
@pytest.mark.parametrize("d", [0.4, 0.95, 1.3])
def test_cell_vjp_matches_central_differences(law, kind, d):
    o = _oracle(law, kind)
    rng = np.random.default_rng(3)
    f = _state(3)
    m = d * np.array([0.01, -0.004, 0.006])
    g = rng.standard_normal(19)
    dpb = rng.standard_normal(3)
    fb, db, mb = o.psm_cell_vjp(f, d, m, g, dpb)

    def phi(ff, dd, mm):
        post, dp = o.psm_cell(ff, dd, mm)
        return g @ post + dpb @ dp

    h = 1e-6
    v = rng.standard_normal(19) * W19
    fd = (phi(f + h * v, d, m) - phi(f - h * v, d, m)) / (2 * h)
    assert abs(fd - fb @ v) <= 1e-7 * max(1.0, abs(fd))
    fd = (phi(f, d + h, m) - phi(f, d - h, m)) / (2 * h)
    assert abs(fd - db) <= 1e-7 * max(1.0, abs(fd))
    dm = rng.standard_normal(3)
    fd = (phi(f, d, m + h * dm) - phi(f, d, m - h * dm)) / (2 * h)
    assert abs(fd - mb @ dm) <= 1e-7 * max(1.0, abs(fd))


def _occupancy(shape, seed):
    """A partly occupied design: zeros outside a blob, values across all branches inside."""
    rng = np.random.default_rng(seed)
    d = random_alpha(shape, seed) * 12.0
    d[rng.random(d.size) < 0.4] = 0.0
    d[rng.random(d.size) < 0.1] = 1.2
    return d


@pytest.mark.parametrize("law", ["psm", "psm_superposition", "brinkman_occupancy"])
def test_step_transpose_and_gradient(law):
    """<lambda, J v> of the coupled step against <J^T lambda, v>, and the gradient of a
    pressure-drop stage cost with respect to the occupancy and the solid momentum
    against central differences (h = 1e-5: the pressure drop's rounding floor is about
    1e-16 on densities near one, against derivatives of 1e-6)."""
    shape = (8, 5, 4)
    n = int(np.prod(shape))
    rng = np.random.default_rng(4)
    m = rng.standard_normal((n, 3)) * 0.01
    prob = orc.Problem(shape=shape, periodic=(False, True, True), omega=1.3, accel=(1e-4, 0.0, 0.0),
                       coupling=law, psm_kappa=0.1, solid_momentum=m, psm_c_alpha=C_ALPHA,
                       psm_drag_shape=DRAG_SHAPE)
    o = orc.Oracle(prob)
    d = _occupancy(shape, 5)
    f = o.run(d, 4)
    lam = rng.standard_normal(n * 19)
    v = rng.standard_normal(n * 19) * 1e-3
    h = 1e-6
    jv = (o.interval(d, f + h * v) - o.interval(d, f - h * v)) / (2 * h)
    fb, _ = o.interval_vjp(d, f, lam)
    assert abs(lam @ jv - fb @ v) <= 1e-7 * abs(fb @ v)

    steps = 10
    ob = orc.Objective(kind="pressure_drop", window=(5, steps), weight=1.0, axis=0, plane_a=1, plane_b=6)
    mbar = np.zeros((n, 3))
    _, db, _, _ = orc.scan_adjoint(o, d, steps, ob, m_bar=mbar)
    for idx in np.nonzero(d > 0)[0][:4]:
        fd, rel = orc.fd_vs_ad_probe(lambda x: orc.objective_value(o, x, steps, ob), d, int(idx), None, db[idx],
                                     step=1e-5)
        assert rel <= 1e-5, (idx, fd, db[idx])
    assert np.all(db[d == 0] == 0.0)

    def with_m(mm):
        p2 = orc.Problem(shape=shape, periodic=(False, True, True), omega=1.3, accel=(1e-4, 0.0, 0.0),
                         coupling=law, psm_kappa=0.1, solid_momentum=mm, psm_c_alpha=C_ALPHA,
                         psm_drag_shape=DRAG_SHAPE)
        return orc.objective_value(orc.Oracle(p2), d, steps, ob)

    x = int(np.nonzero(d > 0)[0][0])
    for a in range(3):
        mp, mm = m.copy(), m.copy()
        mp[x, a] += 1e-5
        mm[x, a] -= 1e-5
        fd = (with_m(mp) - with_m(mm)) / 2e-5
        assert abs(fd - mbar[x, a]) <= 1e-5 * max(abs(fd), 1e-8), (a, fd, mbar[x, a])


@pytest.mark.parametrize("law", ["psm", "psm_superposition", "brinkman_occupancy"])
def test_solid_force_objective_gradient(law):
    """The stage cost w . (momentum exchanged with the solid) over a window of states:
    its value is the sum of the cells' dp, and the design and solid-momentum gradients
    through scan_adjoint match central differences."""
    shape, steps = (8, 5, 4), 10
    n = int(np.prod(shape))
    rng = np.random.default_rng(9)
    m = rng.standard_normal((n, 3)) * 0.01
    kw = dict(shape=shape, periodic=(False, True, True), omega=1.3, accel=(1e-4, 0.0, 0.0), coupling=law,
              psm_kappa=0.1, psm_c_alpha=C_ALPHA, psm_drag_shape=DRAG_SHAPE)
    o = orc.Oracle(orc.Problem(solid_momentum=m, **kw))
    d = _occupancy(shape, 10)
    ob = orc.Objective(kind="solid_force", window=(4, steps - 1), weight=0.5, vector_weights=(1.0, -0.3, 0.2, 0, 0, 0))
    f = o.run(d, 3)
    want = 0.0
    for x in np.nonzero(d > 0)[0]:
        want += np.array([1.0, -0.3, 0.2]) @ o.psm_cell(f.reshape(-1, 19)[x], d[x], m[x])[1]
    assert abs(o.solid_force(d, f, (1.0, -0.3, 0.2)) - want) <= 1e-15 * max(1.0, abs(want)) * 50
    mbar = np.zeros((n, 3))
    _, db, _, _ = orc.scan_adjoint(o, d, steps, ob, m_bar=mbar)
    # cells with d well above zero: the coupling has terms in m^2 / d
    for idx in np.nonzero(d > 0.05)[0][:3]:
        fd, rel = orc.fd_vs_ad_probe(lambda x: orc.objective_value(o, x, steps, ob), d, int(idx), None, db[idx],
                                     step=1e-6)
        assert rel <= 1e-5, (idx, fd, db[idx])
    x = int(np.nonzero(d > 0.05)[0][1])
    for a in range(3):
        mp, mm = m.copy(), m.copy()
        mp[x, a] += 1e-5
        mm[x, a] -= 1e-5
        fp = orc.objective_value(orc.Oracle(orc.Problem(solid_momentum=mp, **kw)), d, steps, ob)
        fm = orc.objective_value(orc.Oracle(orc.Problem(solid_momentum=mm, **kw)), d, steps, ob)
        fd = (fp - fm) / 2e-5
        assert abs(fd - mbar[x, a]) <= 1e-5 * max(abs(fd), 1e-8), (a, fd, mbar[x, a])


def test_occupancy_endpoints():
    """A second endpoint (d1, m1) blended as d0 (1 - theta) + d1 theta over steps t0..t1
    (field.rs:1266-1275): equal endpoints give the single-occupancy states and split its
    gradient into parts that sum to it; distinct endpoints' gradients match central
    differences."""
    shape, steps = (8, 5, 4), 10
    n = int(np.prod(shape))
    rng = np.random.default_rng(12)
    m0 = rng.standard_normal((n, 3)) * 0.01
    d0 = _occupancy(shape, 13)
    kw = dict(shape=shape, periodic=(False, True, True), omega=1.3, accel=(1e-4, 0.0, 0.0), coupling="psm",
              psm_kappa=0.1)
    ob = orc.Objective(kind="pressure_drop", window=(4, steps), weight=1.0, axis=0, plane_a=1, plane_b=6)
    single = orc.Oracle(orc.Problem(solid_momentum=m0, **kw))
    same = orc.Oracle(orc.Problem(solid_momentum=m0, occupancy_end=(d0, m0, 2, 8), **kw))
    assert np.abs(single.run(d0, steps) - same.run(d0, steps)).max() <= 1e-15
    mb = np.zeros((n, 3))
    _, db, _, _ = orc.scan_adjoint(single, d0, steps, ob, m_bar=mb)
    mb0, end = np.zeros((n, 3)), (np.zeros(n), np.zeros((n, 3)))
    _, db0, _, _ = orc.scan_adjoint(same, d0, steps, ob, m_bar=mb0, end_bar=end)
    assert np.abs(db0 + end[0] - db).max() <= 1e-13 * np.abs(db).max()
    assert np.abs(mb0 + end[1] - mb).max() <= 1e-13 * np.abs(mb).max()
    assert np.abs(end[0]).max() > 0.1 * np.abs(db).max()

    d1 = np.clip(d0 + 0.2 * rng.standard_normal(n) * (d0 > 0), 0.0, None)
    m1 = m0 + rng.standard_normal((n, 3)) * 0.005
    mov = orc.Oracle(orc.Problem(solid_momentum=m0, occupancy_end=(d1, m1, 2, 8), **kw))
    end = (np.zeros(n), np.zeros((n, 3)))
    _, db0, _, _ = orc.scan_adjoint(mov, d0, steps, ob, end_bar=end)

    def J(dd1):
        o = orc.Oracle(orc.Problem(solid_momentum=m0, occupancy_end=(dd1, m1, 2, 8), **kw))
        return orc.objective_value(o, d0, steps, ob)

    for idx in np.nonzero((d0 > 0.05) & (d1 > 0.05))[0][:3]:
        fd, rel = orc.fd_vs_ad_probe(J, d1, int(idx), None, end[0][idx], step=1e-5)
        assert rel <= 1e-5, (idx, fd, end[0][idx])
        fd, rel = orc.fd_vs_ad_probe(lambda x: orc.objective_value(mov, x, steps, ob), d0, int(idx), None, db0[idx],
                                     step=1e-5)
        assert rel <= 1e-5, (idx, fd, db0[idx])
