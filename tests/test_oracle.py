"""Pins of the CPU oracle (run without a GPU).

The reference's LBM and AD crates carry no tests, fixtures or golden vectors
(crates/implexity-physics-lbm/Cargo.toml:29,67-68), so the oracle is pinned by the
reference's own tables, closed forms, invariants, special cases and finite
differences, each chosen so that a plausible slip (a wrong sign, index or weight, a
dropped term, a transposed operand) fails at least one of them.
"""
from __future__ import annotations

import itertools

import numpy as np
import pytest

from oracle import oracle as orc
from oracle import revolve as rev

# d3q19.rs:29-61 -- the reference's own tables (a golden vector held by its source).
REF_D3Q19_C = [[0, 0, 0], [-1, -1, 0], [-1, 0, -1], [-1, 0, 0], [-1, 0, 1], [-1, 1, 0], [0, -1, -1], [0, -1, 0],
               [0, -1, 1], [0, 0, -1], [0, 0, 1], [0, 1, -1], [0, 1, 0], [0, 1, 1], [1, -1, 0], [1, 0, -1],
               [1, 0, 0], [1, 0, 1], [1, 1, 0]]
REF_D3Q19_OPP = [0, 18, 17, 16, 15, 14, 13, 12, 11, 10, 9, 8, 7, 6, 5, 4, 3, 2, 1]
REF_D3Q19_W = [1 / 3, 1 / 36, 1 / 36, 1 / 18, 1 / 36, 1 / 36, 1 / 36, 1 / 18, 1 / 36, 1 / 18, 1 / 18, 1 / 36,
               1 / 18, 1 / 36, 1 / 36, 1 / 36, 1 / 18, 1 / 36, 1 / 36]


def test_d3q19_tables_match_reference():
    c, w, opp = orc.lattice(19)
    assert c.tolist() == REF_D3Q19_C
    assert opp.tolist() == REF_D3Q19_OPP
    assert np.array_equal(w, np.array(REF_D3Q19_W))


@pytest.mark.parametrize("Q", [19, 27])
def test_lattice_isotropy(Q):
    c, w, opp = orc.lattice(Q)
    assert abs(w.sum() - 1.0) < 1e-15
    assert np.abs(w @ c).max() < 1e-15
    np.testing.assert_allclose(np.einsum("q,qa,qb->ab", w, c, c), np.eye(3) / 3, atol=1e-15)
    assert np.array_equal(c[opp], -c)
    if Q == 27:
        # moving/lattice.rs:93-116: rest first, then lexicographic (a,b,c) without origin
        lex = [list(t) for t in itertools.product([-1, 0, 1], repeat=3) if t != (0, 0, 0)]
        assert c.tolist() == [[0, 0, 0]] + lex


def _brute_source(shape, periodic, solid, Q):
    """Halfway bounce-back by geometry: the population arriving at x along c_q comes from
    x - c_q unless that cell is solid or lies beyond a closed face; then it is x's own
    population of the opposite direction."""
    c, _, opp = orc.lattice(Q)
    nx, ny, nz = shape
    n = nx * ny * nz
    src = np.zeros(n * Q, np.uint32)
    for x in range(n):
        i, j, k = x // (ny * nz), (x // nz) % ny, x % nz
        for q in range(Q):
            if solid[x]:
                src[x * Q + q] = 0xFFFFFFFF
                continue
            y = [i - c[q, 0], j - c[q, 1], k - c[q, 2]]
            out = any((not periodic[a]) and not (0 <= y[a] < shape[a]) for a in range(3))
            y = [y[a] % shape[a] for a in range(3)]
            yx = (y[0] * ny + y[1]) * nz + y[2]
            src[x * Q + q] = x * Q + opp[q] if (out or solid[yx]) else yx * Q + q
    return src


@pytest.mark.parametrize("Q", [19, 27])
@pytest.mark.parametrize("periodic", [(False, False, False), (True, False, True), (True, True, True)])
def test_stream_table_brute_force(Q, periodic):
    shape = (5, 4, 3)
    solid = np.random.default_rng(1).random(shape) < 0.15
    o = orc.Oracle(orc.Problem(shape=shape, periodic=periodic, Q=Q, solid=solid))
    assert np.array_equal(o.source, _brute_source(shape, periodic, solid.reshape(-1), Q))


def _cell_problem(Q, kind, **kw):
    return orc.Oracle(orc.Problem(shape=(3, 3, 3), periodic=(True, True, True), Q=Q, kind=kind, **kw))


KINDS = [(19, "bgk"), (19, "trt"), (19, "mrt"), (19, "bgk_guo"), (19, "regularized"), (27, "trt"), (27, "mrt"),
         (27, "regularized"), (27, "cumulant")]


@pytest.mark.parametrize("Q,kind", KINDS)
def test_collision_conserves_mass_and_adds_force(Q, kind):
    """sum f' = sum f, and sum c f' = sum c f + F with F = rho (a - alpha u)."""
    c, w, _ = orc.lattice(Q)
    a = np.array([3e-4, -2e-4, 1e-4])
    o = _cell_problem(Q, kind, omega=1.3, accel=tuple(a), bulk_rate=1.0 if kind == "cumulant" else float("nan"))
    rng = np.random.default_rng(2)
    for _ in range(5):
        f = w * (1 + 0.1 * rng.standard_normal(Q))
        alpha = 0.05 * rng.random()
        out = o.cell_collide(f, alpha)
        rho = f.sum()
        j = c.T @ f
        u = (j / rho + a / 2) / (1 + alpha / 2)
        F = rho * (a - alpha * u)
        assert abs(out.sum() - rho) < 1e-14
        np.testing.assert_allclose(c.T @ out - j, F, atol=1e-15)


@pytest.mark.parametrize("Q,kind", [(19, "bgk"), (19, "trt"), (19, "mrt"), (27, "trt"), (27, "mrt")])
def test_equilibrium_is_fixed_point(Q, kind):
    c, w, _ = orc.lattice(Q)
    o = _cell_problem(Q, kind, omega=1.7)
    rho, u = 1.02, np.array([0.03, -0.02, 0.01])
    cu = c @ u
    eq = rho * w * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * u @ u)
    np.testing.assert_allclose(o.cell_collide(eq, 0.0), eq, atol=1e-16)


@pytest.mark.parametrize("Q", [19, 27])
def test_mrt_default_rates_reproduce_trt(Q):
    """options.rs:56-59: MRT with default rates reproduces TRT exactly."""
    c, w, _ = orc.lattice(Q)
    kw = dict(omega=1.4, magic=0.21, accel=(1e-4, 2e-4, -1e-4))
    trt, mrt = _cell_problem(Q, "trt", **kw), _cell_problem(Q, "mrt", **kw)
    rng = np.random.default_rng(3)
    for _ in range(4):
        f = w * (1 + 0.1 * rng.standard_normal(Q))
        np.testing.assert_allclose(mrt.cell_collide(f, 0.02), trt.cell_collide(f, 0.02), atol=2e-16)


def test_bgk_is_trt_with_equal_rates():
    """collision.rs:443: BGK = collide_pairs with omega- = omega+; TRT gives omega- =
    omega+ when magic = (1/omega - 1/2)^2."""
    omega = 1.25
    magic = (1 / omega - 0.5) ** 2
    assert abs(orc.odd_rate(omega, magic) - omega) < 1e-15
    c, w, _ = orc.lattice(19)
    bgk, trt = _cell_problem(19, "bgk", omega=omega), _cell_problem(19, "trt", omega=omega, magic=magic)
    f = w * (1 + 0.05 * np.random.default_rng(4).standard_normal(19))
    np.testing.assert_allclose(trt.cell_collide(f, 0.01), bgk.cell_collide(f, 0.01), atol=1e-16)


@pytest.mark.parametrize("Q", [19, 27])
def test_regularized_at_unit_rate_is_bgk(Q):
    """Regularized relax (collision.rs:376-394) keeps bulk and shear moments at rate omega
    and removes every other moment (rate 1): at omega = 1 every moment has rate 1, which
    is BGK with omega = 1 (collide_pairs, a different code path)."""
    c, w, _ = orc.lattice(Q)
    a = (2e-4, -1e-4, 5e-5)
    reg, bgk = _cell_problem(Q, "regularized", omega=1.0, accel=a), _cell_problem(Q, "bgk", omega=1.0, accel=a)
    f = w * (1 + 0.05 * np.random.default_rng(6).standard_normal(Q))
    np.testing.assert_allclose(reg.cell_collide(f, 0.02), bgk.cell_collide(f, 0.02), rtol=0, atol=1e-15)


@pytest.mark.parametrize("Q", [19, 27])
def test_regularized_differs_from_mrt_only_off_bulk_and_shear(Q):
    """Away from omega = 1 the regularized operator differs from MRT with default rates
    (which is TRT) -- the non-hydrodynamic moments are removed instead of relaxed."""
    c, w, _ = orc.lattice(Q)
    reg, mrt = _cell_problem(Q, "regularized", omega=1.6), _cell_problem(Q, "mrt", omega=1.6)
    f = w * (1 + 0.05 * np.random.default_rng(7).standard_normal(Q))
    out_r, out_m = reg.cell_collide(f, 0.0), mrt.cell_collide(f, 0.0)
    assert np.abs(out_r - out_m).max() > 1e-6
    # mass and momentum are conserved by both
    np.testing.assert_allclose(out_r.sum(), f.sum(), rtol=1e-14)
    np.testing.assert_allclose(c.T @ out_r, c.T @ f, atol=1e-15)


def test_stationary_bgk_guo_equals_moving_bgk_without_brinkman():
    """The stationary collide_cell (solver.rs:298) and the moving BGK collide_pairs
    (collision.rs:457) are the same operator; they agree to rounding."""
    c, w, _ = orc.lattice(19)
    a = (2e-4, 0.0, -1e-4)
    g, m = _cell_problem(19, "bgk_guo", omega=1.25, accel=a), _cell_problem(19, "bgk", omega=1.25, accel=a)
    f = w * (1 + 0.05 * np.random.default_rng(5).standard_normal(19))
    np.testing.assert_allclose(g.cell_collide(f, 0.03), m.cell_collide(f, 0.03), atol=1e-16)


def test_cumulant_unit_rates_give_isotropic_second_moments():
    """With omega = 1 and bulk = 1 the post-collision second central moments are
    rho/3 delta (collision.rs:686-694), whatever the input."""
    c, w, _ = orc.lattice(27)
    o = _cell_problem(27, "cumulant", omega=1.0, bulk_rate=1.0)
    f = w * (1 + 0.1 * np.random.default_rng(6).standard_normal(27))
    out = o.cell_collide(f, 0.0)
    rho = f.sum()
    u = (c.T @ f) / rho
    d = c - u
    M2 = np.einsum("q,qa,qb->ab", out, d, d)
    np.testing.assert_allclose(M2, rho / 3 * np.eye(3), atol=1e-15)


def test_cumulant_wick_table():
    """matchings() of collision.rs:216-239: (2n-1)!! pairings; 3 for order 4, 15 for
    order 6; none for odd orders."""
    import ctypes

    counts = np.zeros(27, np.int32)
    pairs = np.zeros(27 * 16 * 3 * 2, np.int32)
    orc.lib().orc_wick_table(ctypes.c_void_p(counts.ctypes.data), ctypes.c_void_p(pairs.ctypes.data))
    for k in range(27):
        e = (k // 9, (k // 3) % 3, k % 3)
        order = sum(e)
        assert counts[k] == (3 if order == 4 else 15 if order == 6 else 0)


@pytest.mark.parametrize("Q,kind", KINDS)
def test_cell_vjp_central_differences(Q, kind):
    c, w, _ = orc.lattice(Q)
    o = _cell_problem(Q, kind, omega=1.3, accel=(2e-4, -1e-4, 3e-4), bulk_rate=1.0 if kind == "cumulant"
                      else float("nan"))
    rng = np.random.default_rng(7)
    f = w * (1 + 0.1 * rng.standard_normal(Q))
    alpha = 0.04
    g = rng.standard_normal(Q)
    fb, ab = o.cell_vjp(f, alpha, g)
    h = 1e-6
    fd = np.array([(g @ o.cell_collide(f + h * e, alpha) - g @ o.cell_collide(f - h * e, alpha)) / (2 * h)
                   for e in np.eye(Q)])
    fda = (g @ o.cell_collide(f, alpha + h) - g @ o.cell_collide(f, alpha - h)) / (2 * h)
    np.testing.assert_allclose(fb, fd, rtol=1e-7, atol=1e-9 * np.abs(fd).max())
    assert abs(ab - fda) <= 1e-7 * max(abs(fda), 1e-3)


def test_port_nee_reproduces_equilibrium():
    """boundary.rs:445-459: from an equilibrium neighbour, a velocity port returns
    f_eq(rho_nb, u_b) and a pressure port f_eq(rho_b, u_nb)."""
    c, w, _ = orc.lattice(19)
    o = _cell_problem(19, "trt")
    rho, u = 1.01, np.array([0.02, 0.01, -0.01])

    def eq(r, v):
        cu = c @ v
        return r * w * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * v @ v)

    ub = np.array([0.03, 0.0, 0.0])
    np.testing.assert_allclose(o.port_populations(eq(rho, u), 0, ub), eq(rho, ub), atol=1e-16)
    np.testing.assert_allclose(o.port_populations(eq(rho, u), 1, [0.99, 0, 0]), eq(0.99, u), atol=1e-16)


@pytest.mark.parametrize("kind", [0, 1])
def test_port_vjp_central_differences(kind):
    c, w, _ = orc.lattice(19)
    o = _cell_problem(19, "trt")
    rng = np.random.default_rng(8)
    nb = w * (1 + 0.1 * rng.standard_normal(19))
    val = [0.03, 0.01, 0.0] if kind == 0 else [1.01, 0, 0]
    g = rng.standard_normal(19)
    got = o.port_populations_vjp(nb, kind, val, g)
    h = 1e-6
    fd = np.array([(g @ o.port_populations(nb + h * e, kind, val) - g @ o.port_populations(nb - h * e, kind, val))
                   / (2 * h) for e in np.eye(19)])
    np.testing.assert_allclose(got, fd, rtol=1e-7, atol=1e-9)


def test_poiseuille_closed_form():
    """Body-force channel with halfway bounce-back walls and TRT magic 3/16: the steady
    profile is the exact parabola u(y) = a/(2 nu) (y + 1/2)(H - y - 1/2) (options.rs:53-55)."""
    H = 6
    tau = 0.8
    nu = (tau - 0.5) / 3
    a = 1e-5
    prob = orc.Problem(shape=(2, H, 2), periodic=(True, False, True), Q=19, kind="trt", omega=1 / tau,
                       accel=(a, 0, 0))
    o = orc.Oracle(prob)
    alpha = np.zeros(o.n)
    f = o.run(alpha, 6000)
    _, u = o.macroscopic(alpha, f)
    y = np.arange(H)
    exact = a / (2 * nu) * (y + 0.5) * (H - y - 0.5)
    got = u.reshape(2, H, 2, 3)[0, :, 0, 0]
    np.testing.assert_allclose(got, exact, rtol=1e-9)


def test_darcy_closed_form():
    """Uniform Brinkman resistance in a periodic box: steady u = a / alpha."""
    a, alpha0 = 2e-5, 0.05
    o = orc.Oracle(orc.Problem(shape=(3, 3, 3), periodic=(True, True, True), Q=19, kind="trt", omega=1.25,
                               accel=(a, 0, 0)))
    alpha = np.full(o.n, alpha0)
    f = o.run(alpha, 1000)
    _, u = o.macroscopic(alpha, f)
    np.testing.assert_allclose(u[:, 0], a / alpha0, rtol=1e-10)


def test_mass_conserved_in_closed_box():
    shape = (6, 5, 4)
    solid = np.random.default_rng(9).random(shape) < 0.1
    o = orc.Oracle(orc.Problem(shape=shape, Q=19, kind="trt", omega=1.6, accel=(1e-4, 0, 0), solid=solid))
    alpha = np.random.default_rng(10).random(o.n) * 0.05
    fluid = ~solid.reshape(-1)
    f0 = o.initial()
    f = o.run(alpha, 50)
    m0 = f0.reshape(-1, 19)[fluid].sum()
    assert abs(f.reshape(-1, 19)[fluid].sum() - m0) <= 1e-13 * m0


def _interval_pair():
    shape = (6, 5, 4)
    from implexity_amd.design import parabolic_profile

    prof = parabolic_profile(5, 4).reshape(-1)
    prob = orc.Problem(shape=shape, Q=19, kind="trt", omega=1.4,
                       ports=[orc.Port("xmin", "velocity", (0.03, 0, 0), prof), orc.Port("xmax", "pressure", (1.0,))])
    return orc.Oracle(prob)


def test_interval_vjp_dot_product_identity():
    """<lambda, J v> = <J^T lambda, v> with J v from central differences of interval."""
    o = _interval_pair()
    rng = np.random.default_rng(11)
    alpha = rng.random(o.n) * 0.05
    f = o.run(alpha, 3)
    lam = rng.standard_normal(f.size)
    v = rng.standard_normal(f.size) * 1e-3
    h = 1e-4
    jv = (o.interval(alpha, f + h * v) - o.interval(alpha, f - h * v)) / (2 * h)
    fb, _ = o.interval_vjp(alpha, f, lam)
    assert abs(lam @ jv - fb @ v) <= 1e-9 * abs(lam @ jv)


def test_scan_adjoint_central_differences():
    o = _interval_pair()
    alpha = np.random.default_rng(12).random(o.n) * 0.05
    steps = 12
    ob = orc.Objective(kind="pressure_drop", window=(6, 12), weight=1 / 7, axis=0, plane_a=1, plane_b=4)
    J, ab, lam0, _ = orc.scan_adjoint(o, alpha, steps, ob, every=4)
    J_all, ab_all, _, _ = orc.scan_adjoint(o, alpha, steps, ob)
    assert J == J_all and np.array_equal(ab, ab_all)
    for x in np.argsort(-np.abs(ab))[:5]:
        fd, rel = orc.fd_vs_ad_probe(lambda a: orc.objective_value(o, a, steps, ob), alpha, int(x), J, ab[x],
                                     step=1e-6)
        assert rel < 1e-6, (x, ab[x], fd)


def test_dissipation_gradient_central_differences():
    o = _interval_pair()
    rng = np.random.default_rng(13)
    alpha = rng.random(o.n) * 0.05
    f = o.run(alpha, 4)
    fb, ab = o.dissipation_grad(alpha, f)
    h = 1e-7
    for idx in rng.choice(f.size, 6, replace=False):
        e = np.zeros(f.size)
        e[idx] = h
        fd = (o.dissipation(alpha, f + e) - o.dissipation(alpha, f - e)) / (2 * h)
        assert abs(fd - fb[idx]) <= 1e-6 * abs(fd) + 1e-12
    for x in rng.choice(o.n, 4, replace=False):
        a2 = alpha.copy()
        a2[x] += h
        a3 = alpha.copy()
        a3[x] -= h
        fd = (o.dissipation(a2, f) - o.dissipation(a3, f)) / (2 * h)
        assert abs(fd - ab[x]) <= 1e-6 * abs(fd) + 1e-12


SCALES = (0.7, 998.0, 2.5e-3, 1.3e-6)  # velocity, density, step_s, cell volume


def test_vector_closed_form_on_equilibrium_states():
    """vector (solver.rs:874-894) of f = f_eq(rho0, U) with uniform alpha0: the Brinkman
    velocity is U / (1 + alpha0/2) (solver.rs:189-196), so mean u, kinetic energy and
    porous dissipation follow in closed form; the volume is sum(phi) dV."""
    o = _interval_pair()
    c, w, _ = orc.lattice(o.Q)
    rho0, U, alpha0 = 1.02, np.array([0.03, -0.01, 0.005]), 0.04
    cu = c @ U
    eq = w * rho0 * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * U @ U)
    state = np.tile(eq, o.n)
    alpha = np.full(o.n, alpha0)
    phi = np.random.default_rng(3).random(o.n)
    v = o.vector(phi, alpha, state, SCALES)
    vel, dens, dt, dv = SCALES
    u = U / (1 + alpha0 / 2) * vel
    usq = u @ u
    np.testing.assert_allclose(v[:3], u, rtol=1e-13)
    np.testing.assert_allclose(v[3], 0.5 * rho0 * dens * usq * o.n * dv, rtol=1e-12)
    np.testing.assert_allclose(v[4], rho0 * dens * alpha0 / dt * usq * o.n * dv, rtol=1e-12)
    np.testing.assert_allclose(v[5], phi.sum() * dv, rtol=1e-13)


def test_vector_vjp_central_differences():
    """vector_vjp (solver.rs:896-947) against central differences of w . vector."""
    o = _interval_pair()
    rng = np.random.default_rng(17)
    alpha = rng.random(o.n) * 0.05
    phi = rng.random(o.n)
    f = o.run(alpha, 4)
    wv = rng.standard_normal(6)
    fb, ab, pb = o.vector_vjp(phi, alpha, f, SCALES, wv)
    val = lambda ff, aa, pp: float(wv @ o.vector(pp, aa, ff, SCALES))
    h = 1e-7
    for idx in rng.choice(f.size, 6, replace=False):
        e = np.zeros(f.size)
        e[idx] = h
        fd = (val(f + e, alpha, phi) - val(f - e, alpha, phi)) / (2 * h)
        assert abs(fd - fb[idx]) <= 1e-6 * abs(fd) + 1e-10 * np.abs(fb).max()
    for x in rng.choice(o.n, 4, replace=False):
        e = np.zeros(o.n)
        e[x] = h
        fd = (val(f, alpha + e, phi) - val(f, alpha - e, phi)) / (2 * h)
        assert abs(fd - ab[x]) <= 1e-6 * abs(fd) + 1e-10 * np.abs(ab).max()
        fd = (val(f, alpha, phi + e) - val(f, alpha, phi - e)) / (2 * h)
        assert abs(fd - pb[x]) <= 1e-6 * abs(fd) + 1e-12


def test_pressure_drop_uniform_state():
    o = _interval_pair()
    f = o.initial()
    assert o.pressure_drop(f, 0, 1, 4) == 0.0
    g = o.pressure_drop_grad(0, 1, 4)
    # d/df of (mean rho_a - mean rho_b)/3: 1/(3 n_plane) on plane a, minus on plane b
    assert np.isclose(g.reshape(6, 5, 4, 19)[1].mean(), 1 / (3 * 20))
    assert np.isclose(g.reshape(6, 5, 4, 19)[4].mean(), -1 / (3 * 20))


@pytest.mark.parametrize("n,s", [(1, 1), (2, 1), (5, 1), (9, 2), (30, 3), (64, 4), (100, 9), (257, 5)])
def test_revolve_replay_and_optimality(n, s):
    acts, _, fwd, peak = rev.binomial_schedule(n, s)
    assert fwd == rev.optimal_forward_steps(n, s)  # revolve.rs:291
    assert rev.replay_check(n, s) == fwd
    assert peak <= s


def test_revolve_closed_forms():
    # beta(s, r) = C(s+r, r)
    from math import comb

    for s in range(1, 8):
        for r in range(0, 6):
            assert rev.beta(s, r) == comb(s + r, r)
    # one slot: each reversal recomputes from state 0: n + (n-2) + ... + 1 + 0 steps
    for n in range(2, 12):
        assert rev.optimal_forward_steps(n, 1) == n * (n - 1) // 2 + 1
    # enough slots: one forward sweep
    assert rev.optimal_forward_steps(50, 50) == 50
