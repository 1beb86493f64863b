"""Oracle pins of the sponge layers (moving/sponge.rs, blended after collision as
field.rs:1461-1475; far-field targets field.rs:1015-1036) on the stationary step."""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import oracle as orc
from tests.lbm_cases import random_alpha


def test_ramp_is_the_quintic_smoothstep():
    """ramp is the quintic with value 0/1 and vanishing first and second derivatives at
    both ends (solved here from those six conditions), clamped outside [0, 1]; the
    signal's ramp (boundary.rs:112-118) is the same polynomial."""
    A = np.array([[1, 0, 0, 0, 0, 0], [0, 1, 0, 0, 0, 0], [0, 0, 2, 0, 0, 0],
                  [1, 1, 1, 1, 1, 1], [0, 1, 2, 3, 4, 5], [0, 0, 2, 6, 12, 20]], float)
    coef = np.linalg.solve(A, np.array([0, 0, 0, 1, 0, 0], float))
    for s in np.linspace(0.0, 1.0, 101):
        want = sum(c * s ** k for k, c in enumerate(coef))
        assert abs(orc.sponge_ramp(s) - want) <= 1e-14
        r, _ = orc.signal_factors(10.0, 0.0, 0.0, 10.0 * s)
        assert abs(r - want) <= 1e-14
    assert orc.sponge_ramp(-0.3) == 0.0 and orc.sponge_ramp(1.7) == 1.0
    assert orc.signal_factors(10.0, 0.0, 0.0, 25.0) == (1.0, 0.0)


def test_sigma_profile_mirrors_between_opposite_faces():
    """Equal sponges on opposite faces give a mirror-symmetric strength that is largest
    at the face, falls monotonically inward, covers exactly `thickness` layers and is
    zero elsewhere."""
    shape = (11, 3, 4)
    for t in (1, 2, 5):
        sp = [orc.Sponge("xmin", t, 0.6), orc.Sponge("xmax", t, 0.6)]
        cells, sigma, off, part, weight, phase = orc.sponge_field(shape, (False, True, True), sp)
        full = np.zeros(np.prod(shape))
        full[cells] = sigma
        prof = full.reshape(shape)
        assert np.all(prof == prof[:, :1, :1])  # depends on x only
        line = prof[:, 0, 0]
        assert np.array_equal(line, line[::-1])
        assert np.count_nonzero(line) == 2 * t
        assert np.all(np.diff(line[:t]) < 0) and line[0] < 0.6
        assert np.all(line[t:shape[0] - t] == 0.0)
        assert np.array_equal(np.diff(off), np.ones(len(cells), np.int32))
        assert np.array_equal(phase.reshape(-1, 2), np.tile([1.0, 0.0], (len(cells), 1)))


def test_corner_cells_sum_their_sponges():
    shape = (6, 6, 3)
    sp = [orc.Sponge("xmin", 2, 0.3), orc.Sponge("ymax", 3, 0.4)]
    cells, sigma, off, part, weight, _ = orc.sponge_field(shape, (False, False, True), sp)
    for r, x in enumerate(cells):
        w = weight[off[r]:off[r + 1]]
        assert list(part[off[r]:off[r + 1]]) == sorted(part[off[r]:off[r + 1]])
        assert sigma[r] == (w.sum() if len(w) == 1 else w[0] + w[1])
    i, j = cells // 18, (cells // 3) % 6
    two = (i < 2) & (j >= 3)
    assert np.array_equal(np.diff(off)[two], np.full(two.sum(), 2))
    assert np.array_equal(np.diff(off)[~two], np.ones((~two).sum(), np.int32))


@pytest.mark.parametrize("sp,periodic,msg", [
    (orc.Sponge("xmin", 2, 0.5), (True, False, False), "periodic axis"),
    (orc.Sponge("ymax", 0, 0.5), (False, False, False), "half the lattice"),
    (orc.Sponge("ymax", 4, 0.5), (False, False, False), "half the lattice"),
    (orc.Sponge("zmin", 2, 1.0), (False, False, False), r"\(0, 1\)"),
    (orc.Sponge("zmin", 2, 0.0), (False, False, False), r"\(0, 1\)"),
    (orc.Sponge("zmin", 2, 0.5, density=float("nan")), (False, False, False), "finite"),
    (orc.Sponge("zmin", 2, 0.5, wave=orc.Wave((0.01, 0, 0), (1.0, 1.0, 0), 0.1, 0.01)), (False, False, False),
     "unit direction"),
    (orc.Sponge("zmin", 2, 0.5, wave=orc.Wave((0.01, 0, 0), (1.0, 0, 0), 0.0, 0.01)), (False, False, False),
     "positive speed"),
])
def test_contract(sp, periodic, msg):
    with pytest.raises(ValueError, match=msg):
        orc.sponge_field((6, 7, 6), periodic, [sp])


def test_summed_strength_below_one():
    sp = [orc.Sponge("xmin", 2, 0.7), orc.Sponge("xmin", 2, 0.7)]
    with pytest.raises(ValueError, match="below one"):
        orc.sponge_field((6, 4, 4), (False, True, True), sp)


def _problem(shape, sponges, periodic=(False, True, True), **kw):
    return orc.Problem(shape=shape, periodic=periodic, omega=1.3, sponges=sponges, **kw)


def test_far_field_equilibrium_is_a_fixed_point():
    """A resting state at the sponge's far-field density is kept by collision, walls and
    the blend alike."""
    shape = (8, 4, 5)
    rho = 1.02
    o = orc.Oracle(_problem(shape, [orc.Sponge("xmin", 3, 0.5, density=rho), orc.Sponge("xmax", 2, 0.3,
                                                                                         density=rho)]))
    c, w, _ = orc.lattice(19)
    f = np.tile(w * rho, o.n)
    after = o.interval(np.zeros(o.n), f, t=3)
    assert np.abs(after - f).max() <= 1e-15


def test_blend_moves_mass_and_momentum_to_the_target():
    """At every sponge cell the blend's density and momentum are (1 - sigma) times the
    collided ones plus sum s_l rho_l (and sum s_l rho_l u_l), with the travelling wave's
    velocity mean + A ramp sin(2 pi f t + phi - theta(x)); other cells are untouched."""
    shape = (9, 5, 4)
    wave = orc.Wave((0.01, -0.004, 0.002), (0.6, 0.8, 0.0), 0.3, 0.02, phase=0.4, ramp=20.0, origin=(1.0, 2.0, 0.0))
    sp = [orc.Sponge("xmin", 3, 0.45, density=1.01, velocity=(0.02, 0.0, 0.0), wave=wave),
          orc.Sponge("xmax", 2, 0.25, density=0.99)]
    with_s = orc.Oracle(_problem(shape, sp, accel=(1e-4, 0.0, 0.0)))
    without = orc.Oracle(_problem(shape, [], accel=(1e-4, 0.0, 0.0)))
    alpha = random_alpha(shape, 3)
    f = without.run(alpha, 7)
    t = 13
    _, _, post_s = with_s.interval(alpha, f, want_post=True, t=t)
    _, _, post_0 = without.interval(alpha, f, want_post=True, t=t)
    c, _, _ = orc.lattice(19)
    cells, sigma, off, part, weight, _ = with_s._sp
    covered = np.zeros(with_s.n, bool)
    covered[cells] = True
    P_s, P_0 = post_s.reshape(-1, 19), post_0.reshape(-1, 19)
    assert np.array_equal(P_s[~covered], P_0[~covered])
    ramp = orc.signal_factors(wave.ramp, wave.frequency, wave.phase, float(t))[0]
    for r, x in enumerate(cells):
        i, j, k = with_s.coords(int(x))
        pos = np.array([i, j, k], float) + 0.5
        m_want = (1.0 - sigma[r]) * P_0[x].sum()
        j_want = (1.0 - sigma[r]) * (P_0[x] @ c)
        for q in range(off[r], off[r + 1]):
            s = sp[part[q]]
            u = np.array(s.velocity, float)
            if s.wave is not None:
                theta = 2 * math.pi * wave.frequency * ((pos - wave.origin) @ np.array(wave.direction)) / wave.speed
                u = u + np.array(wave.amplitude) * ramp * math.sin(2 * math.pi * wave.frequency * t + wave.phase
                                                                     - theta)
            m_want += weight[q] * s.density
            j_want = j_want + weight[q] * s.density * u
        assert abs(P_s[x].sum() - m_want) <= 1e-14
        assert np.abs(P_s[x] @ c - j_want).max() <= 1e-15


def test_travelling_wave_moves_along_its_direction():
    """After the ramp, the target at a cell one cell further along the direction equals
    the earlier target delayed by 1 / speed steps."""
    shape = (4, 4, 12)
    wave = orc.Wave((0.0, 0.0, 0.01), (0.0, 0.0, 1.0), 0.25, 0.01, phase=0.3)
    sp = [orc.Sponge("xmin", 2, 0.4, wave=wave)]
    cells, _, off, part, _, phase = orc.sponge_field(shape, (False, True, True), sp)
    ph = {int(x): phase[2 * off[r]: 2 * off[r] + 2] for r, x in enumerate(cells)}

    def u(x, t):
        tg = orc.sponge_targets(sp, t)[0]
        return tg[1:4] + tg[4:7] * ph[x][0] - tg[7:10] * ph[x][1]

    for t in (0.0, 7.0, 31.0):
        for k in range(0, 10):
            x0, x1 = k, k + 1  # (0, 0, k) and (0, 0, k + 1)
            assert np.abs(u(x1, t + 4.0) - u(x0, t)).max() <= 1e-15
    assert np.abs(u(0, 2.0) - u(0, 0.0)).max() > 1e-4


def test_step_vjp_dot_product_and_gradient():
    """<lambda, J v> of the sponge step (central differences) against <J^T lambda, v>,
    and the design gradient of a stage cost through sponge steps against finite
    differences."""
    shape = (8, 5, 4)
    wave = orc.Wave((0.006, 0.0, 0.0), (1.0, 0.0, 0.0), 0.2, 0.03, ramp=4.0)
    sp = [orc.Sponge("xmin", 2, 0.5, density=1.0, velocity=(0.03, 0.0, 0.0), wave=wave),
          orc.Sponge("xmax", 3, 0.4, density=1.005)]
    prob = _problem(shape, sp, periodic=(False, True, False))
    o = orc.Oracle(prob)
    alpha = random_alpha(shape, 5)
    f = o.run(alpha, 5)
    rng = np.random.default_rng(2)
    lam = rng.standard_normal(o.n * 19)
    v = rng.standard_normal(o.n * 19)
    h = 1e-6
    jv = (o.interval(alpha, f + h * v, t=5) - o.interval(alpha, f - h * v, t=5)) / (2 * h)
    fb, _ = o.interval_vjp(alpha, f, lam, t=5)
    assert abs(lam @ jv - fb @ v) <= 1e-7 * abs(fb @ v)

    steps = 12
    ob = orc.Objective(kind="dissipation", window=(6, steps), weight=1.0)
    _, ab, _, _ = orc.scan_adjoint(o, alpha, steps, ob)
    for idx in (3, 61, 150):
        fd, rel = orc.fd_vs_ad_probe(lambda a: orc.objective_value(o, a, steps, ob), alpha, idx, None, ab[idx])
        assert rel <= 1e-6, (idx, fd, ab[idx])
