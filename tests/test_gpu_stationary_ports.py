"""GPU parity of the stationary solver's port rules (crates/implexity-physics-lbm/src/
ports.rs) through the C ABI against the FP64 oracle: the moment reconstruction of
"legacy_moment" ports (:343-406), the neighbour non-equilibrium pressure rule
(:484-497), on masked faces (ports.rs:170-200).

Bars as for the other ports: the FP64 forward bit-exact (with the port cells and the
sources bit-exact too), the step VJP within 1e-12 of the oracle's apply_ports_vjp
(:582-621; analytic transposes on the device, the dual-number cell Jacobian in the
oracle), the gradient within 1e-11, the mass drift with the port exchange at rounding,
FP32 and FP16 storage within the usual bounds of the largest population deviation.
"""
from __future__ import annotations

import numpy as np
import pytest

from implexity_amd.lbm import Lattice, Objective
from oracle import oracle as orc
from tests.lbm_cases import make_pair, random_alpha, random_solid, stationary_ports

pytestmark = pytest.mark.gpu

CASES = [
    # shape, periodic, collision, inlet, outlet, accel, tangential, inset, with_solid
    ((8, 6, 5), (False, False, False), "bgk_guo", "velocity_moment", "pressure_neighbor", (0, 0, 0), (0.0, 0.0),
     (1, 1), False),
    ((8, 6, 8), (False, False, True), "bgk_guo", "velocity_moment", "pressure_moment", (1e-4, 2e-5, 0),
     (0.0, 0.0), (1, 0), True),
    ((7, 6, 5), (False, True, True), "trt", "pressure_moment", "pressure_neighbor", (0, 0, 0), (0.003, -0.002),
     (0, 0), False),
    ((7, 5, 6), (False, False, False), "mrt", "velocity_moment", "velocity_moment", (0, 0, 0), (0.0, 0.0),
     (1, 1), True),
    ((8, 6, 5), (False, True, False), "bgk", "pressure_moment", "pressure_moment", (2e-4, 0, -1e-4),
     (0.001, 0.0), (0, 1), False),
]


def _case(shape, periodic, coll, inlet, outlet, accel, tangential, inset, with_solid, precision="f64"):
    ports = stationary_ports(shape, 0.025, 0.997, tangential, inlet, outlet, inset)
    solid = None
    if with_solid:
        solid = random_solid(shape, 11).copy()
        solid[:2] = False
        solid[-2:] = False
    return make_pair(shape, periodic, Q=19, collision=coll, precision=precision, accel=accel, solid=solid,
                     ports=ports), solid


@pytest.mark.parametrize("shape,periodic,coll,inlet,outlet,accel,tangential,inset,with_solid", CASES)
def test_forward_f64_bit_exact(shape, periodic, coll, inlet, outlet, accel, tangential, inset, with_solid):
    (cfg, prob), _ = _case(shape, periodic, coll, inlet, outlet, accel, tangential, inset, with_solid)
    o = orc.Oracle(prob)
    alpha = random_alpha(shape, 5)
    want = o.run(alpha, 14)
    cells, nbs, _, _ = o.port_cells()
    with Lattice(cfg) as lat:
        got_cells, got_nbs = lat.port_cells()
        lat.set_design(alpha)
        lat.initial_populations()
        lat.interval(14)
        got = lat.state()
    assert np.array_equal(got_cells, cells) and np.array_equal(got_nbs, nbs)
    assert np.array_equal(got, want), np.abs(got - want).max()


@pytest.mark.parametrize("shape,periodic,coll,inlet,outlet,accel,tangential,inset,with_solid", CASES)
def test_step_vjp_f64(shape, periodic, coll, inlet, outlet, accel, tangential, inset, with_solid):
    (cfg, prob), solid = _case(shape, periodic, coll, inlet, outlet, accel, tangential, inset, with_solid)
    o = orc.Oracle(prob)
    alpha = random_alpha(shape, 9)
    f = o.run(alpha, 4)
    rng = np.random.default_rng(1)
    lam = rng.standard_normal(o.n * 19)
    if solid is not None:
        lam.reshape(-1, 19)[solid.reshape(-1)] = 0.0
    want_f, want_a = o.interval_vjp(alpha, f, lam)
    with Lattice(cfg) as lat:
        lat.set_design(alpha)
        lat.set_state(f)
        got_f, got_a = lat.interval_vjp(lam)
    fluid = np.ones(o.n, bool) if solid is None else ~solid.reshape(-1)
    gf, wf = got_f.reshape(-1, 19)[fluid], want_f.reshape(-1, 19)[fluid]
    assert np.abs(gf - wf).max() <= 1e-12 * np.abs(wf).max()
    assert np.abs(got_a - want_a).max() <= 1e-12 * np.abs(want_a).max()


@pytest.mark.parametrize("case", [0, 1, 2])
def test_gradient_f64_matches_oracle(case):
    shape, periodic, coll, inlet, outlet, accel, tangential, inset, with_solid = CASES[case]
    (cfg, prob), _ = _case(shape, periodic, coll, inlet, outlet, accel, tangential, inset, with_solid)
    steps = 20
    o = orc.Oracle(prob)
    alpha = random_alpha(shape, 21)
    ob = Objective(kind="pressure_drop", axis=0, plane_a=1, plane_b=shape[0] - 2, t_begin=steps // 2, t_end=steps,
                   weight=1.0 / (steps - steps // 2 + 1))
    oob = orc.Objective(kind=ob.kind, window=(ob.t_begin, ob.t_end), weight=ob.weight, axis=ob.axis,
                        plane_a=ob.plane_a, plane_b=ob.plane_b)
    J_want, ab_want, lam_want, _ = orc.scan_adjoint(o, alpha, steps, oob, every=5)
    with Lattice(cfg) as lat:
        nbytes = lat.state_bytes
    cfg.checkpoint_budget_bytes = nbytes * 4
    with Lattice(cfg) as lat:
        lat.set_design(alpha)
        lat.initial_populations()
        J, ab, lam0, _ = lat.gradient(steps, ob, want_lambda0=True)
    assert abs(J - J_want) <= 1e-12 * abs(J_want)
    assert np.abs(ab - ab_want).max() <= 1e-11 * np.abs(ab_want).max()
    assert np.abs(lam0 - lam_want).max() <= 1e-11 * np.abs(lam_want).max()


def test_step_jvp_dot_product_with_stationary_ports():
    """<lambda, J v> = <J^T lambda, v>: the tangent runs the rules on dual numbers, the
    transpose is analytic."""
    shape, periodic, coll, inlet, outlet, accel, tangential, inset, with_solid = CASES[2]
    (cfg, prob), _ = _case(shape, periodic, coll, inlet, outlet, accel, tangential, inset, with_solid)
    o = orc.Oracle(prob)
    alpha = random_alpha(shape, 4)
    f = o.run(alpha, 3)
    rng = np.random.default_rng(2)
    v = rng.standard_normal(o.n * 19) * 1e-3
    lam = rng.standard_normal(o.n * 19)
    with Lattice(cfg) as lat:
        lat.set_design(alpha)
        lat.set_state(f)
        jv = lat.interval_jvp(v)
        lam_prev, _ = lat.interval_vjp(lam)
    lhs, rhs = float(lam @ jv), float(lam_prev @ v)
    assert abs(lhs - rhs) <= 1e-13 * float(np.abs(lam).sum() * np.abs(jv).max())


def test_mass_drift_includes_stationary_port_exchange_f64():
    shape = (12, 8, 8)
    cfg, _ = make_pair(shape, (False, False, False), collision="bgk_guo", precision="f64",
                       ports=stationary_ports(shape, 0.03, 0.998, inset=(1, 1)))
    with Lattice(cfg) as lat:
        lat.set_design(random_alpha(shape, 3))
        lat.initial_populations()
        _, st = lat.history(80)
    assert st.mass_drift <= 1e-13, st.mass_drift


@pytest.mark.parametrize("precision,tol", [("f32", 2e-5), ("f16s", 2e-3), ("bf16s", 1.5e-2)])
def test_forward_reduced_precision(precision, tol):
    shape, periodic, coll, inlet, outlet, accel, tangential, inset, with_solid = CASES[1]
    (cfg, prob), _ = _case(shape, periodic, coll, inlet, outlet, accel, tangential, inset, with_solid,
                           precision=precision)
    o = orc.Oracle(prob)
    alpha = random_alpha(shape, 5)
    want = o.run(alpha, 14)
    with Lattice(cfg) as lat:
        lat.set_design(alpha)
        lat.initial_populations()
        lat.interval(14)
        got = lat.state()
    _, w, _ = orc.lattice(19)
    dev = want.reshape(-1, 19) - w
    err = np.abs(got.reshape(-1, 19) - want.reshape(-1, 19)).max() / np.abs(dev).max()
    assert err < tol, err
