"""GPU parity of the regularized stationary port rules ("regularized_known_stress",
reconstruct_regularized, crates/implexity-physics-lbm/src/ports.rs:455-482, with the
stress fit of known_stress_operator, :409-440) through the C ABI against the FP64 oracle.

The rule reads the port cell's own pre-port populations and is not affine, so its
transpose needs them after the port pass has overwritten the cell: the library keeps
them beside every state buffer.  The gradient cases run with a checkpoint budget of four
state buffers, so the reverse sweep reads them from recomputed states and snapshots.

Bars: the FP64 forward bit-exact (the stress operator is the correctly rounded exact
operator on both sides), the step VJP within 1e-12 of the oracle's apply_ports_vjp
(:582-621; analytic transposes on the device, the dual-number cell Jacobian in the
oracle), the gradient within 1e-11, the tangent/transpose dot product within 1e-13, FP64
mass drift with the port exchange at rounding, reduced precision within the usual bounds
of the largest population deviation (forward), the FP32 gradient within 2e-5 and the
FP16-storage gradient within 2e-2 relative L2.
"""
from __future__ import annotations

import numpy as np
import pytest

from implexity_amd.lbm import ContractError, Lattice, Objective
from oracle import oracle as orc
from tests.lbm_cases import make_pair, random_alpha, random_solid, stationary_ports

pytestmark = pytest.mark.gpu


def z_ports(shape, u_in=0.02, rho_out=0.997):
    """Regularized ports on the z faces of an x/y-periodic box."""
    m_in = np.zeros(shape, bool)
    m_out = np.zeros(shape, bool)
    m_in[:, :, 0] = True
    m_out[:, :, -1] = True
    return [{"face": "zmin", "kind": "velocity_regularized", "value": (0.003, -0.002, u_in), "mask": m_in},
            {"face": "zmax", "kind": "pressure_regularized", "value": (rho_out, 0.001, 0.0), "mask": m_out}]


CASES = {
    # shape, periodic, collision, ports, with_solid
    "vreg_neighbor": ((8, 6, 5), (False, False, False), "bgk_guo",
                      stationary_ports((8, 6, 5), 0.025, 0.997, (0.0, 0.0), "velocity_regularized",
                                       "pressure_neighbor", (1, 1)), False),
    "moment_preg": ((8, 6, 8), (False, False, True), "trt",
                    stationary_ports((8, 6, 8), 0.02, 0.998, (0.002, -0.001), "velocity_moment",
                                     "pressure_regularized", (1, 0)), True),
    "both_mrt": ((7, 5, 6), (False, False, False), "mrt",
                 stationary_ports((7, 5, 6), 0.03, 1.002, (0.0, 0.0), "velocity_regularized",
                                  "pressure_regularized", (1, 1)), True),
    "z_faces": ((6, 6, 8), (True, True, False), "bgk", z_ports((6, 6, 8)), False),
}


def _case(name, precision="f64"):
    shape, periodic, coll, ports, with_solid = CASES[name]
    solid = None
    if with_solid:
        solid = random_solid(shape, 11).copy()
        solid[:2] = False
        solid[-2:] = False
    cfg, prob = make_pair(shape, periodic, Q=19, collision=coll, precision=precision, solid=solid, ports=ports)
    return cfg, prob, solid


@pytest.mark.parametrize("name", list(CASES))
def test_forward_f64_bit_exact(name):
    cfg, prob, _ = _case(name)
    o = orc.Oracle(prob)
    alpha = random_alpha(prob.shape, 5)
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


@pytest.mark.parametrize("name", list(CASES))
def test_step_vjp_f64(name):
    cfg, prob, solid = _case(name)
    o = orc.Oracle(prob)
    alpha = random_alpha(prob.shape, 9)
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


def _gradient(name, precision="f64", steps=20):
    cfg, prob, _ = _case(name, precision)
    shape = prob.shape
    o = orc.Oracle(prob)
    alpha = random_alpha(shape, 21)
    axis = 2 if name == "z_faces" else 0
    ob = Objective(kind="pressure_drop", axis=axis, plane_a=1, plane_b=shape[axis] - 2, t_begin=steps // 2,
                   t_end=steps, weight=1.0 / (steps - steps // 2 + 1))
    oob = orc.Objective(kind=ob.kind, window=(ob.t_begin, ob.t_end), weight=ob.weight, axis=ob.axis,
                        plane_a=ob.plane_a, plane_b=ob.plane_b)
    J_want, ab_want, lam_want, _ = orc.scan_adjoint(o, alpha, steps, oob, every=5)
    with Lattice(cfg) as lat:
        nbytes = lat.state_bytes
    cfg.checkpoint_budget_bytes = nbytes * 4
    with Lattice(cfg) as lat:
        lat.set_design(alpha)
        lat.initial_populations()
        J, ab, lam0, st = lat.gradient(steps, ob, want_lambda0=True)
    return (J, ab, lam0, st), (J_want, ab_want, lam_want)


@pytest.mark.parametrize("name", ["vreg_neighbor", "both_mrt", "z_faces"])
def test_gradient_f64_matches_oracle(name):
    (J, ab, lam0, _), (J_want, ab_want, lam_want) = _gradient(name)
    assert abs(J - J_want) <= 1e-12 * abs(J_want)
    assert np.abs(ab - ab_want).max() <= 1e-11 * np.abs(ab_want).max()
    assert np.abs(lam0 - lam_want).max() <= 1e-11 * np.abs(lam_want).max()


def test_gradient_f32_matches_oracle():
    (J, ab, _, st), (J_want, ab_want, _) = _gradient("both_mrt", "f32")
    err = np.linalg.norm(ab - ab_want) / np.linalg.norm(ab_want)
    assert err <= 2e-5, err
    assert abs(J - J_want) <= 2e-5 * abs(J_want)
    assert np.isfinite(st.mass_drift)


@pytest.mark.parametrize("name", ["moment_preg", "z_faces"])
def test_gradient_f16s_matches_oracle(name):
    """FP16 storage (the pre-port populations kept in FP16 too) with scaled FP16
    cotangents, against the FP64 oracle."""
    (J, ab, _, st), (J_want, ab_want, _) = _gradient(name, "f16s")
    err = np.linalg.norm(ab - ab_want) / np.linalg.norm(ab_want)
    assert err <= 2e-2, err
    assert abs(J - J_want) <= 2e-2 * abs(J_want)
    assert np.isfinite(st.mass_drift)


def test_step_jvp_dot_product_with_regularized_ports():
    """<lambda, J v> = <J^T lambda, v>: the tangent runs the rules on dual numbers, the
    transpose analytically at the kept pre-port populations."""
    cfg, prob, _ = _case("both_mrt")
    o = orc.Oracle(prob)
    alpha = random_alpha(prob.shape, 4)
    f = o.run(alpha, 3)
    rng = np.random.default_rng(2)
    v = rng.standard_normal(o.n * 19) * 1e-3
    lam = rng.standard_normal(o.n * 19)
    solid = prob.solid.reshape(-1)
    v.reshape(-1, 19)[solid] = 0.0
    lam.reshape(-1, 19)[solid] = 0.0
    with Lattice(cfg) as lat:
        lat.set_design(alpha)
        lat.set_state(f)
        jv = lat.interval_jvp(v)
        lam_prev, _ = lat.interval_vjp(lam)
    lhs, rhs = float(lam @ jv), float(lam_prev @ v)
    assert abs(lhs - rhs) <= 1e-13 * float(np.abs(lam).sum() * np.abs(jv).max())


def test_mass_drift_includes_regularized_port_exchange_f64():
    shape = (12, 8, 8)
    cfg, _ = make_pair(shape, (False, False, False), collision="bgk_guo", precision="f64",
                       ports=stationary_ports(shape, 0.03, 0.998, (0.0, 0.0), "velocity_regularized",
                                              "pressure_regularized", (1, 1)))
    with Lattice(cfg) as lat:
        lat.set_design(random_alpha(shape, 3))
        lat.initial_populations()
        _, st = lat.history(80)
    assert st.mass_drift <= 1e-13, st.mass_drift


@pytest.mark.parametrize("precision,tol", [("f32", 2e-5), ("f16s", 2e-3), ("bf16s", 1.5e-2)])
def test_forward_reduced_precision(precision, tol):
    cfg, prob, _ = _case("both_mrt", precision)
    o = orc.Oracle(prob)
    alpha = random_alpha(prob.shape, 5)
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


def test_regularized_ports_refused_with_acceleration():
    shape = (8, 6, 5)
    cfg, _ = make_pair(shape, (False, False, False), collision="bgk_guo", accel=(1e-5, 0.0, 0.0),
                       ports=stationary_ports(shape, 0.02, 0.998, (0.0, 0.0), "velocity_regularized",
                                              "pressure_neighbor", (1, 1)))
    with pytest.raises(ContractError, match="regularized_known_stress ports require zero acceleration"):
        Lattice(cfg).close()
