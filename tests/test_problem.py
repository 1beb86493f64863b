"""LbmProblem::normalize (crates/implexity-physics-lbm/src/solver.rs:414-550) and
normalize_ports (ports.rs:107-270) on the host: the reference's problem template
(provider.rs:74-91) normalises, every contract check refuses with the reference's message
when its one field is broken, and the derived lattice quantities take their closed forms.
"""
from __future__ import annotations

import copy

import numpy as np
import pytest

from implexity_amd import design, problem
from implexity_amd.lbm import ContractError

SHAPE = (4, 6, 4)


def template():
    """provider.rs:74-91 problem_template."""
    n = SHAPE
    return {
        "shape": list(SHAPE), "spacing_m": 0.001, "step_s": 0.1, "step_count": 10, "origin_m": [0.0, 0.0, 0.0],
        "density_kg_m3": 1000.0, "kinematic_viscosity_m2_s": 1e-6, "acceleration_m_s2": [1e-5, 0.0, 0.0],
        "periodic_axes": [True, False, True], "solid_mask": np.zeros(n, bool).tolist(),
        "design_region": np.ones(n, bool).tolist(), "fixed_design": np.ones(n).tolist(), "drag_max_per_s": 10.0,
        "drag_shape": 0.1, "mach_limit": 0.1, "history_byte_budget": 16_000_000, "ports": [],
        "topology_map": {"filter_radius_m": 0.0015, "projection_beta": 2.0, "projection_eta": 0.5},
    }


def test_template_normalises_with_closed_forms():
    p = problem.normalize(template())
    assert p.shape == SHAPE and p.step_count == 10
    assert abs(p.tau - 0.8) < 1e-15  # 1/2 + 3 nu dt / dx^2 = 1/2 + 3e-6 * 0.1 / 1e-6
    a = p.acceleration_lattice
    assert abs(a[0] - 1e-4) < 1e-19 and a[1] == 0.0 and a[2] == 0.0  # a dt^2 / dx
    assert p.alpha_max == 0.1 * 10.0
    assert p.vector_scales() == (0.001 / 0.1, 1000.0, 0.1, 0.001 ** 3)
    assert p.topology_map == problem.TopologyMap(0.0015, 2.0, 0.5)
    phi = np.linspace(0, 1, 7)
    assert np.array_equal(p.resistance(phi), design.resistance(phi, 1.0, 0.1))
    assert p.resistance(np.array([1.0]))[0] == 0.0 and p.resistance(np.array([0.0]))[0] == 1.0
    cfg = p.lattice_config("f32")
    assert cfg.collision == "bgk_guo" and cfg.lattice == 19 and cfg.mach_limit == 0.1
    assert cfg.periodic == (True, False, True) and cfg.solid is None
    assert cfg.checkpoint_budget_bytes == 16_000_000
    assert cfg.tau == p.tau  # handed over as tau, not as a reciprocal


def _broken(path, value):
    p = template()
    if path is None:
        return value(p)
    d = p
    for k in path[:-1]:
        d = d[k]
    if value is KeyError:
        del d[path[-1]]
    else:
        d[path[-1]] = value
    return p


MISMATCH = [
    (("spacing_m",), KeyError, "exactly these problem keys"),
    (("extra",), 1, "exactly these problem keys"),
    (("shape",), [4, 1, 4], "three integers >= 2"),
    (("shape",), [4.0, 6, 4], "three integers >= 2"),
    (("spacing_m",), 0.0, "spacing_m must be a finite positive scalar"),
    (("step_s",), float("inf"), "step_s must be a finite positive scalar"),
    (("density_kg_m3",), True, "density_kg_m3 must be a finite positive scalar"),
    (("mach_limit",), 0.25, "not a compressible gas solver"),
    (("step_count",), 0, r"step_count must be an integer in \[1,1000000\]"),
    (("step_count",), 10.0, r"step_count must be an integer"),
    (("history_byte_budget",), (10 + 1) * 96 * 19 * 8 - 1, "exceeds explicit byte budget"),
    (("solid_mask",), np.zeros((4, 6, 3), bool).tolist(), "solid_mask must be a shape-matched boolean array"),
    (("design_region",), np.ones(SHAPE).tolist(), "design_region must be a shape-matched boolean array"),
    (("fixed_design",), np.full(SHAPE, 1.5).tolist(), r"fixed_design must be finite and in \[0,1\]"),
    (("solid_mask",), np.ones(SHAPE, bool).tolist(), "Need fluid cells"),
    (("acceleration_m_s2",), [0.0, float("nan"), 0.0], "three finite components"),
    (("periodic_axes",), [1, 0, 1], "three booleans"),
    (("kinematic_viscosity_m2_s",), 1e-9, "relaxation time"),
    (("topology_map", "filter_radius_m"), 0.0041, r"\[0,4\*spacing_m\]"),
    (("topology_map", "projection_eta"), 1.0, r"eta in \(0,1\)"),
    (("topology_map", "extra"), 0.0, "requires filter_radius_m"),
    (("geometry_source",), {}, "geometry_source"),
    (("field_registration",), {}, "field_registration"),
]


@pytest.mark.parametrize("path,value,msg", MISMATCH)
def test_contract_refusals(path, value, msg):
    with pytest.raises(ContractError, match=msg):
        problem.normalize(_broken(path, value))


def test_budget_at_the_bound_is_accepted():
    p = template()
    p["history_byte_budget"] = (10 + 1) * 96 * 19 * 8
    problem.normalize(p)


def test_solid_designable_cell_refused():
    p = template()
    s = np.zeros(SHAPE, bool)
    s[1, 2, 1] = True
    p["solid_mask"] = s.tolist()
    with pytest.raises(ContractError, match="cannot be designable"):
        problem.normalize(p)
    d = np.ones(SHAPE, bool)
    d[1, 2, 1] = False
    p["design_region"] = d.tolist()
    assert problem.normalize(p).solid_mask[1, 2, 1]


# --- ports -------------------------------------------------------------------------------
def _channel(recon_in=None, recon_out="neighbor_non_equilibrium", gauge=2e-4):
    """The template turned into a y channel: ports on ymin / ymax (x and z periodic), the
    design region kept off the port planes and the outlet's inward plane."""
    p = template()
    p["acceleration_m_s2"] = [0.0, 0.0, 0.0]
    d = np.ones(SHAPE, bool)
    d[:, :2] = False
    d[:, -2:] = False
    p["design_region"] = d.tolist()
    m_in = np.zeros(SHAPE, bool)
    m_in[:, 0] = True
    m_out = np.zeros(SHAPE, bool)
    m_out[:, -1] = True
    inlet = {"id": "in", "kind": "velocity", "face": "ymin", "mask": m_in.tolist(), "velocity_m_s": [0.0, 1e-4, 0.0]}
    if recon_in:
        inlet["reconstruction"] = recon_in
    outlet = {"id": "out", "kind": "pressure", "face": "ymax", "mask": m_out.tolist(),
              "velocity_m_s": [0.0, 0.0, 0.0], "gauge_pressure_Pa": gauge, "reconstruction": recon_out}
    p["ports"] = [inlet, outlet]
    return p


def test_ports_map_to_stationary_kinds():
    p = problem.normalize(_channel())
    inlet, outlet = p.ports
    assert inlet.face == "ymin" and inlet.kind == "velocity_moment"
    assert inlet.value == (0.0, 1e-4 * 0.1 / 0.001, 0.0)
    assert outlet.face == "ymax" and outlet.kind == "pressure_neighbor"
    unit = 0.001 / 0.1
    rho = 1.0 + 2e-4 * 3.0 / (1000.0 * (unit * unit))  # p = c_s^2 (rho - 1) rho_phys unit^2
    assert outlet.value == (rho, 0.0, 0.0)
    assert outlet.mask[:, -1].all() and outlet.mask.sum() == SHAPE[0] * SHAPE[2]
    p2 = problem.normalize(_channel(recon_out="legacy_moment"))
    assert p2.ports[1].kind == "pressure_moment"
    p3 = problem.normalize(_channel(recon_in="regularized_known_stress", recon_out="regularized_known_stress"))
    assert [q.kind for q in p3.ports] == ["velocity_regularized", "pressure_regularized"]
    assert p3.ports[0].value == inlet.value and p3.ports[1].value == (rho, 0.0, 0.0)


def test_regularized_ports_refused_with_acceleration():
    p = _channel(recon_in="regularized_known_stress", recon_out="legacy_moment")
    p["acceleration_m_s2"] = [0.0, 0.0, 1e-3]
    with pytest.raises(ContractError, match="regularized_known_stress ports require zero acceleration"):
        problem.normalize(p)


PORT_MISMATCH = [
    (lambda p: p["ports"][0].update(kind="inlet"), "Port kind must be velocity or pressure"),
    (lambda p: p["ports"][0].pop("velocity_m_s"), "exact port fields"),
    (lambda p: p["ports"][1].update(id="in"), "Unique port id"),
    (lambda p: p["ports"][0].update(reconstruction="zou_he"), "Port reconstruction must be"),
    (lambda p: p["ports"][0].update(reconstruction="neighbor_non_equilibrium"), "requires a pressure port"),
    (lambda p: p["ports"][0].update(face="xmin"), "A periodic face cannot be a port"),
    (lambda p: p["ports"][0].update(face="left"), "one Cartesian face"),
    (lambda p: p["ports"][0].update(mask=np.zeros(SHAPE, bool).tolist()), "Nonempty shape-matched"),
    (lambda p: p["ports"][1].update(face="ymin"), "must lie on its face"),
    (lambda p: p["ports"][0].update(velocity_m_s=[0.0, 1e-2, 0.0]), "exceeds Mach limit"),
    (lambda p: p["ports"][1].update(velocity_m_s=[0.0, 1e-5, 0.0]), "normal component must be zero"),
    (lambda p: p["ports"][1].update(gauge_pressure_Pa=1.0), "within 10% of reference"),
    (lambda p: p["ports"][1].update(gauge_pressure_Pa="1"), "Finite scalar gauge pressure"),
]


@pytest.mark.parametrize("mutate,msg", PORT_MISMATCH)
def test_port_refusals(mutate, msg):
    p = _channel()
    mutate(p)
    with pytest.raises(ContractError, match=msg):
        problem.normalize(p)


def test_port_cells_must_be_protected_fluid():
    p = _channel()
    d = np.asarray(p["design_region"])
    d[0, 0, 0] = True
    p["design_region"] = d.tolist()
    with pytest.raises(ContractError, match="protected, unobstructed"):
        problem.normalize(p)
    p = _channel()
    d = np.asarray(p["design_region"])
    d[0, -2, 0] = True  # the outlet's inward plane
    p["design_region"] = d.tolist()
    with pytest.raises(ContractError, match="inward fluid cells"):
        problem.normalize(p)


def test_port_corners_on_closed_faces_refused():
    p = _channel()
    p["periodic_axes"] = [False, False, True]
    with pytest.raises(ContractError, match="corner closure"):
        problem.normalize(p)


def test_neighbor_rule_cannot_sample_another_port():
    p = _channel()
    p["shape"] = [4, 2, 4]
    for k in ("solid_mask", "design_region", "fixed_design"):
        p[k] = np.asarray(p[k])[:, :2].tolist()
    p["history_byte_budget"] = 10 ** 8
    p["design_region"] = np.zeros((4, 2, 4), bool).tolist()
    for port, j in zip(p["ports"], (0, 1)):
        m = np.zeros((4, 2, 4), bool)
        m[:, j] = True
        port["mask"] = m.tolist()
    with pytest.raises(ContractError, match="cannot sample another port"):
        problem.normalize(p)


def test_normalize_does_not_modify_input():
    p = _channel()
    q = copy.deepcopy(p)
    problem.normalize(p)
    assert p == q
