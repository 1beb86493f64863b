"""GPU tangent of one step (lbm_step_jvp, the drop-in for ScanStep::jvp of
crates/implexity-ad/src/scan.rs:46-80) through the C ABI, in FP64:

* against central differences of the oracle's step (solver.rs:603-614 restated) in the
  same direction (state and design together), to the accuracy of h = 1e-6;
* the dot-product identity <lambda, J v> = <J^T lambda, v> with the device's own step
  VJP (lbm_step_vjp) on the other side, to rounding -- two independent device codes,
  forward mode through the step kernels and the hand-written transposed kernels.

The operators: BGK, TRT, MRT (default and non-default rates), regularized, BGK-Guo,
cumulant, the Smagorinsky (psm.rs:161-197) and WALE (moving/wale.rs) closures on BGK
and TRT, sponge layers (moving/sponge.rs) with a steady and a travelling-wave far field,
and the occupancy couplings (the Psm law on D3Q19; PsmSuperposition and the moving
Brinkman law on both lattices).
"""
from __future__ import annotations

import numpy as np
import pytest

from implexity_amd.lbm import ContractError, Lattice
from oracle import oracle as orc
from tests.lbm_cases import channel_ports, make_pair, random_alpha, random_solid

pytestmark = pytest.mark.gpu

JVP_CASES = [
    # shape, periodic, Q, collision, with_solid, with_ports, (bulk_rate, even_rate)
    ((7, 6, 5), (True, False, False), 19, "trt", True, False, (None, None)),
    ((7, 6, 5), (False, False, False), 19, "trt", False, True, (None, None)),
    ((7, 6, 5), (True, False, False), 19, "bgk", True, False, (None, None)),
    ((7, 6, 5), (True, False, False), 19, "bgk_guo", True, False, (None, None)),
    ((6, 5, 6), (True, True, False), 19, "mrt", True, False, (1.1, 0.7)),
    ((6, 5, 6), (False, False, False), 19, "mrt", False, True, (None, None)),
    ((6, 5, 6), (True, True, False), 19, "regularized", True, False, (None, None)),
    ((6, 5, 4), (True, False, False), 27, "trt", True, False, (None, None)),
    ((6, 5, 4), (True, False, False), 27, "mrt", False, False, (1.3, 0.8)),
    ((6, 5, 4), (True, False, False), 27, "cumulant", True, False, (0.8, None)),
    ((7, 5, 4), (False, False, False), 27, "cumulant", False, True, (None, None)),
    # the Smagorinsky closure (rates = (les constant, norm floor) for these)
    ((7, 6, 5), (True, False, False), 19, "trt_les", True, False, (0.15, 1e-6)),
    ((7, 6, 5), (False, False, False), 19, "bgk_les", False, True, (0.15, 1e-6)),
    ((6, 5, 4), (True, False, False), 27, "trt_les", False, False, (0.15, 1e-6)),
    # the WALE closure (moving/wale.rs; rates = (WALE constant, floor))
    ((7, 6, 5), (True, False, False), 19, "trt_wale", True, False, (0.5, 1e-12)),
    ((7, 6, 5), (False, False, False), 19, "bgk_wale", False, True, (0.5, 1e-12)),
    ((6, 5, 4), (True, False, False), 27, "trt_wale", False, False, (0.5, 1e-12)),
    # sponge layers (moving/sponge.rs): an affine blend after the collision
    ((8, 6, 5), (False, True, True), 19, "trt_sponge", True, False, (None, None)),
    ((7, 6, 6), (False, False, False), 19, "bgk_sponge", False, True, (None, None)),
    ((6, 5, 6), (True, True, False), 27, "trt_sponge", False, False, (None, None)),
]

_WAVE = orc.Wave((0.01, -0.004, 0.002), (0.6, 0.8, 0.0), 0.3, 0.02, phase=0.4, ramp=6.0, origin=(1.0, 2.0, 0.0))


def _sponges(periodic, with_ports):
    if with_ports:  # the ports hold the x faces
        return [orc.Sponge("ymax", 2, 0.4, density=1.002, velocity=(0.01, 0.0, 0.0), wave=_WAVE)]
    if not periodic[0]:  # x faces, one with a travelling wave
        return [orc.Sponge("xmin", 2, 0.5, density=1.0, velocity=(0.02, 0.0, 0.0), wave=_WAVE),
                orc.Sponge("xmax", 2, 0.4, density=1.002)]
    return [orc.Sponge("zmax", 2, 0.35, density=1.01, velocity=(0.01, 0.0, 0.0), wave=_WAVE)]


def _setup(shape, periodic, Q, coll, with_solid, with_ports, rates, omega=1.4):
    solid = random_solid(shape, 13) if with_solid else None
    ports = channel_ports(shape, 0.03) if with_ports else ()
    if with_ports and solid is not None:
        solid = solid.copy()
        solid[:2] = False
        solid[-2:] = False
    les, wale, sponges = (0.0, 0.0), (0.0, 1e-12), ()
    if coll.endswith("_les"):  # the closure on BGK / TRT
        coll, les, rates, omega = coll[:-4], rates, (None, None), 1.8
    if coll.endswith("_wale"):
        coll, wale, rates, omega = coll[:-5], rates, (None, None), 1.8
    if coll.endswith("_sponge"):
        coll, sponges = coll[:-7], _sponges(periodic, with_ports)
    bulk = rates[0] if rates[0] is not None else (1.0 if coll == "cumulant" else float("nan"))
    even = rates[1] if rates[1] is not None else float("nan")
    cfg, prob = make_pair(shape, periodic, Q=Q, collision=coll, precision="f64", omega=omega,
                          accel=(0, 0, 0) if with_ports else (1e-4, -3e-5, 0), solid=solid, ports=ports,
                          bulk_rate=bulk, even_rate=even, les=les, wale=wale, sponges=sponges)
    return cfg, prob, solid


def _directions(o, Q, solid, seed):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(o.n * Q) * 1e-3
    a = rng.standard_normal(o.n) * 1e-2
    lam = rng.standard_normal(o.n * Q)
    if solid is not None:
        s = solid.reshape(-1)
        v.reshape(-1, Q)[s] = 0.0
        lam.reshape(-1, Q)[s] = 0.0
    return v, a, lam


@pytest.mark.parametrize("shape,periodic,Q,coll,with_solid,with_ports,rates", JVP_CASES)
def test_step_jvp_central_differences(shape, periodic, Q, coll, with_solid, with_ports, rates):
    cfg, prob, solid = _setup(shape, periodic, Q, coll, with_solid, with_ports, rates)
    o = orc.Oracle(prob)
    alpha = random_alpha(shape, 3)
    f = o.run(alpha, 4)
    v, a, _ = _directions(o, Q, solid, 5)
    h = 1e-6
    fd = (o.interval(alpha + h * a, f + h * v) - o.interval(alpha - h * a, f - h * v)) / (2 * h)
    with Lattice(cfg) as lat:
        lat.set_design(alpha)
        lat.set_state(f)
        got = lat.interval_jvp(v, a)
        after = lat.state()  # the state is not advanced
    fluid = np.ones(o.n, bool) if solid is None else ~solid.reshape(-1)
    g, w = got.reshape(-1, Q)[fluid], fd.reshape(-1, Q)[fluid]
    assert np.abs(g - w).max() <= 1e-7 * np.abs(w).max(), np.abs(g - w).max() / np.abs(w).max()
    assert np.abs(after.reshape(-1, Q)[fluid] - f.reshape(-1, Q)[fluid]).max() == 0.0


@pytest.mark.parametrize("shape,periodic,Q,coll,with_solid,with_ports,rates", JVP_CASES)
def test_step_jvp_vjp_dot_product(shape, periodic, Q, coll, with_solid, with_ports, rates):
    """<lambda, J_x v + J_alpha a> = <J_x^T lambda, v> + <J_alpha^T lambda, a>, both sides
    on the device."""
    cfg, prob, solid = _setup(shape, periodic, Q, coll, with_solid, with_ports, rates)
    o = orc.Oracle(prob)
    alpha = random_alpha(shape, 7)
    f = o.run(alpha, 3)
    v, a, lam = _directions(o, Q, solid, 9)
    with Lattice(cfg) as lat:
        lat.set_design(alpha)
        lat.set_state(f)
        jv = lat.interval_jvp(v, a)
        lam_prev, abar = lat.interval_vjp(lam)
    fluid = np.ones(o.n, bool) if solid is None else ~solid.reshape(-1)
    lhs = float(np.dot(lam.reshape(-1, Q)[fluid].ravel(), jv.reshape(-1, Q)[fluid].ravel()))
    rhs = float(np.dot(lam_prev.reshape(-1, Q)[fluid].ravel(), v.reshape(-1, Q)[fluid].ravel()) + abar @ a)
    scale = float(np.abs(lam).sum() * np.abs(jv).max())
    assert abs(lhs - rhs) <= 1e-13 * scale, (lhs, rhs)


COUPLED = [
    # shape, periodic, Q, collision, law, with solid, with ports (moving/psm.rs:213-266)
    ((7, 6, 5), (True, False, False), 19, "trt", "psm", True, False),
    ((9, 6, 6), (False, False, False), 19, "bgk", "psm_superposition", False, True),
    ((7, 6, 5), (True, False, False), 19, "trt", "brinkman_occupancy", True, False),
    ((6, 5, 4), (True, False, True), 27, "trt", "psm_superposition", True, False),
    ((6, 5, 4), (True, False, True), 27, "bgk", "brinkman_occupancy", False, False),
]


@pytest.mark.parametrize("shape,periodic,Q,coll,law,with_solid,with_ports", COUPLED)
def test_step_jvp_occupancy_couplings(shape, periodic, Q, coll, law, with_solid, with_ports):
    """The occupancy d as the design and the solid momentum as a constant: central
    differences of the oracle's step and the dot product with the device's step VJP
    (the direction in d only where the coupling is on, d > 0)."""
    from tests.test_gpu_psm import _case as psm_case
    from tests.test_gpu_psm import _lattice as psm_lattice

    cfg, prob, solid, d, m = psm_case(shape, periodic, Q, coll, law, with_solid, with_ports)
    o = orc.Oracle(prob)
    f = o.run(d, 4)
    v, a, lam = _directions(o, Q, solid, 11)
    a = a * (d > 0)
    h = 1e-6
    fd = (o.interval(d + h * a, f + h * v) - o.interval(d - h * a, f - h * v)) / (2 * h)
    with psm_lattice(cfg, d, m) as lat:
        lat.set_state(f)
        jv = lat.interval_jvp(v, a)
        lam_prev, dbar = lat.interval_vjp(lam)
    fluid = np.ones(o.n, bool) if solid is None else ~solid.reshape(-1)
    g, w = jv.reshape(-1, Q)[fluid], fd.reshape(-1, Q)[fluid]
    assert np.abs(g - w).max() <= 1e-7 * np.abs(w).max(), np.abs(g - w).max() / np.abs(w).max()
    lhs = float(np.dot(lam.reshape(-1, Q)[fluid].ravel(), g.ravel()))
    rhs = float(np.dot(lam_prev.reshape(-1, Q)[fluid].ravel(), v.reshape(-1, Q)[fluid].ravel()) + dbar @ a)
    assert abs(lhs - rhs) <= 1e-13 * float(np.abs(lam).sum() * np.abs(jv).max()), (lhs, rhs)


def test_step_jvp_refuses_d3q27_psm_law():
    cfg, _ = make_pair((6, 5, 4), (True, False, False), Q=27, precision="f64", coupling=("psm", 0.1))
    with Lattice(cfg) as lat:
        with pytest.raises(ContractError, match="lbm_step_jvp covers"):
            lat.interval_jvp(np.zeros(lat.n * 27))
