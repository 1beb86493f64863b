"""The golden vectors (tests/golden/lbm_small.npz, written by tests/golden/make_golden.py).

CPU: the oracle still reproduces every vector bit for bit.  GPU: the HIP library's FP64
path reproduces the forward states bit for bit and the gradients to 1e-11, from the
fixture data alone.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from oracle import oracle as orc
from oracle import revolve as orc_rev
from oracle.thermal import CoupledObjective, Thermal, ThermalOracle
from tests.golden.cases import CASES, build, gpu_config
from tests.golden.make_golden import encode_actions

GOLD = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "lbm_small.npz"))


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_reproduces_golden(name):
    spec = CASES[name]
    prob, alpha, phi = build(spec)
    o = orc.Oracle(prob)
    if spec["kind"] == "forward":
        assert np.array_equal(o.run(alpha, spec["steps"]), GOLD[f"{name}/state"])
    elif spec["kind"] == "gradient":
        ob = orc.Objective(kind="pressure_drop", window=tuple(spec["window"]), weight=spec["weight"], axis=0,
                           plane_a=1, plane_b=prob.shape[0] - 2)
        J, ab, lam, final = orc.scan_adjoint(o, alpha, spec["steps"], ob)
        assert J == GOLD[f"{name}/J"][0]
        assert np.array_equal(ab, GOLD[f"{name}/alpha_bar"])
        assert np.array_equal(lam, GOLD[f"{name}/lambda0"])
        assert np.array_equal(final, GOLD[f"{name}/state"])
    else:
        t = ThermalOracle(o, Thermal(**spec["thermal"]))
        f, g = t.run(alpha, phi, spec["steps"])
        assert np.array_equal(f, GOLD[f"{name}/state"]) and np.array_equal(g, GOLD[f"{name}/tstate"])
        ob = CoupledObjective(window=tuple(spec["window"]), weight=spec["weight"], w_T=1.0, w_dp=0.1, plane_a=1,
                              plane_b=prob.shape[0] - 2)
        J, ab, pb, _, _ = t.scan_adjoint(ob, alpha, phi, spec["steps"])
        assert J == GOLD[f"{name}/J"][0]
        assert np.array_equal(ab, GOLD[f"{name}/alpha_bar"]) and np.array_equal(pb, GOLD[f"{name}/phi_bar"])


def test_oracle_revolve_reproduces_golden():
    acts, rstart, fwd, peak = orc_rev.binomial_schedule(30, 4)
    assert np.array_equal(encode_actions(acts), GOLD["revolve_30_4/actions"])
    assert [rstart, fwd, peak] == GOLD["revolve_30_4/info"].tolist()


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_matches_golden(name):
    from implexity_amd.lbm import Lattice, Objective

    spec = CASES[name]
    _, alpha, phi = build(spec)
    cfg = gpu_config(spec)
    nx = spec["shape"][0]
    with Lattice(cfg) as lat:
        lat.set_design(alpha)
        if phi is not None:
            lat.set_thermal_design(phi)
        lat.initial_populations()
        if spec["kind"] == "forward":
            lat.interval(spec["steps"])
            assert np.array_equal(lat.state(), GOLD[f"{name}/state"])
            return
        lo, hi = spec["window"]
        if spec["kind"] == "gradient":
            ob = Objective(kind="pressure_drop", axis=0, plane_a=1, plane_b=nx - 2, t_begin=lo, t_end=hi,
                           weight=spec["weight"])
        else:
            ob = Objective(kind="thermal", axis=0, plane_a=1, plane_b=nx - 2, t_begin=lo, t_end=hi,
                           weight=spec["weight"], vector_weights=(1.0, 0.1, 0, 0, 0, 0))
        J, ab, lam, _ = lat.gradient(spec["steps"], ob, want_lambda0=spec["kind"] == "gradient")
        J_want = GOLD[f"{name}/J"][0]
        assert abs(J - J_want) <= 1e-12 * abs(J_want)
        want = GOLD[f"{name}/alpha_bar"]
        assert np.abs(ab - want).max() <= 1e-11 * np.abs(want).max()
        if spec["kind"] == "gradient":
            want = GOLD[f"{name}/lambda0"]
            assert np.abs(lam - want).max() <= 1e-11 * np.abs(want).max()
        else:
            want = GOLD[f"{name}/phi_bar"]
            assert np.abs(lat.thermal_design_bar() - want).max() <= 1e-11 * np.abs(want).max()
        # the forward state after the sweep is the initial one again; replay it
        lat.initial_populations()
        lat.interval(spec["steps"])
        assert np.array_equal(lat.state(), GOLD[f"{name}/state"])
        if spec["kind"] == "thermal":
            assert np.array_equal(lat.thermal_state(), GOLD[f"{name}/tstate"])


@pytest.mark.gpu
def test_gpu_revolve_matches_golden():
    from implexity_amd.lbm import BinomialSchedule

    s = BinomialSchedule(30, 4)
    assert np.array_equal(encode_actions(s.actions), GOLD["revolve_30_4/actions"])
