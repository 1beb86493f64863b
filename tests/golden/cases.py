"""The tiny problems of the golden vectors (shared by make_golden.py and the tests)."""
from __future__ import annotations

import numpy as np

from implexity_amd import design
from tests.lbm_cases import channel_ports, make_pair, random_alpha, random_solid

CASES = {
    # D3Q19 TRT channel with solids and NEE ports, 10 steps
    "trt_channel": dict(kind="forward", shape=(7, 5, 4), periodic=(False, False, False), Q=19, collision="trt",
                        ports=True, solid=True, accel=(0.0, 0.0, 0.0), steps=10),
    # D3Q27 cumulant, periodic x with a body force, 6 steps
    "cumulant_periodic": dict(kind="forward", shape=(5, 4, 4), periodic=(True, False, False), Q=27,
                              collision="cumulant", ports=False, solid=False, accel=(1e-4, 0.0, 0.0), steps=6),
    # D3Q19 MRT all-periodic with a body force, 8 steps
    "mrt_periodic": dict(kind="forward", shape=(5, 4, 3), periodic=(True, True, True), Q=19, collision="mrt",
                         ports=False, solid=False, accel=(1e-4, 2e-5, 0.0), steps=8),
    # pressure-drop gradient through 12 steps of the TRT channel
    "trt_gradient": dict(kind="gradient", shape=(8, 5, 4), periodic=(False, False, False), Q=19, collision="trt",
                         ports=True, solid=False, accel=(0.0, 0.0, 0.0), steps=12, window=(6, 12),
                         weight=1.0 / 7),
    # coupled D3Q7 temperature on the TRT channel, 9 steps
    "thermal_channel": dict(kind="thermal", shape=(7, 5, 4), periodic=(False, False, False), Q=19,
                            collision="trt", ports=True, solid=False, accel=(0.0, 0.0, 0.0), steps=9, window=(5, 9),
                            weight=0.2, thermal=dict(kappa=0.02, ratio=10.0, source=1e-3, inlet=0.3)),
}


def build(spec, precision="f64"):
    """(oracle Problem, alpha, thermal design or None) -- or with want_cfg the GPU config."""
    shape = spec["shape"]
    solid = None
    if spec["solid"]:
        solid = random_solid(shape, 21)
        solid[:2] = False
        solid[-2:] = False
    ports = channel_ports(shape, 0.03) if spec["ports"] else ()
    _, prob = make_pair(shape, spec["periodic"], Q=spec["Q"], collision=spec["collision"], precision=precision,
                        omega=1.25, accel=spec["accel"], solid=solid, ports=ports,
                        bulk_rate=1.0 if spec["collision"] == "cumulant" else float("nan"))
    alpha = random_alpha(shape, 23)
    phi = np.random.default_rng(24).random(int(np.prod(shape))) if spec["kind"] == "thermal" else None
    return prob, alpha, phi


def gpu_config(spec, precision="f64"):
    shape = spec["shape"]
    solid = None
    if spec["solid"]:
        solid = random_solid(shape, 21)
        solid[:2] = False
        solid[-2:] = False
    ports = channel_ports(shape, 0.03) if spec["ports"] else ()
    cfg, _ = make_pair(shape, spec["periodic"], Q=spec["Q"], collision=spec["collision"], precision=precision,
                       omega=1.25, accel=spec["accel"], solid=solid, ports=ports,
                       bulk_rate=1.0 if spec["collision"] == "cumulant" else float("nan"))
    if spec["kind"] == "thermal":
        cfg.thermal = True
        th = spec["thermal"]
        cfg.thermal_kappa, cfg.thermal_ratio = th["kappa"], th["ratio"]
        cfg.thermal_source, cfg.thermal_inlet = th["source"], th["inlet"]
    return cfg


_ = design  # the design helpers are used through tests.lbm_cases
