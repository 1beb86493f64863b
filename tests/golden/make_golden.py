"""Writes tests/golden/lbm_small.npz: small FP64 regression vectors of the oracle.

The reference holds no fixtures or golden vectors for this path (its LBM and AD tests
are disabled and it cannot be built here), so these vectors are not reference outputs:
they freeze the oracle's FP64 results for a few tiny cases (each oracle function being
pinned separately by closed forms, invariants and central differences in
tests/test_oracle*.py), so that a change of the oracle shows up as a fixture mismatch
and the GPU parity of these cases can be checked against data alone.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as orc  # noqa: E402
from oracle import revolve as orc_rev  # noqa: E402
from oracle.thermal import CoupledObjective, Thermal, ThermalOracle  # noqa: E402
from tests.golden.cases import CASES, build  # noqa: E402


KINDS = ("advance", "snapshot", "restore", "reverse", "release")


def encode_actions(acts):
    """(kind, a, b) rows, b = -1 for one-argument actions."""
    rows = [(KINDS.index(a[0]), a[1], a[2] if len(a) > 2 else -1) for a in acts]
    return np.array(rows, np.int64)


def main():
    out = {}
    for name, spec in CASES.items():
        prob, alpha, extra = build(spec)
        o = orc.Oracle(prob)
        if spec["kind"] == "forward":
            out[f"{name}/state"] = o.run(alpha, spec["steps"])
        elif spec["kind"] == "gradient":
            ob = orc.Objective(kind="pressure_drop", window=tuple(spec["window"]), weight=spec["weight"], axis=0,
                               plane_a=1, plane_b=prob.shape[0] - 2)
            J, ab, lam, final = orc.scan_adjoint(o, alpha, spec["steps"], ob)
            out[f"{name}/J"] = np.array([J])
            out[f"{name}/alpha_bar"] = ab
            out[f"{name}/lambda0"] = lam
            out[f"{name}/state"] = final
        elif spec["kind"] == "thermal":
            t = ThermalOracle(o, Thermal(**spec["thermal"]))
            phi = extra
            f, g = t.run(alpha, phi, spec["steps"])
            out[f"{name}/state"] = f
            out[f"{name}/tstate"] = g
            ob = CoupledObjective(window=tuple(spec["window"]), weight=spec["weight"], w_T=1.0, w_dp=0.1,
                                  plane_a=1, plane_b=prob.shape[0] - 2)
            J, ab, pb, _, _ = t.scan_adjoint(ob, alpha, phi, spec["steps"])
            out[f"{name}/J"] = np.array([J])
            out[f"{name}/alpha_bar"] = ab
            out[f"{name}/phi_bar"] = pb
    acts, rstart, fwd, peak = orc_rev.binomial_schedule(30, 4)
    out["revolve_30_4/actions"] = encode_actions(acts)
    out["revolve_30_4/info"] = np.array([rstart, fwd, peak], np.int64)
    path = os.path.join(ROOT, "tests", "golden", "lbm_small.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: {len(out)} arrays, {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
