"""The benchmark's algorithmic-byte accounting of alpha_bar (bench.abar_bytes): FP64
storage reads and writes the FP64 sum every reverse step (16 B); with FP32 arithmetic the
gradient driver gathers runs of ABAR_RUN reverse steps in an FP32 partial -- a run's first
step writes it (4 B), its middle steps read and write it (8 B), its last step reads it and
updates the FP64 sum (4 + 8 + 8 B), a run of one step updates the FP64 sum directly
(16 B), and the sweep's last reverse step (k = 1) always closes a run.  The closed form is
checked against that rule enumerated step by step over the sweep k = n .. 1."""
from __future__ import annotations

import pytest

import bench


def _enumerated(steps, run=bench.ABAR_RUN):
    total, in_run = 0, 0
    for k in range(steps, 0, -1):
        fold = k == 1 or in_run + 1 == run
        if in_run == 0 and fold:
            total += 16  # a run of one step: the FP64 sum directly
        elif in_run == 0:
            total += 4
        elif fold:
            total += 20
        else:
            total += 8
        in_run = 0 if fold else in_run + 1
    return total / steps


@pytest.mark.parametrize("steps", [1, 2, 3, 20, 31, 32, 33, 63, 64, 65, 1000, 8000])
def test_abar_bytes_matches_the_run_rule(steps):
    assert bench.abar_bytes("f16s", steps) == pytest.approx(_enumerated(steps), rel=1e-15)
    assert bench.abar_bytes("f32", steps) == pytest.approx(_enumerated(steps), rel=1e-15)


def test_abar_bytes_fp64_and_isolated_launches():
    assert bench.abar_bytes("f64", 8000) == 16.0
    assert bench.abar_bytes("f16s") == 8.0  # a middle step, as the isolated kernel timing runs it
    assert bench.algorithmic_bytes("adjoint", "f16s", steps=8000) == pytest.approx(19 * 2 * 3 + 4 + 8.25)
