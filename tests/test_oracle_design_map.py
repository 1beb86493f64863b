"""Pins of the design-map restatement (DesignMap, design.rs:95-280)."""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle.design_map import DesignMapOracle


def test_small_radius_is_the_identity():
    rng = np.random.default_rng(0)
    raw = rng.random(60)
    solid = rng.random(60) < 0.1
    for r in (0.0, 0.5, 1.0):
        m = DesignMapOracle((3, 4, 5), radius=r, solid=solid)
        want = np.where(solid, 0.0, raw)
        assert np.array_equal(m.map(raw), want)


def test_seven_point_cone():
    """radius 1.4 (below the edge distance sqrt 2): weight 1.4 on the cell and 0.4 on its
    six face neighbours inside the lattice (closed axes) or across a periodic axis."""
    shape = (4, 5, 6)
    rng = np.random.default_rng(1)
    raw = rng.random(shape)
    m = DesignMapOracle(shape, periodic=(False, True, False), radius=1.4)
    got = m.map(raw).reshape(shape)
    for i in range(4):
        for j in range(5):
            for k in range(6):
                num, den = 1.4 * raw[i, j, k], 1.4
                for di, dj, dk in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
                    a, b, c = i + di, (j + dj) % 5, k + dk
                    if 0 <= a < 4 and 0 <= c < 6:
                        num += (1.4 - 1.0) * raw[a, b, c]
                        den += 1.4 - 1.0
                assert abs(got[i, j, k] - num / den) <= 2e-16 * 8


@pytest.mark.parametrize("r", [1.7, 2.5, 4.0])
def test_constants_are_reproduced(r):
    shape = (6, 5, 4)
    rng = np.random.default_rng(2)
    solid = rng.random(shape) < 0.15
    region = rng.random(shape) < 0.8
    fixed = rng.random(shape)
    m = DesignMapOracle(shape, periodic=(True, False, False), radius=r, solid=solid, region=region, fixed=fixed)
    phi = m.map(np.full(m.n, 0.37))
    ed = m.editable
    assert np.abs(phi[ed] - 0.37).max() <= 1e-15
    assert np.all(phi[solid.reshape(-1)] == 0.0)
    other = ~ed & ~solid.reshape(-1)
    assert np.array_equal(phi[other], fixed.reshape(-1)[other])


def test_projection_end_points_and_small_beta():
    m = DesignMapOracle((2, 2, 2), beta=8.0, eta=0.3)
    assert np.allclose(m.map(np.zeros(8)), 0.0, atol=1e-15) and np.allclose(m.map(np.ones(8)), 1.0, atol=1e-15)
    f = np.linspace(0.0, 1.0, 8)
    assert np.all(np.diff(m.map(f)) > 0)
    m2 = DesignMapOracle((2, 2, 2), beta=1e-6, eta=0.3)
    assert np.abs(m2.map(f) - f).max() <= 1e-6
    a, den = math.tanh(8 * 0.3), math.tanh(8 * 0.3) + math.tanh(8 * 0.7)
    assert abs(m.map(np.full(8, 0.3))[0] - a / den) <= 1e-15


@pytest.mark.parametrize("r,beta", [(0.0, 0.0), (2.2, 0.0), (2.2, 6.0), (3.0, 12.0)])
def test_vjp_matches_differences_and_partial(r, beta):
    shape = (5, 4, 5)
    rng = np.random.default_rng(3)
    solid = rng.random(shape) < 0.1
    region = rng.random(shape) < 0.85
    m = DesignMapOracle(shape, periodic=(False, False, True), radius=r, beta=beta, eta=0.45, solid=solid,
                        region=region, fixed=rng.random(shape))
    raw = rng.random(m.n)
    pb = rng.standard_normal(m.n)
    g = m.vjp(raw, pb)
    assert np.abs(g - m.partial(raw).T @ pb).max() <= 1e-13 * np.abs(g).max()
    v = rng.standard_normal(m.n)
    h = 1e-6
    fd = (pb @ m.map(raw + h * v) - pb @ m.map(raw - h * v)) / (2 * h)
    assert abs(fd - g @ v) <= 1e-7 * max(1.0, abs(fd))
    assert np.all(g[~m.editable] == 0.0)
