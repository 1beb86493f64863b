"""CPU restatement of the reference's design map -- TEST INFRASTRUCTURE ONLY.

Only tests/ (and the checker legs of __graft_entry__.smoke() / bench.py) use it, as the
checker of the device design map.  It restates DesignMap::{new, map, vjp, partial}
(crates/implexity-physics-lbm/src/design.rs:95-280) with the filter radius given in
cells (filter_radius_m / spacing_m), rows built in the reference's (a, b, c) loop order
with Grid::inside / Grid::wrap (d3q19.rs:101-128).

Pins (tests/test_oracle_design_map.py): radius <= 1 is the identity filter; the 7-point
cone at radius 1.4 against its explicit stencil; constants are reproduced on the
editable cells (the filter is a partition of unity); the projection maps 0 -> 0,
1 -> 1 and tends to the identity as beta -> 0; vjp against central differences and
against the transpose of partial.
"""
from __future__ import annotations

import math

import numpy as np


class DesignMapOracle:
    def __init__(self, shape, periodic=(False, False, False), radius=0.0, beta=0.0, eta=0.5, solid=None,
                 region=None, fixed=None):
        if not (0.0 <= radius <= 4.0):
            raise ValueError("filter radius must lie in [0,4] cells")
        if not (0.0 <= beta <= 32.0) or not (0.0 < eta < 1.0):
            raise ValueError("Require beta in [0,32] and eta in (0,1)")
        self.shape = tuple(int(s) for s in shape)
        nx, ny, nz = self.shape
        self.n = n = nx * ny * nz
        self.periodic = tuple(bool(p) for p in periodic)
        self.solid = np.zeros(n, bool) if solid is None else np.asarray(solid, bool).reshape(-1)
        reg = np.ones(n, bool) if region is None else np.asarray(region, bool).reshape(-1)
        self.editable = reg & ~self.solid
        self.fixed = np.zeros(n) if fixed is None else np.asarray(fixed, np.float64).reshape(-1)
        self.beta, self.eta = float(beta), float(eta)
        self.denominators = np.ones(n)
        self.rows = None
        if radius > 0.0:
            reach = int(math.ceil(radius))
            offsets = []
            for a in range(-reach, reach + 1):
                for b in range(-reach, reach + 1):
                    for c in range(-reach, reach + 1):
                        norm = math.sqrt(float(a * a + b * b + c * c))
                        weight = max(radius - norm, 0.0)
                        if weight != 0.0:
                            offsets.append(((a, b, c), weight))
            self.offsets = offsets
            rows = []
            for x in range(n):
                i, j, k = x // (ny * nz), (x // nz) % ny, x % nz
                terms = []
                den = 0.0
                for (a, b, c), w in offsets:
                    v = [i - a, j - b, k - c]
                    ok = True
                    for t in range(3):
                        if not self.periodic[t] and not (0 <= v[t] < self.shape[t]):
                            ok = False
                        v[t] %= self.shape[t]
                    if not ok:
                        continue
                    s = (v[0] * ny + v[1]) * nz + v[2]
                    if not self.editable[s]:
                        continue
                    terms.append((s, w))
                    den += w
                self.denominators[x] = den if den > 0.0 else 1.0
                rows.append(terms)
            self.rows = rows

    def filtered(self, raw, x):
        if self.rows is None:
            return float(raw[x])
        num = 0.0
        for s, w in self.rows[x]:
            num += w * float(raw[s])
        return num / self.denominators[x]

    def _projection(self):
        if self.beta > 0.0:
            a = math.tanh(self.beta * self.eta)
            return a, a + math.tanh(self.beta * (1.0 - self.eta))
        return None

    def map(self, raw):
        raw = np.asarray(raw, np.float64).reshape(-1)
        proj = self._projection()
        out = np.zeros(self.n)
        for x in range(self.n):
            if self.solid[x]:
                out[x] = 0.0
            elif self.editable[x]:
                f = self.filtered(raw, x)
                out[x] = (proj[0] + math.tanh(self.beta * (f - self.eta))) / proj[1] if proj else f
            else:
                out[x] = self.fixed[x]
        return out

    def vjp(self, raw, phi_bar):
        raw = np.asarray(raw, np.float64).reshape(-1)
        phi_bar = np.asarray(phi_bar, np.float64).reshape(-1)
        proj = self._projection()
        out = np.zeros(self.n)
        for x in range(self.n):
            if not self.editable[x] or phi_bar[x] == 0.0:
                continue
            g = phi_bar[x]
            if proj:
                t = math.tanh(self.beta * (self.filtered(raw, x) - self.eta))
                g *= self.beta * (1.0 - t * t) / proj[1]
            if self.rows is None:
                out[x] += g
            else:
                scale = g / self.denominators[x]
                for s, w in self.rows[x]:
                    out[s] += w * scale
        return out

    def partial(self, raw):
        """Dense d phi / d raw (DesignMap::partial as a matrix; small problems)."""
        raw = np.asarray(raw, np.float64).reshape(-1)
        proj = self._projection()
        J = np.zeros((self.n, self.n))
        for x in range(self.n):
            if not self.editable[x]:
                continue
            slope = 1.0
            if proj:
                t = math.tanh(self.beta * (self.filtered(raw, x) - self.eta))
                slope = self.beta * (1.0 - t * t) / proj[1]
            if self.rows is None:
                J[x, x] = slope
            else:
                for s, w in self.rows[x]:
                    J[x, s] += slope * w / self.denominators[x]
        return J
