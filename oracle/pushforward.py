"""CPU restatement of the push-forward of body points into the lattice occupancy --
TEST INFRASTRUCTURE ONLY (the checker of the device push-forward).

Restates crates/implexity-physics-lbm/src/moving/pushforward.rs: the frame checks of
Frame::new / with_mirrors (:113-165; a periodic axis is at least the kernel support
long unless it is one cell, planar, with weight 1 on its cell; a closed axis is longer
than one cell; a symmetry face sits on a closed axis of at least the support), the
kernels m4 / m4_derivative (:41-70) and peskin4 / peskin4_derivative (:88-100), the
per-axis weights of Stencils::new_with_kernel (:205-318; the closed-form cubic branch
for width 1, the kernel formula otherwise, folded by fold_axis, :344-362, into the
lattice at a symmetry face the support crosses), the scatter of a_q and a_q v_q into d
and m (:766-860, summed here in point order), and endpoint_vjp (:946-967) for the
cotangents of the points, their weights and velocities.  Positions are lattice
coordinates (cell c at coordinate c; Frame::to_lattice, :186-189, maps metres to them).

Pins (tests/test_oracle_pushforward.py): both kernels are partitions of unity and
reproduce linear functions, so the total occupancy is sum a_q and its first moment
sum a_q xi_q; the width-1 closed form equals m4; Peskin's even/odd sums are 1/2; a
symmetry face equals the unfolded scatter of the point and its mirror image restricted
to the lattice (the kernels are even); a planar axis equals the sum over a periodic
axis of the support's length; the transposes against central differences.
"""
from __future__ import annotations

import math

import numpy as np


def m4(s):
    a = -s if s < 0.0 else s
    if a < 1.0:
        return (4.0 - a * a * 6.0 + a * a * a * 3.0) / 6.0
    if a < 2.0:
        t = 2.0 - a
        return t * t * t / 6.0
    return 0.0


def m4_derivative(s):
    neg = s < 0.0
    a = -s if neg else s
    if a < 1.0:
        d = -a * 2.0 + a * a * 1.5
    elif a < 2.0:
        t = 2.0 - a
        d = -(t * t) * 0.5
    else:
        d = 0.0
    return -d if neg else d


def peskin4(s):
    r = -s if s < 0.0 else s
    if r < 1.0:
        return (3.0 - r * 2.0 + math.sqrt(1.0 + r * 4.0 - r * r * 4.0)) / 8.0
    if r < 2.0:
        return (5.0 - r * 2.0 - math.sqrt(-7.0 + r * 12.0 - r * r * 4.0)) / 8.0
    return 0.0


def peskin4_derivative(s):
    neg = s < 0.0
    r = -s if neg else s
    if r < 1.0:
        d = (-2.0 + (2.0 - r * 4.0) / math.sqrt(1.0 + r * 4.0 - r * r * 4.0)) / 8.0
    elif r < 2.0:
        d = (-2.0 - (6.0 - r * 4.0) / math.sqrt(-7.0 + r * 12.0 - r * r * 4.0)) / 8.0
    else:
        d = 0.0
    return -d if neg else d


def axis_weights(x, h, kernel):
    """(first cell, weights[4h], d weights / d x[4h]) of one coordinate."""
    b = math.floor(x)
    first = b - 2 * h + 1
    n = 4 * h
    w, dw = np.zeros(n), np.zeros(n)
    if h == 1 and kernel == "cubic":
        phi = x - b
        r = 1.0 - phi
        p2, r2 = phi * phi, r * r
        w[:] = [r2 * r / 6.0, (p2 * (phi * 3.0 - 6.0) + 4.0) / 6.0,
                (((-phi * 3.0 + 3.0) * phi + 3.0) * phi + 1.0) / 6.0, p2 * phi / 6.0]
        dw[:] = [-r2 * 0.5, phi * (phi * 1.5 - 2.0), (-phi * 1.5 + 1.0) * phi + 0.5, p2 * 0.5]
    else:
        k, kd = (m4, m4_derivative) if kernel == "cubic" else (peskin4, peskin4_derivative)
        for t in range(n):
            r = (float(first + t) - x) / h
            w[t] = k(r) / h
            dw[t] = -kd(r) / (h * h)
    return first, w, dw


def fold(w, first, start, n):
    """fold_axis: weights of cells first + t reflected into the lattice (c < 0 -> -1 - c,
    c >= n -> 2n - 1 - c) and summed at offsets from start."""
    out = np.zeros_like(w)
    for t in range(len(w)):
        c = first + t
        m = -1 - c if c < 0 else (2 * n - 1 - c if c >= n else c)
        out[m - start] += w[t]
    return out


class PushForwardOracle:
    def __init__(self, shape, periodic=(False, False, False), width=1, kernel="cubic",
                 mirrored=((False, False), (False, False), (False, False))):
        self.shape = tuple(int(s) for s in shape)
        self.periodic = tuple(bool(p) for p in periodic)
        self.h, self.kernel = int(width), kernel
        self.mirrored = tuple((bool(lo), bool(hi)) for lo, hi in mirrored)
        self.n = int(np.prod(self.shape))
        span = 4 * self.h
        for a in range(3):
            if self.periodic[a] and 1 < self.shape[a] < span:
                raise ValueError(f"periodic axis {a} is shorter than the push-forward kernel support")
            if not self.periodic[a] and self.shape[a] == 1:
                raise ValueError(f"a one-cell lattice axis ({a}) must be periodic (planar)")
            if any(self.mirrored[a]) and (self.periodic[a] or self.shape[a] < span):
                raise ValueError(f"push-forward mirror face on axis {a} needs a non-periodic axis")

    def planar(self, a):
        return self.periodic[a] and self.shape[a] == 1

    def support(self, a):
        return 1 if self.planar(a) else 4 * self.h

    def _check(self, xi):
        for q, x in enumerate(xi):
            for a in range(3):
                v = float(x[a])
                if not math.isfinite(v):
                    raise ValueError(f"push-forward point {q} is not finite")
                if self.periodic[a]:
                    continue
                lo, hi = self.mirrored[a]
                n = self.shape[a]
                if (lo and v < -0.5) or (hi and v > n - 0.5):
                    raise ValueError(f"point {q} lies beyond a symmetry face")
                if (not lo and v - 2.0 * self.h <= -1.0) or (not hi and v + 2.0 * self.h >= n):
                    raise ValueError(f"point {q} has kernel support outside the lattice")

    def _axis(self, x, a):
        """(first cell, weights, d weights / d x) of coordinate x on axis a, folded at a
        symmetry face the support crosses."""
        if self.planar(a):
            return 0, np.ones(1), np.zeros(1)
        first, w, dw = axis_weights(x, self.h, self.kernel)
        n, span = self.shape[a], 4 * self.h
        lo, hi = self.mirrored[a]
        start = 0 if (lo and first < 0) else (n - span if (hi and first + span > n) else None)
        if start is not None:
            w, dw, first = fold(w, first, start, n), fold(dw, first, start, n), start
        return first, w, dw

    def _stencil(self, x):
        """cells (flat, per tensor entry), weights, and the three partial weight products."""
        per = [self._axis(float(x[a]), a) for a in range(3)]
        out = []
        nx, ny, nz = self.shape
        for tx in range(len(per[0][1])):
            for ty in range(len(per[1][1])):
                for tz in range(len(per[2][1])):
                    c = [per[0][0] + tx, per[1][0] + ty, per[2][0] + tz]
                    c = [c[a] % self.shape[a] for a in range(3)]
                    cell = (c[0] * ny + c[1]) * nz + c[2]
                    wx, wy, wz = per[0][1][tx], per[1][1][ty], per[2][1][tz]
                    dx, dy, dz = per[0][2][tx], per[1][2][ty], per[2][2][tz]
                    out.append((cell, wx * wy * wz, (dx * wy * wz, wx * dy * wz, wx * wy * dz)))
        return out

    def scatter(self, xi, a, v=None):
        xi = np.asarray(xi, np.float64).reshape(-1, 3)
        self._check(xi)
        d = np.zeros(self.n)
        m = np.zeros((self.n, 3))
        for q in range(len(xi)):
            for cell, W, _ in self._stencil(xi[q]):
                d[cell] += a[q] * W
                if v is not None:
                    m[cell] += a[q] * W * np.asarray(v[q], np.float64)
        return d, m

    def vjp(self, xi, a, v, d_bar, m_bar):
        """(xi_bar, a_bar, v_bar) of <d_bar, d> + <m_bar, m>."""
        xi = np.asarray(xi, np.float64).reshape(-1, 3)
        v = np.asarray(v, np.float64).reshape(-1, 3)
        m_bar = np.asarray(m_bar, np.float64).reshape(-1, 3)
        xb, ab, vb = np.zeros_like(xi), np.zeros(len(xi)), np.zeros_like(v)
        for q in range(len(xi)):
            ks, kv, gs, gv = 0.0, np.zeros(3), np.zeros(3), np.zeros((3, 3))
            for cell, W, dW in self._stencil(xi[q]):
                ks += W * d_bar[cell]
                kv += W * m_bar[cell]
                for e in range(3):
                    gs[e] += dW[e] * d_bar[cell]
                    gv[e] += dW[e] * m_bar[cell]
            ab[q] = ks + kv @ v[q]
            xb[q] = a[q] * (gs + gv @ v[q])
            vb[q] = a[q] * kv
        return xb, ab, vb
