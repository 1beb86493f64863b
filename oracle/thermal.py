"""Coupled D3Q19 flow + D3Q7 temperature oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/ and __graft_entry__.smoke() may import this module, as the checker.  It
wraps the thermal section of ``oracle/lbm_oracle.c`` (FP64, -ffp-contract=off).

The thermal model is not in the reference (its thermal coupling is finite-volume energy
transport on D3Q19 links, crates/implexity-physics-lbm/src/lbm/thermal.rs:91-134); it is
defined here and in lbm_oracle.c:

* T = sum_i g_i, g_i^eq = w_i T (1 + 4 c_i.u) on D3Q7 (w = 1/4, 1/8; c_s^2 = 1/4), with
  u the Brinkman velocity of the flow state that the same interval collides;
* kappa(phi) = kappa_f (1 + (ratio - 1) phi), tau_T = 1/2 + 4 kappa, S(phi) = source phi;
* g* = g - omega_T (g - g^eq) + w S, halfway bounce-back (adiabatic walls), then at the
  flow's port cells: velocity ports hold T = T_in (g_b = w T_in + g_nb - w T_nb),
  pressure ports are zero-gradient (g_b = g_nb).

A coupled interval maps (f, g) to (interval(f), thermal_interval(f, g)): the thermal
step reads the flow state before the flow step.  The coupled objective of a state is
w_T * mean T over the cells with phi > 1/2 + w_dp * pressure drop of the flow.

Pinned by (tests/test_oracle_thermal.py): lattice moments, heat conservation with and
without sources, decay of a diffusive Fourier mode and drift of an advected one against
the closed forms, the ports' boundary values, central differences of every VJP and the
dot-product identity of the coupled step.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from oracle.oracle import Oracle, _ptr, lib


@dataclass
class Thermal:
    kappa: float = 0.01
    ratio: float = 10.0
    source: float = 1e-5
    inlet: float = 0.0


class _Th(ctypes.Structure):
    _fields_ = [("kappa", ctypes.c_double), ("ratio", ctypes.c_double), ("source", ctypes.c_double),
                ("inlet", ctypes.c_double)]


@dataclass
class CoupledObjective:
    """Stage cost weight * (w_T * mean solid T + w_dp * pressure drop) on states s in
    window (indexed 1..n), plus terminal_weight * the same on the final state."""

    window: tuple = (1, 1)
    weight: float = 1.0
    terminal_weight: float = 0.0
    w_T: float = 1.0
    w_dp: float = 0.1
    axis: int = 0
    plane_a: int = 1
    plane_b: int = -2

    def stage_weight(self, s):
        lo, hi = self.window
        return self.weight if lo <= s <= hi else 0.0


class ThermalOracle:
    """The flow oracle plus the D3Q7 temperature lattice (AoS g, index x*7+q)."""

    def __init__(self, orc: Oracle, th: Thermal):
        if orc.prob.solid is not None and np.any(orc.prob.solid):
            raise ValueError("the thermal model has no hard solid cells")
        self.orc = orc
        self.th = th
        self.n = orc.n
        self._t = _Th(float(th.kappa), float(th.ratio), float(th.source), float(th.inlet))
        self._tp = ctypes.byref(self._t)
        self.source7 = np.zeros(self.n * 7, np.uint32)
        lib().orc_thermal_stream_table(orc._pp, _ptr(self.source7))
        lib().orc_obj_solid_temperature.restype = ctypes.c_double

    def omega(self, phi):
        phi = np.asarray(phi, np.float64)
        return 1.0 / (0.5 + 4.0 * (self.th.kappa * (1.0 + (self.th.ratio - 1.0) * phi)))

    def initial(self):
        return np.zeros(self.n * 7)

    def cell(self, g, u, om, S):
        out = np.zeros(7)
        lib().orc_thermal_cell(_ptr(np.ascontiguousarray(g, np.float64)),
                               _ptr(np.ascontiguousarray(u, np.float64)), ctypes.c_double(om), ctypes.c_double(S),
                               _ptr(out))
        return out

    def cell_vjp(self, g, u, om, gb):
        g_bar = np.zeros(7)
        u_bar = np.zeros(3)
        ob = ctypes.c_double(0.0)
        sb = ctypes.c_double(0.0)
        lib().orc_thermal_cell_vjp(_ptr(np.ascontiguousarray(g, np.float64)),
                                   _ptr(np.ascontiguousarray(u, np.float64)), ctypes.c_double(om),
                                   _ptr(np.ascontiguousarray(gb, np.float64)), _ptr(g_bar), _ptr(u_bar),
                                   ctypes.byref(ob), ctypes.byref(sb))
        return g_bar, u_bar, ob.value, sb.value

    def interval(self, alpha, phi, f_before, g_before):
        post = np.empty(self.n * 7)
        out = np.empty(self.n * 7)
        lib().orc_thermal_interval(self.orc._pp, self._tp, _ptr(self.source7),
                                   _ptr(np.ascontiguousarray(alpha, np.float64)),
                                   _ptr(np.ascontiguousarray(phi, np.float64)),
                                   _ptr(np.ascontiguousarray(f_before, np.float64)),
                                   _ptr(np.ascontiguousarray(g_before, np.float64)), _ptr(post), _ptr(out))
        return out

    def interval_vjp(self, alpha, phi, f_before, g_before, g_after_bar):
        """Returns (g_bar, f_bar, alpha_bar, phi_bar) of <g_after_bar, thermal_interval>."""
        gab = np.array(g_after_bar, np.float64)
        post_bar = np.zeros(self.n * 7)
        g_bar = np.zeros(self.n * 7)
        f_bar = np.zeros(self.n * self.orc.Q)
        a_bar = np.zeros(self.n)
        p_bar = np.zeros(self.n)
        lib().orc_thermal_interval_vjp(self.orc._pp, self._tp, _ptr(self.source7),
                                       _ptr(np.ascontiguousarray(alpha, np.float64)),
                                       _ptr(np.ascontiguousarray(phi, np.float64)),
                                       _ptr(np.ascontiguousarray(f_before, np.float64)),
                                       _ptr(np.ascontiguousarray(g_before, np.float64)), _ptr(gab), _ptr(post_bar),
                                       _ptr(g_bar), _ptr(f_bar), _ptr(a_bar), _ptr(p_bar))
        return g_bar, f_bar, a_bar, p_bar

    def temperature(self, g):
        T = np.zeros(self.n)
        lib().orc_temperature(self.orc._pp, _ptr(np.ascontiguousarray(g, np.float64)), _ptr(T))
        return T

    def solid_temperature(self, phi, g):
        return lib().orc_obj_solid_temperature(self.orc._pp, _ptr(np.ascontiguousarray(phi, np.float64)),
                                               _ptr(np.ascontiguousarray(g, np.float64)))

    def solid_temperature_grad(self, phi, weight):
        gb = np.zeros(self.n * 7)
        lib().orc_obj_solid_temperature_grad(self.orc._pp, _ptr(np.ascontiguousarray(phi, np.float64)),
                                             ctypes.c_double(weight), _ptr(gb))
        return gb

    # --- coupled problem ----------------------------------------------------------------
    def step(self, alpha, phi, f, g, t=0):
        return self.orc.interval(alpha, f, t=t), self.interval(alpha, phi, f, g)

    def run(self, alpha, phi, steps, f0=None, g0=None):
        f = self.orc.initial() if f0 is None else np.array(f0, np.float64)
        g = self.initial() if g0 is None else np.array(g0, np.float64)
        for t in range(steps):
            f, g = self.step(alpha, phi, f, g, t)
        return f, g

    def objective(self, ob: CoupledObjective, alpha, phi, f, g):
        a, b = ob.plane_a % self.orc.prob.shape[ob.axis], ob.plane_b % self.orc.prob.shape[ob.axis]
        v = 0.0
        if ob.w_T:
            v += ob.w_T * self.solid_temperature(phi, g)
        if ob.w_dp:
            v += ob.w_dp * self.orc.pressure_drop(f, ob.axis, a, b)
        return v

    def objective_grad(self, ob: CoupledObjective, phi, weight):
        a, b = ob.plane_a % self.orc.prob.shape[ob.axis], ob.plane_b % self.orc.prob.shape[ob.axis]
        fb = np.zeros(self.n * self.orc.Q)
        gb = np.zeros(self.n * 7)
        if ob.w_T:
            gb += self.solid_temperature_grad(phi, weight * ob.w_T)
        if ob.w_dp:
            fb += self.orc.pressure_drop_grad(ob.axis, a, b, weight * ob.w_dp)
        return fb, gb

    def objective_value(self, ob: CoupledObjective, alpha, phi, steps):
        f, g = self.orc.initial(), self.initial()
        J = 0.0
        for t in range(steps):
            f, g = self.step(alpha, phi, f, g, t)
            w = ob.stage_weight(t + 1)
            if w:
                J += w * self.objective(ob, alpha, phi, f, g)
        if ob.terminal_weight:
            J += ob.terminal_weight * self.objective(ob, alpha, phi, f, g)
        return J

    def step_vjp(self, alpha, phi, f, g, lam_f, lam_g, t=0):
        """Transpose of one coupled interval at (f, g) read at step t: (f_bar, g_bar,
        alpha_bar, phi_bar)."""
        gb, fb_t, ab_t, pb = self.interval_vjp(alpha, phi, f, g, lam_g)
        fb, ab = self.orc.interval_vjp(alpha, f, lam_f, t=t)
        return fb + fb_t, gb, ab + ab_t, pb

    def scan_adjoint(self, ob: CoupledObjective, alpha, phi, steps):
        """Forward + reverse sweep with every state kept (small problems): returns
        (J, alpha_bar, phi_bar, lambda_f0, lambda_g0)."""
        alpha = np.asarray(alpha, np.float64)
        phi = np.asarray(phi, np.float64)
        fs, gs = [self.orc.initial()], [self.initial()]
        J = 0.0
        for t in range(steps):
            f, g = self.step(alpha, phi, fs[-1], gs[-1], t)
            fs.append(f)
            gs.append(g)
            w = ob.stage_weight(t + 1)
            if w:
                J += w * self.objective(ob, alpha, phi, f, g)
        lam_f = np.zeros(self.n * self.orc.Q)
        lam_g = np.zeros(self.n * 7)
        abar = np.zeros(self.n)
        pbar = np.zeros(self.n)
        if ob.terminal_weight:
            J += ob.terminal_weight * self.objective(ob, alpha, phi, fs[-1], gs[-1])
            fb, gb = self.objective_grad(ob, phi, ob.terminal_weight)
            lam_f += fb
            lam_g += gb
        for t in range(steps - 1, -1, -1):
            w = ob.stage_weight(t + 1)
            if w:
                fb, gb = self.objective_grad(ob, phi, w)
                lam_f = lam_f + fb
                lam_g = lam_g + gb
            lam_f, lam_g, ab, pb = self.step_vjp(alpha, phi, fs[t], gs[t], lam_f, lam_g, t)
            abar += ab
            pbar += pb
        return J, abar, pbar, lam_f, lam_g
