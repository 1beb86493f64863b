"""Design sensitivities of the LBM responses through the device path.

``LbmSensitivities`` follows LatticeBoltzmannProvider::{parts, evaluate_value,
sensitivities_value} (crates/implexity-physics-lbm/src/provider.rs:249-365) with the
library doing the lattice work:

    raw -> DesignMap (device) -> phi -> alpha = resistance(phi) (design.rs:282-294)
        -> history (lbm_forward) -> vector(phi, alpha, final state)
    per response r: (f_bar, alpha_bar_final, phi_bar) = vector_vjp(w = e_r)
        alpha_bar = history_adjoint seeded with f_bar (lbm_gradient_seeded)
        phi_bar += (alpha_bar + alpha_bar_final) * resistance'(phi)
        gradient = DesignMap::vjp(raw, phi_bar)

``LbmSensitivities.from_problem`` builds the chain from the reference's JSON problem
through LbmProblem::normalize (implexity_amd.problem): the lattice (D3Q19 BGK-Guo, the
problem's ports), the design map, the step count, the response scales and the resistance,
with admission (LbmProblem::admit, solver.rs:707-793) at the problem's Mach limit, so a
trajectory the reference refuses raises ConvergenceError here too.  The optimiser stays
with the caller.
"""
from __future__ import annotations

import numpy as np

from . import design
from .lbm import DesignMap, Lattice, LatticeConfig, LbmError
from .problem import LbmProblem, normalize

RESPONSES = ("mean_velocity_x_m_s", "mean_velocity_y_m_s", "mean_velocity_z_m_s", "kinetic_energy_J",
             "porous_dissipation_W", "fluid_volume_m3")


class LbmSensitivities:
    """cfg: the lattice (its design is set here); steps: the history length; design_map:
    a DesignMap of the same shape; scales: (velocity, density, step_s, cell_volume) of
    the responses (LbmProblem::vector_scales); alpha_max, q: the resistance's
    step_s * drag_max_per_s and drag shape."""

    def __init__(self, cfg: LatticeConfig, steps: int, design_map: DesignMap, scales, alpha_max=0.1, q=0.1):
        if design_map.shape != tuple(cfg.shape):
            raise LbmError("design map and lattice shapes differ")
        self.cfg, self.steps, self.map, self.scales = cfg, int(steps), design_map, tuple(scales)
        self.alpha_max, self.q = float(alpha_max), float(q)
        self.lattice = Lattice(cfg)

    @classmethod
    def from_problem(cls, problem, precision="f64", **config):
        """From a JSON problem dict (normalised here) or a normalised LbmProblem, as
        LatticeBoltzmannProvider::parts does (provider.rs:249-262); `config` overrides
        fields of the lattice configuration (adjoint storage, checkpointing)."""
        prob = problem if isinstance(problem, LbmProblem) else normalize(problem)
        dm = prob.design_map()
        try:
            out = cls(prob.lattice_config(precision, **config), prob.step_count, dm, prob.vector_scales(),
                      alpha_max=prob.alpha_max, q=prob.drag_shape)
        except Exception:
            dm.close()
            raise
        out.problem, out._owns_map = prob, True
        return out

    def close(self):
        self.lattice.close()
        if getattr(self, "_owns_map", False):
            self.map.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _parts(self, raw):
        phi = self.map.map(raw)
        alpha = design.resistance(phi, self.alpha_max, self.q)
        self.lattice.set_design(alpha)
        self.lattice.initial_populations()
        self.lattice.history(self.steps)
        return phi, alpha

    def evaluate(self, raw):
        """Responses of the final state (evaluate_value's `responses`)."""
        phi, _ = self._parts(raw)
        return dict(zip(RESPONSES, self.lattice.vector(self.scales, phi)))

    def sensitivities(self, raw, responses=RESPONSES):
        """(values, gradients): per response name its value and d value / d raw."""
        if not responses or len(set(responses)) != len(responses) or any(r not in RESPONSES for r in responses):
            raise LbmError("Unknown, duplicated or empty LBM responses")
        raw = np.ascontiguousarray(raw, np.float64).reshape(-1)
        phi, _ = self._parts(raw)
        values = self.lattice.vector(self.scales, phi)
        dres = design.resistance_derivative(phi, self.alpha_max, self.q)
        out_v, out_g = {}, {}
        for name in responses:
            w = np.zeros(len(RESPONSES))
            w[RESPONSES.index(name)] = 1.0
            f_bar, ab_final, phi_bar = self.lattice.vector_vjp(self.scales, w, phi)
            self.lattice.initial_populations()
            _, ab, _, _ = self.lattice.gradient(self.steps, None, lambda_final=f_bar)
            phi_bar = phi_bar + (ab + ab_final) * dres
            g = self.map.vjp(raw, phi_bar)
            if not np.all(np.isfinite(g)):
                raise LbmError("Nonfinite LBM sensitivity")
            out_v[name] = float(values[RESPONSES.index(name)])
            out_g[name] = g
        return out_v, out_g
