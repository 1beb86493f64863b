"""Design fields and the Brinkman coefficient (host side).

``resistance`` / ``resistance_derivative`` restate crates/implexity-physics-lbm/src/
design.rs:282-294 with step_s * drag_max_per_s = alpha_max:
    alpha(phi) = alpha_max * q * (1 - phi) / (q + phi).
The synthetic design fields are the workload generators of the benchmark problems:
i.i.d. U(0,1) from a seed, Gaussian smoothing (periodic axes wrap, walls reflect),
then a sigmoid projection; or a sum of random Gaussian blobs, smoothed and sigmoided.
"""
from __future__ import annotations

import numpy as np
from scipy import ndimage


def resistance(phi, alpha_max=0.1, q=0.1):
    phi = np.asarray(phi, np.float64)
    return alpha_max * q * (1.0 - phi) / (q + phi)


def resistance_derivative(phi, alpha_max=0.1, q=0.1):
    phi = np.asarray(phi, np.float64)
    d = q + phi
    return alpha_max * q * (-d - (1.0 - phi)) / (d * d)


def sigmoid(s, slope=10.0, center=0.5):
    return 1.0 / (1.0 + np.exp(-slope * (s - center)))


def smoothed_random_design(shape, sigma, seed, periodic=(True, False, False), slope=10.0):
    """U(0,1) noise, Gaussian smoothing with standard deviation sigma (cells), sigmoid at
    0.5 with the given slope.  The smoothed field is rescaled to [0,1] first so the
    sigmoid sees the full range (smoothing shrinks the variance of white noise)."""
    rng = np.random.default_rng(seed)
    raw = rng.random(shape)
    modes = ["wrap" if p else "reflect" for p in periodic]
    s = raw
    for axis in range(3):
        s = ndimage.gaussian_filter1d(s, sigma, axis=axis, mode=modes[axis])
    lo, hi = s.min(), s.max()
    s = (s - lo) / (hi - lo) if hi > lo else np.full_like(s, 0.5)
    return sigmoid(s, slope)


def blob_design(shape, n_blobs, seed, radius=(6.0, 20.0), sigma=2.0, slope=10.0, periodic=(False, False, False)):
    """Sum of n random Gaussian blobs (centres uniform, radii U(radius)), smoothed,
    normalised to [0,1] and sigmoided: high values are fluid, blobs carve solid."""
    rng = np.random.default_rng(seed)
    grids = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")
    field_ = np.zeros(shape)
    for _ in range(n_blobs):
        c = [rng.uniform(0, n) for n in shape]
        r = rng.uniform(*radius)
        d2 = sum((g - ci) ** 2 for g, ci in zip(grids, c))
        field_ += np.exp(-d2 / (2.0 * r * r))
    modes = ["wrap" if p else "reflect" for p in periodic]
    for axis in range(3):
        field_ = ndimage.gaussian_filter1d(field_, sigma, axis=axis, mode=modes[axis])
    lo, hi = field_.min(), field_.max()
    s = (field_ - lo) / (hi - lo) if hi > lo else np.zeros_like(field_)
    return sigmoid(1.0 - s, slope)


def parabolic_profile(ny, nz):
    """Duct inlet profile with walls at the half links: 16 s(1-s) t(1-t), s = (j+1/2)/ny,
    t = (k+1/2)/nz; 1 at the centre.  Ordered like the face's cells (j slow, k fast)."""
    s = (np.arange(ny) + 0.5) / ny
    t = (np.arange(nz) + 0.5) / nz
    return (4.0 * s * (1.0 - s))[:, None] * (4.0 * t * (1.0 - t))[None, :]
