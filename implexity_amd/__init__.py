"""MI355X-native differentiable lattice-Boltzmann hot path (D3Q19/D3Q27, TRT/MRT/
cumulant, halfway bounce-back, Brinkman design coupling, discrete adjoint).

The compute path is the HIP library ``libimplexity_lbm.so`` behind the C ABI in
``include/implexity_lbm.h``; this package is its Python binding and host-side helpers.
"""
from .lbm import (  # noqa: F401
    BinomialSchedule,
    ContractError,
    Lattice,
    LatticeConfig,
    LbmError,
    Objective,
    Port,
    copy_bandwidth,
    lib,
    odd_rate,
    optimal_forward_steps,
)
from . import design  # noqa: F401
