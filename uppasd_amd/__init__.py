"""uppasd_amd -- MI355X (gfx950) engine for UppASD's spin-update hot path.

Effective field (exchange, DMI, uniaxial/cubic anisotropy, Zeeman), the
Depondt and semi-implicit midpoint stochastic LLG steps and coloured
Metropolis / heat-bath Monte Carlo sweeps, as hand-written HIP kernels
behind the C-ABI of ``libuasd_hip.so`` (include/uasd_hip.h).
"""
from .engine import (Engine, UasdError, moment_update, device_count, set_tune_cache,  # noqa: F401
                     UASD_NOISE_NONE, UASD_NOISE_PHILOX, UASD_NOISE_INJECTED,
                     UASD_SDE_MIDPOINT, UASD_SDE_HEUN, UASD_SDE_DEPONDT, UASD_SDE_SIB,
                     UASD_SDE_SPHMID)
from .lattice import Lattice, Shells, build_pair_list, bcc, triangular, kagome  # noqa: F401
