"""MI355X (gfx950) H/C transport step of MantlePar: ML-WENO reconstruction,
Lax-Friedrichs face fluxes and SSP Runge-Kutta stages fused with the binary
eutectic phase equilibrium, behind the C ABI of include/mantlepar_amd.h."""
from ._lib import LIB_PATH, MantleError, lib  # noqa: F401
from .transport import (BC_INFLOW_OUTFLOW, BC_OUTFLOW, BC_ZERO_FLUX, PREC_FP64,  # noqa: F401
                        PREC_MIXED, SCHEME_FE, SCHEME_SSPRK2, SCHEME_SSPRK3, EutecticModel,
                        Transport, default_phase, default_weno)
from .batch import BATCH_EAGER, BATCH_GRAPH, Batch, Server  # noqa: F401,E402
from .coupled import Coupled, default_coupled, run_batch as coupled_run_batch  # noqa: F401,E402
