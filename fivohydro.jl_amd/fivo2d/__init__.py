"""fivo2d -- MI355X-native 2+1D finite-volume hydro core (FiVoHydro.jl's src2d hot path).

The compute runs in hand-written FP64 HIP kernels for gfx950 behind a C ABI
(include/fivo2d.h, csrc/); this package is the host-side mirror of the
reference's Julia API over that ABI.
"""
from ._abi import (FIVO_BC_OUTFLOW, FIVO_BC_PERIODIC, FIVO_SSPRK2, FIVO_SSPRK3, Lib, load_hip,
                   HIP_LIB_PATH)
from .model import (ConformalHQEOS, LatticeHRGEOS, Grid2D, make_grid2d, make_grid_2d,
                    make_layout2d, build_model_2d, allocate_state, resolve_terms, without, check_terms,
                    TERM_REGISTER, T_MIN, E_FLOOR,
                    HBARC)
from .solver import Solver, FivoError
from .batch import Ensemble
from .mlp import Mlp
from .background import export_background2d, export_backgrounds2d, load_background
from .fields_io import save_fields, load_fields
from . import ic
from . import freezeout
from .freezeout import FreezeOut, freeze_surface, pion_yield

__all__ = ["FIVO_BC_OUTFLOW", "FIVO_BC_PERIODIC", "FIVO_SSPRK2", "FIVO_SSPRK3", "Lib", "load_hip",
           "HIP_LIB_PATH", "ConformalHQEOS", "LatticeHRGEOS", "Grid2D", "make_grid2d",
           "make_grid_2d", "make_layout2d", "build_model_2d", "allocate_state", "resolve_terms", "without", "check_terms", "TERM_REGISTER",
           "Solver", "FivoError", "Ensemble", "Mlp", "export_background2d", "export_backgrounds2d", "load_background", "ic", "T_MIN", "E_FLOOR", "HBARC", "save_fields", "load_fields",
           "freezeout", "FreezeOut", "freeze_surface", "pion_yield"]
