"""MI355X (gfx950) implementation of Tichi_SPH's per-step particle hot path.

Hand-written HIP kernels (tichi_sph_amd/csrc) behind the C ABI of
include/tichi_sph.h, driven from Python through ctypes on torch device
tensors.  The Python surface keeps the reference's names: Config,
Fluid, Ngrid, NeighborSearch, the sph.py kernel functions and sph_step().
"""
from .config import Config, read_param  # noqa: F401
from .sph_obj import Fluid, Ngrid  # noqa: F401
from .solver import (EventTimer, GlobalVariable, Simulation, StepStats, current_simulation,  # noqa: F401
                     inlet_helper_get_dt, load_simulation, simulation_from_dicts, use_config)
from .NeighborSearch import NeighborSearch  # noqa: F401
from .grid import Grid  # noqa: F401
