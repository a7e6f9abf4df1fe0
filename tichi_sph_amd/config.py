"""Solver configuration: the reference's `Config` (sph_config.py:83-232).

Reads the same two JSON documents (solver config and scenario) with the same
keys, and derives the same constants with the same rounding points: values
the reference stores in ti.f32 fields are rounded to float32, values it
computes with Python floats are computed in float64 first
(sph_config.py:145-224).
"""
import json
import math

import numpy as np

from . import _lib


def read_param(param, paramname):
    """sph_config.py:13-16: a missing or null key raises."""
    if param is None:
        raise Exception("The parameter " + paramname + " is missing, or invalid parameter")
    return param


def _f32(x):
    return float(np.float32(x))


def r2_in_range_max(h, dtype):
    """Largest r^2 (in `dtype`) for which the reference's kernel support test
    fl(fl(sqrt(r^2)) / h) < 1 holds (sph_util.py:18-23: q = r / h, q < 1).

    sqrt and division are monotone and correctly rounded, so the in-range set
    is exactly {r^2 <= r2max}: the GPU tests r^2 against this threshold and
    classifies every pair as the reference's arithmetic does."""
    dt = np.dtype(dtype).type
    h = dt(h)
    one = dt(1)
    inside = lambda x: (np.sqrt(dt(x)) / h) < one  # noqa: E731
    x = dt(h) * dt(h)
    while not inside(x):
        x = np.nextafter(x, dt(0))
    while inside(np.nextafter(x, dt(np.inf))):
        x = np.nextafter(x, dt(np.inf))
    return float(x)


class Field0D(float):
    """A float the reference keeps in a 0-D field: `config.x[None]` reads it
    (sph_config.py: ti.field(float, ())); otherwise it is a plain float."""

    def __getitem__(self, k):
        if k is None:
            return float(self)
        raise TypeError("a 0-D field is indexed with [None]")


class IntField0D(int):
    """An int the reference keeps in a 0-D field (`config.x[None]`)."""

    def __getitem__(self, k):
        if k is None:
            return int(self)
        raise TypeError("a 0-D field is indexed with [None]")


# Config values the reference reads as 0-D fields from the run loop
# (main.py:9-10, 199-234, 348; sph_config.py:145-194)
_FIELDS_F = ("divergence_threshold", "compression_threshold", "dt", "cfl_factor", "dynamic_viscosity",
             "surface_tension_gamma", "fbm_diffusion_term", "fbm_convection_term")
_FIELDS_I = ("fluid_max_part_num", "bound_max_part_num", "iter_threshold_min", "iter_threshold_max", "cs",
             "wc_gamma", "node_num", "gui_fps")


class Config:
    """Constants of one scene, from config.json + scenario.json."""

    def __init__(self, config_buffer, scenario_buffer):
        self.config_buffer = config_buffer
        self.scenario_buffer = scenario_buffer
        env = scenario_buffer["sim_env"]
        # module globals, sph_config.py:77-80
        self.sim_dim = int(read_param(env["sim_dim"], "sim_dim"))
        self.phase_num = int(read_param(env["phase_num"], "phase_num"))
        self.solver_type = read_param(config_buffer.get("solver_type"), "solver_type")
        self.neighb_range = int(read_param(config_buffer.get("solver_neighb_range"), "solver_neighb_range"))
        if self.solver_type not in ("DFSPH", "VFSPH"):
            raise Exception("solver_type must be DFSPH or VFSPH, got " + str(self.solver_type))
        if self.sim_dim not in (2, 3):
            raise Exception("sim_dim must be 2 or 3")
        if not 1 <= self.phase_num <= _lib.TSPH_MAX_PHASE:
            raise Exception("phase_num must be 1..%d" % _lib.TSPH_MAX_PHASE)
        self.arch = config_buffer.get("arch")
        self.device_memory_GB = config_buffer.get("device_memory_GB")
        self.init_sim_env()
        self.init_solver()
        self.init_gui()
        self.calculate_kernel_param()
        self.init_neighb_param()
        self.calculate_neighb_param()
        for name in _FIELDS_F:
            setattr(self, name, Field0D(getattr(self, name)))
        for name in _FIELDS_I:
            if getattr(self, name, None) is not None:
                setattr(self, name, IntField0D(getattr(self, name)))

    @classmethod
    def from_files(cls, config_file, scenario_file):
        with open(config_file) as f:
            cb = json.load(f)
        with open(scenario_file) as f:
            sb = json.load(f)
        return cls(cb, sb)

    # sph_config.py:145-163
    def init_sim_env(self):
        env = self.scenario_buffer["sim_env"]
        d = self.sim_dim
        self.dim = d
        self.part_size = [0.0] * 5
        self.part_size[1] = _f32(read_param(env["global_part_size"], "global_part_size"))
        self.dynamic_viscosity = _f32(read_param(env["global_dynamic_viscosity"], "global_dynamic_viscosity"))
        self.sim_space_lb = [_f32(v) for v in read_param(env["sim_space_lb"], "sim_space_lb")]
        self.sim_space_rt = [_f32(v) for v in read_param(env["sim_space_rt"], "sim_space_rt")]
        self.gravity = [_f32(v) for v in read_param(env["gravity"], "gravity")]
        if not (len(self.sim_space_lb) == len(self.sim_space_rt) == len(self.gravity) == d):
            raise Exception("sim_space_lb / sim_space_rt / gravity must have sim_dim entries")
        rho = read_param(env["phase_rest_density"], "phase_rest_density")
        if len(rho) != self.phase_num:
            raise Exception("phase_rest_density must have phase_num entries")
        self.phase_rest_density = [_f32(v) for v in rho]
        # Extension (not in the reference, which has one global mu,
        # sph_config.py:148): per-phase dynamic viscosities; the viscous force
        # of particle i then uses mu_i = sum_k volume_frac_ik * mu_k.
        mus = env.get("phase_dynamic_viscosity")
        if mus is not None and len(mus) != self.phase_num:
            raise Exception("phase_dynamic_viscosity must have phase_num entries")
        self.phase_dynamic_viscosity = None if mus is None else [_f32(v) for v in mus]
        self.fluid_max_part_num = int(read_param(self.scenario_buffer["fluid"]["max_part_num"], "fluid_max_part_num"))
        self.bound_max_part_num = int(read_param(self.scenario_buffer["bound"]["max_part_num"], "bound_max_part_num"))
        hexes = env.get("phase_color_hex", ["0x6F7DBC"] * self.phase_num)
        self.phase_rgb = []
        for i in range(self.phase_num):
            hx = int(hexes[i], 16)
            self.phase_rgb.append([((hx & 0xFF0000) >> 16) / 255.0, ((hx & 0x00FF00) >> 8) / 255.0, (hx & 0x0000FF) / 255.0])
        self.part_size[2] = _f32(math.pow(self.part_size[1], 2))
        self.part_size[3] = _f32(math.pow(self.part_size[1], 3))
        self.part_size[4] = _f32(math.pow(self.part_size[1], 4))

    # sph_config.py:165-185
    def init_solver(self):
        cb = self.config_buffer
        self.wc_gamma = int(read_param(cb.get("solver_wc_gamma"), "solver_wc_gamma"))
        self.divergence_threshold = _f32(read_param(cb.get("solver_divergence_threshold"), "solver_divergence_threshold"))
        self.compression_threshold = _f32(read_param(cb.get("solver_compression_threshold"), "solver_compression_threshold"))
        self.iter_threshold_min = int(read_param(cb.get("solver_iter_threshold_min"), "solver_iter_threshold_min"))
        self.iter_threshold_max = int(read_param(cb.get("solver_iter_threshold_max"), "solver_iter_threshold_max"))
        self.if_cfl = int(read_param(cb.get("solver_if_cfl"), "solver_if_cfl"))
        self.cs = int(read_param(cb.get("solver_cs"), "solver_cs"))
        self.cfl_factor = _f32(read_param(cb.get("solver_cfl_factor"), "solver_cfl_factor"))
        self.neighb_grid_size_TO_global_part_size = int(read_param(
            cb.get("solver_neighb_grid_size_TO_global_part_size"), "solver_neighb_grid_size_TO_global_part_size"))
        self.fbm_diffusion_term = _f32(read_param(cb.get("solver_fbm_diffusion_term"), "solver_fbm_diffusion_term"))
        self.fbm_convection_term = _f32(read_param(cb.get("solver_fbm_convection_term"), "solver_fbm_convection_term"))
        self.surface_tension_gamma = _f32(read_param(cb.get("solver_surface_tension_gamma"), "solver_surface_tension_gamma"))
        self.dt = _f32(self.part_size[1] / float(np.float32(self.cs))) if self.cs else 0.0

    # sph_config.py:187-194 (GUI keys are read for drop-in compatibility only)
    def init_gui(self):
        cb = self.config_buffer
        self.gui_res = cb.get("gui_res")
        self.gui_part_zoomer = cb.get("gui_part_zoomer")
        self.gui_fps = cb.get("gui_fps")
        self.gui_canvas_bgcolor = cb.get("gui_canvas_bgcolor")

    # sph_config.py:196-202
    def calculate_kernel_param(self):
        self.kernel_h = [0.0] * 5
        self.kernel_sig = [0.0] * 4
        self.kernel_h[1] = _f32(self.part_size[1] * 2)
        self.kernel_h[2] = _f32(math.pow(self.kernel_h[1], 2))
        self.kernel_h[3] = _f32(math.pow(self.kernel_h[1], 3))
        self.kernel_h[4] = _f32(math.pow(self.kernel_h[1], 4))
        self.kernel_sig[3] = _f32(8 / math.pi / self.kernel_h[1] ** 3)
        self.kernel_sig[2] = _f32(40 / 7 / math.pi / self.kernel_h[1] ** 2)

    # sph_config.py:204-208: ceil((rt - lb) / h) in f32 inside a Taichi kernel
    def init_neighb_param(self):
        f = np.float32
        self.neighb_grid_size = [
            float(np.ceil((f(self.sim_space_rt[i]) - f(self.sim_space_lb[i])) / f(self.kernel_h[1])))
            for i in range(self.dim)]

    # sph_config.py:210-224
    def calculate_neighb_param(self):
        self.node_num = 1
        self.neighb_grid_coder = [1] * self.dim
        for i in range(self.dim):
            self.neighb_grid_coder[i] = self.node_num
            self.node_num *= int(self.neighb_grid_size[i])
        side = self.neighb_range * 2 + 1
        dice = [-self.neighb_range + i for i in range(side)]
        self.neighb_search_template = []
        for j in range(side ** self.dim):
            tmp = j
            cell = [0] * self.dim
            for d in range(self.dim):
                digit = tmp // (side ** (self.dim - d - 1))
                tmp = tmp % (side ** (self.dim - d - 1))
                cell[self.dim - d - 1] = dice[digit]
            self.neighb_search_template.append(cell)
        self.r2_max32 = r2_in_range_max(self.kernel_h[1], np.float32)
        self.r2_max64 = r2_in_range_max(self.kernel_h[1], np.float64)

    def params(self, precision=4, list_cap_f=64, list_cap_b=48):
        """The C parameter block (include/tichi_sph.h TsphParams)."""
        p = _lib.TsphParams()
        p.dim = self.dim
        p.phase_num = self.phase_num
        p.solver_type = _lib.TSPH_VFSPH if self.solver_type == "VFSPH" else _lib.TSPH_DFSPH
        p.precision = precision
        p.neighb_range = self.neighb_range
        p.ntemp = len(self.neighb_search_template)
        p.node_num = self.node_num
        p.list_cap_f = list_cap_f
        p.list_cap_b = list_cap_b
        for d in range(self.dim):
            p.grid_size[d] = int(self.neighb_grid_size[d])
            p.coder[d] = self.neighb_grid_coder[d]
            p.lb[d] = self.sim_space_lb[d]
            p.rt[d] = self.sim_space_rt[d]
            p.gravity[d] = self.gravity[d]
        p.cs = self.cs
        p.wc_gamma = self.wc_gamma
        p.iter_min = self.iter_threshold_min
        p.iter_max = self.iter_threshold_max
        for k in range(self.phase_num):
            p.phase_rest_density[k] = self.phase_rest_density[k]
        for k in range(5):
            p.part_size[k] = self.part_size[k]
            p.kernel_h[k] = self.kernel_h[k]
        for k in range(4):
            p.kernel_sig[k] = self.kernel_sig[k]
        p.r2_max = self.r2_max32 if precision == 4 else self.r2_max64
        p.dynamic_viscosity = self.dynamic_viscosity
        if self.phase_dynamic_viscosity is not None:
            p.mu_per_phase = 1
            for k, v in enumerate(self.phase_dynamic_viscosity):
                p.phase_viscosity[k] = v
        p.surface_tension_gamma = self.surface_tension_gamma
        p.fbm_diffusion = self.fbm_diffusion_term
        p.fbm_convection = self.fbm_convection_term
        p.cfl_factor = self.cfl_factor
        p.divergence_threshold = self.divergence_threshold
        p.compression_threshold = self.compression_threshold
        return p
