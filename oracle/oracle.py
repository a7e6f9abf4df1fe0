"""ctypes loader of the CPU oracle (oracle/sph_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the tichi_sph_amd package.

The oracle derives its constants (sph_config.py:145-232) and its initial
particle state (sph_obj.py:102-180, cube/box objects) from the raw JSON
documents itself, independently of the product's Config and scene builders.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
MAXP = 8

STEP_CFL, STEP_FBM, STEP_WCSPH = 1, 2, 4


class RawInit(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int), ("phase_num", ctypes.c_int), ("solver_type", ctypes.c_int),
        ("neighb_range", ctypes.c_int),
        ("global_part_size", ctypes.c_double), ("dynamic_viscosity", ctypes.c_double),
        ("lb", ctypes.c_double * 3), ("rt", ctypes.c_double * 3), ("gravity", ctypes.c_double * 3),
        ("phase_rest_density", ctypes.c_double * MAXP),
        ("fluid_max", ctypes.c_int), ("bound_max", ctypes.c_int), ("wc_gamma", ctypes.c_int),
        ("divergence_threshold", ctypes.c_double), ("compression_threshold", ctypes.c_double),
        ("iter_min", ctypes.c_int), ("iter_max", ctypes.c_int), ("cs", ctypes.c_int),
        ("cfl_factor", ctypes.c_double),
        ("fbm_diffusion", ctypes.c_double), ("fbm_convection", ctypes.c_double),
        ("surface_tension_gamma", ctypes.c_double),
        ("mu_per_phase", ctypes.c_int), ("phase_viscosity", ctypes.c_double * MAXP),
    ]


class OracleConsts(ctypes.Structure):
    _fields_ = [
        ("part_size", ctypes.c_double * 5), ("kernel_h", ctypes.c_double * 5),
        ("kernel_sig", ctypes.c_double * 4), ("grid_size", ctypes.c_double * 3),
        ("coder", ctypes.c_int * 3), ("node_num", ctypes.c_int), ("ntemp", ctypes.c_int),
        ("tmpl", (ctypes.c_int * 3) * 125),
    ]


class StepStats(ctypes.Structure):
    _fields_ = [
        ("div_iter", ctypes.c_int), ("incom_iter", ctypes.c_int), ("fbm_rounds", ctypes.c_int),
        ("pad", ctypes.c_int), ("dt", ctypes.c_double), ("ke", ctypes.c_double), ("pe", ctypes.c_double),
        ("div_err", ctypes.c_double * 128), ("incom_err", ctypes.c_double * 128),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.check_call(["make", "-s", "-C", HERE])
        L = ctypes.CDLL(LIB)
        L.oracle_new.restype = ctypes.c_void_p
        L.oracle_new.argtypes = [ctypes.POINTER(RawInit), ctypes.c_int]
        L.oracle_free.argtypes = [ctypes.c_void_p]
        L.oracle_set_part_num.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.oracle_set.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_long]
        L.oracle_get.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_long]
        L.oracle_scalar.restype = ctypes.c_double
        L.oracle_scalar.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.oracle_kernel.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
        L.oracle_sph_step.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(StepStats)]
        L.oracle_consts.argtypes = [ctypes.POINTER(RawInit), ctypes.POINTER(OracleConsts)]
        L.oracle_map_velocity.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_double,
                                          ctypes.c_void_p]
        L.oracle_kernel_fn.restype = ctypes.c_double
        L.oracle_kernel_fn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.oracle_lattice.restype = ctypes.c_long
        L.oracle_lattice.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                     ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                     ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        L.oracle_ns_total.restype = ctypes.c_long
        L.oracle_ns_total.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.oracle_ns_get.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 4
        L.oracle_set_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def raw_init(config_buffer, scenario_buffer):
    env = scenario_buffer["sim_env"]
    r = RawInit()
    r.dim = int(env["sim_dim"])
    r.phase_num = int(env["phase_num"])
    r.solver_type = 1 if config_buffer["solver_type"] == "VFSPH" else 0
    r.neighb_range = int(config_buffer["solver_neighb_range"])
    r.global_part_size = env["global_part_size"]
    r.dynamic_viscosity = env["global_dynamic_viscosity"]
    for d in range(r.dim):
        r.lb[d] = env["sim_space_lb"][d]
        r.rt[d] = env["sim_space_rt"][d]
        r.gravity[d] = env["gravity"][d]
    for k in range(r.phase_num):
        r.phase_rest_density[k] = env["phase_rest_density"][k]
    r.fluid_max = int(scenario_buffer["fluid"]["max_part_num"])
    r.bound_max = int(scenario_buffer["bound"]["max_part_num"])
    cb = config_buffer
    r.wc_gamma = int(cb["solver_wc_gamma"])
    r.divergence_threshold = cb["solver_divergence_threshold"]
    r.compression_threshold = cb["solver_compression_threshold"]
    r.iter_min = int(cb["solver_iter_threshold_min"])
    r.iter_max = int(cb["solver_iter_threshold_max"])
    r.cs = int(cb["solver_cs"])
    r.cfl_factor = cb["solver_cfl_factor"]
    r.fbm_diffusion = cb["solver_fbm_diffusion_term"]
    r.fbm_convection = cb["solver_fbm_convection_term"]
    r.surface_tension_gamma = cb["solver_surface_tension_gamma"]
    mus = env.get("phase_dynamic_viscosity")  # the per-phase viscosity extension
    if mus is not None:
        r.mu_per_phase = 1
        for k in range(r.phase_num):
            r.phase_viscosity[k] = mus[k]
    return r


def consts(config_buffer, scenario_buffer):
    c = OracleConsts()
    lib().oracle_consts(ctypes.byref(raw_init(config_buffer, scenario_buffer)), ctypes.byref(c))
    return c


def lattice(dim, part_size_f32, relaxing, start, end, mode=0, layers=0):
    """Initial positions of a cube (mode 0) or hollow box (mode 1) object."""
    st = (ctypes.c_double * 3)(*[float(v) for v in start])
    en = (ctypes.c_double * 3)(*[float(v) for v in end])
    n = lib().oracle_lattice(dim, part_size_f32, relaxing, st, en, mode, layers, None)
    out = np.zeros((n, dim), dtype=np.float64)
    lib().oracle_lattice(dim, part_size_f32, relaxing, st, en, mode, layers, out.ctypes.data)
    return out


class Oracle:
    FIELDS_VEC = ("pos", "vel", "vel_adv", "acce", "acce_adv", "alpha_1", "normal", "W_grad",
                  "pressure_force")

    def __init__(self, config_buffer, scenario_buffer, precision=4, fluid_max=None, bound_max=None):
        self.cb, self.sb = config_buffer, scenario_buffer
        self.raw = raw_init(config_buffer, scenario_buffer)
        if fluid_max is not None:
            self.raw.fluid_max = fluid_max
        if bound_max is not None:
            self.raw.bound_max = bound_max
        self.precision = precision
        self.dtype = np.float32 if precision == 4 else np.float64
        self.h = lib().oracle_new(ctypes.byref(self.raw), precision)
        self.dim = self.raw.dim
        self.P = self.raw.phase_num
        self.n = [0, 0]
        self.c = consts(config_buffer, scenario_buffer)

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_free(self.h)
            self.h = None

    def _width(self, name):
        if name in self.FIELDS_VEC:
            return self.dim
        if name in ("volume_frac", "volume_frac_tmp"):
            return self.P
        if name == "drift_vel":
            return self.P * self.dim
        if name == "neighb_cell_structured_seq":
            return self.dim
        return 1

    INT = ("flag", "neighb_cell_seq", "neighb_in_cell_seq", "neighb_cell_structured_seq")

    def set_part_num(self, obj, n):
        assert lib().oracle_set_part_num(self.h, obj, n) == 0
        self.n[obj] = n

    def set(self, obj, name, arr):
        dt = np.int32 if name in self.INT else self.dtype
        a = np.ascontiguousarray(np.asarray(arr, dtype=dt).reshape(-1))
        rc = lib().oracle_set(self.h, obj, name.encode(), a.ctypes.data, a.size)
        assert rc == 0, (name, rc)

    def get(self, obj, name):
        w = self._width(name)
        dt = np.int32 if name in self.INT else self.dtype
        n = self.n[obj]
        out = np.zeros(n * w, dtype=dt)
        rc = lib().oracle_get(self.h, obj, name.encode(), out.ctypes.data, out.size)
        assert rc == 0, (name, rc)
        if name == "drift_vel":
            return out.reshape(n, self.P, self.dim)
        return out.reshape(n, w) if w > 1 else out

    def grid(self, name, count):
        out = np.zeros(count, dtype=np.int32)
        assert lib().oracle_get(self.h, 2, name.encode(), out.ctypes.data, count) == 0
        return out

    def scalar(self, which, obj=0, value=None):
        idx = {"compression": 0, "dt": 1, "general_flag": 2, "ke": 3, "pe": 4}[which]
        return lib().oracle_scalar(self.h, obj, idx, 0 if value is None else 1, 0.0 if value is None else value)

    def kernel(self, name, obj, nobj=-1):
        assert lib().oracle_kernel(self.h, name.encode(), obj, nobj) == 0, name

    def map_velocity(self, obj, shape, dist):
        """map_velocity (sph.py:382-397) of object `obj` over a Grid of
        `shape` nodes spaced `dist` from sim_space_lb: [prod(shape), dim]."""
        shape = np.ascontiguousarray(shape, dtype=np.int32)
        out = np.zeros((int(np.prod(shape)), len(shape)), dtype=self.dtype)
        lib().oracle_map_velocity(self.h, obj, shape.ctypes.data, float(dist), out.ctypes.data)
        return out

    def step(self, flags=STEP_CFL):
        st = StepStats()
        lib().oracle_sph_step(self.h, flags, ctypes.byref(st))
        return st

    # ---- initial state from the scenario (sph_obj.py:102-180, main.py:62-79)
    def init_scenario(self):
        env = self.sb["sim_env"]
        gps32 = float(np.float32(env["global_part_size"]))
        vol = float(np.float32(pow(gps32, self.dim)))  # part_size[dim], sph_config.py:161-163
        rho = [np.float32(v) for v in env["phase_rest_density"]]
        for obj, part in ((0, "fluid"), (1, "bound")):
            pos_l, vel_l, vf_l = [], [], []
            for p in self.sb[part]["objs"]:
                if p.get("time", 0) > 0:
                    continue
                if p["type"] == "cube":
                    x = lattice(self.dim, gps32, p["particle_relaxing_factor"], p["start_pos"], p["end_pos"], 0)
                elif p["type"] == "box":
                    x = lattice(self.dim, gps32, p["particle_relaxing_factor"], p["start_pos"], p["end_pos"], 1,
                                p["layers"])
                else:
                    raise NotImplementedError(p["type"])
                pos_l.append(x)
                vel_l.append(np.tile(np.asarray(p["vel"], dtype=np.float32), (len(x), 1)))
                vf_l.append(np.tile(np.asarray(p["volume_frac"], dtype=np.float32), (len(x), 1)))
            if not pos_l:  # an object list with no particles (main.py:62-79 adds nothing)
                self.set_part_num(obj, 0)
                continue
            pos = np.concatenate(pos_l).astype(np.float32)
            n = len(pos)
            self.set_part_num(obj, n)
            self.set(obj, "pos", pos)
            self.set(obj, "vel", np.concatenate(vel_l))
            vf = np.concatenate(vf_l).astype(np.float32)
            self.set(obj, "volume_frac", vf)
            rv = np.full(n, vol, dtype=np.float32)
            self.set(obj, "rest_volume", rv)
            rd = rho[0] * vf[:, 0]
            for k in range(1, self.P):
                rd = (rd + rho[k] * vf[:, k]).astype(np.float32)
            self.set(obj, "rest_density", rd)
            self.set(obj, "mass", (rd * rv).astype(np.float32))

    def state_from(self, obj, fields):
        """Load a state dict of numpy arrays (reference order)."""
        for k, v in fields.items():
            self.set(obj, k, v)
