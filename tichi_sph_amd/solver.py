"""The per-step driver: the reference's `sph_step()` (main.py:175-256).

`Simulation.sph_step()` is the fused MI355X path: one counting sort, one
pass that computes density / alpha and records neighbour lists, predicated
Jacobi iterations whose stop state lives on the device, advection, optional
FBM rounds and the tail.  The host only reads the 8-byte stop state once per
launched chunk of iterations (the reference reads `fluid.compression` after
every iteration, main.py:201 / 222).

`Simulation.sph_step_reference()` issues the reference's kernels one by one
in main.py's order through the kernel surface in `sph.py`; it exists for
drop-in use and as a parity anchor.
"""
import ctypes
import json
import sys
import warnings

import numpy as np
import torch

from . import _lib, binding
from .config import Config
from .sph_obj import Fluid, Ngrid


class StepStats:
    def __init__(self, raw, precision):
        self.div_iter = raw.div_iter
        self.incom_iter = raw.incom_iter
        self.fbm_rounds = raw.fbm_rounds
        self.dt = raw.dt
        self.compression = raw.compression
        self.kinetic_energy = raw.ke
        self.gravity_potential_energy = raw.pe
        self.div_err = list(raw.err_div[: min(raw.div_iter, _lib.TSPH_MAX_ITER_LOG)])
        self.incom_err = list(raw.err_incom[: min(raw.incom_iter, _lib.TSPH_MAX_ITER_LOG)])
        self.overflow_f = raw.overflow_f
        self.overflow_b = raw.overflow_b

    def __repr__(self):
        return (f"StepStats(div_iter={self.div_iter}, incom_iter={self.incom_iter}, dt={self.dt:.6g}, "
                f"compression={self.compression:.4g})")


class EventTimer:
    """Per-call HIP-event timing: each timed library call records a start and
    an end event on the launch stream.  Events come from a pool created up
    front (`reserve`), so timing adds no event creation inside a timed loop.
    With `only` (a set of call names) the other calls are not timed."""

    def __init__(self, reserve=0, only=None):
        self.pool = []
        self.used = 0
        self.records = []  # (name, start event, end event)
        self.only = None if only is None else set(only)
        self.reserve(reserve)

    def wants(self, name):
        return self.only is None or name in self.only

    def reserve(self, n):
        while len(self.pool) < n:
            self.pool.append(torch.cuda.Event(enable_timing=True))

    def pair(self):
        if self.used + 2 > len(self.pool):
            self.reserve(self.used + 256)
        e0, e1 = self.pool[self.used], self.pool[self.used + 1]
        self.used += 2
        return e0, e1

    def per_name(self):
        """{name: [total ms, calls]} (synchronises on the recorded events)."""
        if self.records:
            self.records[-1][2].synchronize()
        out = {}
        for name, e0, e1 in self.records:
            d = out.setdefault(name, [0.0, 0])
            d[0] += e0.elapsed_time(e1)
            d[1] += 1
        return out


_FASTQ_CACHE = {}


def _fast_quotient_ok(lib, params, device):
    """Whether pass A's reduced sqrt and r / h sequences give bit-identical
    W and r > 0 for these parameters (tsph_verify_fast_quotient sweeps every
    FP32 r^2 in [0, r2_max] on the device, about 2 ms); cached per kernel
    constants."""
    key = (params.dim, tuple(params.kernel_h), tuple(params.kernel_sig), params.r2_max)
    if key not in _FASTQ_CACHE:
        out = torch.zeros(3, dtype=torch.int32, device=device)
        st = ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
        _lib.check(lib.tsph_verify_fast_quotient(ctypes.byref(params), ctypes.c_void_p(out.data_ptr()), st),
                   "tsph_verify_fast_quotient")
        _FASTQ_CACHE[key] = int(out[0].item()) == 0
    return _FASTQ_CACHE[key]


class GlobalVariable:
    """The run-time counters of sph_obj.py:347-369 that the step loop and the
    writers use.  step_counter counts sph_steps (the reference declares it
    but never advances it)."""

    def __init__(self):
        self.time_count = 0.0
        self.time_counter = 0
        self.step_counter = 0
        self.frame_div_iter = 0
        self.frame_incom_iter = 0
        self.div_iter_count = 0
        self.incom_iter_count = 0


def inlet_helper_get_dt(norm, speed, relaxing_factor, config):
    """main.py:24-39: the time for an inlet's sheet to advance one particle
    spacing, (sqrt(v^2 + 2 a r) - v) / a with a = the gravity component
    along the inlet's normal (r / v without acceleration), times the
    relaxing factor."""
    norm = np.array(norm) / np.linalg.norm(np.array(norm))
    if speed < 0:
        norm, speed = -norm, -speed
    acc_vec = np.array(config.gravity[: config.dim], dtype=np.float32)
    acc = np.dot(norm, acc_vec)
    r = config.part_size[1]
    if acc < 1e-6:
        if speed > 0:
            return r * relaxing_factor / speed
        return float("inf")
    return ((speed ** 2 + 2 * acc * r) ** 0.5 - speed) / acc * relaxing_factor


class Simulation:
    """Owns the two particle objects (fluid uid 0, boundary uid 1), the
    control block, the workspace and the reference-layout Ngrid.

    device="cpu" gives a host-only Simulation for building and inspecting
    scenes; every compute call on it raises (the library computes on the
    GPU only)."""

    def __init__(self, config, device="cuda", precision=4, fbm=False, wcsph=False, cfl=True,
                 list_cap_f=64, list_cap_b=48, fluid_max=None, bound_max=None, fast_quotient=True,
                 fluid=None, bound=None):
        self.config = config
        self.device = torch.device(device)
        self.precision = precision
        self.lib = _lib.load()
        self.fbm = fbm
        self.wcsph = wcsph
        self.cfl = cfl
        self.params = config.params(precision, list_cap_f, list_cap_b)
        if fast_quotient and precision == 4 and self.device.type == "cuda":
            if _fast_quotient_ok(self.lib, self.params, self.device):
                self.params.flags |= _lib.TSPH_FLAG_FAST_QUOTIENT
        fm = config.fluid_max_part_num if fluid_max is None else fluid_max
        bm = config.bound_max_part_num if bound_max is None else bound_max
        # objects built by the caller (the reference's Fluid(max_part_num),
        # main.py:9-10) are adopted; otherwise the Simulation builds its own
        for o, u in ((fluid, 0), (bound, 1)):
            if o is not None and (o.uid != u or o.precision != precision or o.config is not config
                                  or o.device != self.device):
                raise ValueError(f"object uid {o.uid} does not fit this Simulation (uid {u}, same Config, "
                                 "precision and device)")
        self.fluid = fluid if fluid is not None else Fluid(config, fm, 0, device, precision, list_cap_f,
                                                           list_cap_b)
        self.bound = bound if bound is not None else Fluid(config, bm, 1, device, precision, list_cap_f,
                                                           list_cap_b)
        fm, bm = self.fluid.max_part_num, self.bound.max_part_num
        self.fluid.sim = self.bound.sim = self
        cap = max(fm, 1) + max(bm, 1)
        self.ws = torch.zeros(int(self.lib.tsph_workspace_bytes(ctypes.byref(self.params), cap)),
                              dtype=torch.uint8, device=self.device)
        self.ctrl = torch.zeros(int(self.lib.tsph_control_bytes()), dtype=torch.uint8, device=self.device)
        pin = self.device.type == "cuda"
        self._ctrl_host = torch.zeros(self.ctrl.numel(), dtype=torch.uint8, pin_memory=pin)
        self._stop_host = torch.zeros(16, dtype=torch.uint8, pin_memory=pin)
        # early stop signal of the Jacobi solves (tsph_solver_begin_signal):
        # one uint64 of pinned, device-mapped, coherent host memory
        self._token = 0
        self._signal_host = self._signal_dev = None
        self._signal_np = np.zeros(1, dtype=np.uint64)
        if pin:
            h, d = ctypes.c_void_p(), ctypes.c_void_p()
            _lib.check(self.lib.tsph_signal_alloc(8, ctypes.byref(h), ctypes.byref(d)), "tsph_signal_alloc")
            self._signal_host, self._signal_dev = h, d
            self._signal_np = np.ctypeslib.as_array(ctypes.cast(h, ctypes.POINTER(ctypes.c_uint64)), shape=(1,))
        self.ngrid = Ngrid(self)
        self.bound_dirty = True
        self.pred = [max(config.iter_threshold_min, 1), max(config.iter_threshold_min, 1)]
        if config.iter_threshold_max + 1 > _lib.TSPH_MAX_ITER_LOG:
            raise Exception(f"solver_iter_threshold_max must be < {_lib.TSPH_MAX_ITER_LOG}")
        self.set_dt(config.dt)
        self.globalvar = GlobalVariable()
        self.scenario_wait_buffer = []  # objects with a positive `time`, added later
        self.inlets = []                # inlets pouring particles in (main.py:14)
        # optional per-call HIP-event timing (an EventTimer)
        self.timer = None
        self.calls = 0  # library calls issued through _call (a change counter)

    # ----------------------------------------------------------------- helpers
    def _require_gpu(self, what):
        if self.device.type != "cuda":
            raise _lib.TsphError(f"{what}: this Simulation is host-only (device={self.device}); "
                                 "the library computes on the GPU")

    def stream_ptr(self):
        self._require_gpu("compute")
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _call(self, name, fn, *args):
        self._require_gpu(name)
        self.calls += 1
        if self.timer is None or not self.timer.wants(name):
            _lib.check(fn(*args), name)
            return
        e0, e1 = self.timer.pair()
        e0.record()
        _lib.check(fn(*args), name)
        e1.record()
        self.timer.records.append((name, e0, e1))

    def ctrl_ptr(self):
        return ctypes.c_void_p(self.ctrl.data_ptr())

    def _ws(self):
        return ctypes.c_void_p(self.ws.data_ptr()), self.ws.numel()

    def control(self):
        """The control block (include/tichi_sph.h TsphControl), synchronising."""
        self._ctrl_host.copy_(self.ctrl, non_blocking=True)
        torch.cuda.current_stream(self.device).synchronize()
        return _lib.TsphControl.from_buffer_copy(self._ctrl_host.numpy().tobytes())

    @property
    def time_count(self):
        return self.globalvar.time_count

    def set_dt(self, dt):
        raw = self.control() if self.ctrl.is_cuda else _lib.TsphControl()
        raw.dt = dt
        self._write_ctrl(raw)

    def set_compression(self, c):
        raw = self.control()
        raw.compression = c
        self._write_ctrl(raw)

    def _write_ctrl(self, raw):
        buf = torch.frombuffer(bytearray(bytes(raw)), dtype=torch.uint8)
        self.ctrl.copy_(buf.to(self.device))

    def __del__(self):
        h = getattr(self, "_signal_host", None)
        if h is not None and not sys.is_finalizing():
            self.lib.tsph_signal_free(h)
            self._signal_host = None

    def _wait_signal(self, token, launched):
        """(it, stop) of the running solve once its update half-steps have
        decided on the last of the `launched` iterations, read from the early
        signal without draining the stream (a solve that enters no iteration
        signals (0, stop) from its begin); falls back to the control block
        once the stream has drained."""
        a = self._signal_np
        stream = torch.cuda.current_stream(self.device)
        spins = 0
        while True:
            v = int(a[0])
            if v >> 32 == token:
                w = v & 0xFFFFFFFF
                if w & 1 or (w >> 1) >= launched:
                    return w >> 1, w & 1
            spins += 1
            if spins % 64 == 0 and stream.query():
                it, stop, _ = self._stop_state()
                return it, stop

    def _stop_state(self):
        # it, stop are the first two int32 after two doubles
        self._stop_host.copy_(self.ctrl[16:32], non_blocking=True)
        torch.cuda.current_stream(self.device).synchronize()
        a = self._stop_host.numpy().view(np.int32)
        return int(a[0]), int(a[1]), int(a[3])  # it, stop, general_flag

    # --------------------------------------------------------------- scenario
    def get_obj_from_str(self, s):
        """main.py:16-22."""
        if s == "fluid":
            return self.fluid
        if s == "bound":
            return self.bound
        print("scenario WARNING: object does not exist (name:", s, ")")
        return None

    def add_to_scenario(self, obj_str, param):
        """main.py:42-60: `cube`, `box`, `ply` (vertices of a PLY file) and
        `inlet` (registered; refresh_inlets emits its sheets)."""
        from .ply import read_ply

        obj = self.get_obj_from_str(obj_str)
        if obj is None:
            print("scenario WARNING: obj '", obj_str, "' non-exist.")
            return
        kind = param["type"]
        if kind == "cube":
            obj.scene_add_cube(param["start_pos"], param["end_pos"], param["volume_frac"], param["vel"],
                               int(param["color"], 16), param["particle_relaxing_factor"])
        elif kind == "box":
            obj.scene_add_box(param["start_pos"], param["end_pos"], param["layers"], param["volume_frac"],
                              param["vel"], int(param["color"], 16), param["particle_relaxing_factor"])
        elif kind == "ply":
            verts = read_ply(param["file_name"])
            obj.push_part_from_ply(len(verts), verts, param["volume_frac"], param["vel"], int(param["color"], 16))
        elif kind == "inlet":
            if np.linalg.norm(np.array(param["norm"])) < 1e-6:
                raise Exception("inlet ERROR: magnitute of attribute 'norm' should not be zero.")
            param["obj"] = obj_str
            param["dt"] = inlet_helper_get_dt(param["norm"], param["speed"], param["particle_relaxing_factor"],
                                              self.config)
            param["t_pre"] = self.globalvar.time_count - param["dt"]
            self.inlets.append(param)
        else:
            raise Exception("scenario ERROR: object type '%s' is not supported" % kind)
        if obj is self.bound and kind != "inlet":
            self.bound_dirty = True

    def init_scenario(self):
        """main.py:62-79: objects without a positive `time` are added now,
        the others wait in scenario_wait_buffer."""
        sb = self.config.scenario_buffer
        try:
            for part in sb:
                if part == "sim_env":
                    continue
                for param in sb[part]["objs"]:
                    if "time" not in param or param["time"] <= 0.0:
                        self.add_to_scenario(part, param)
                    else:
                        param["_part"] = part
                        self.scenario_wait_buffer.append(param)
        except Exception as e:
            raise Exception("scenario ERROR: no scenario file or scenario file invalid.") from e

    def refresh_inlets(self):
        """main.py:82-101: every inlet whose interval has passed emits one
        sheet; inlets past their end_time are removed."""
        for i in range(len(self.inlets) - 1, -1, -1):
            inlet = self.inlets[i]
            t = self.globalvar.time_count
            if "end_time" in inlet and t >= inlet["end_time"]:
                self.inlets.pop(i)
            elif t >= inlet["t_pre"] + inlet["dt"]:
                try:
                    inlet["t_pre"] = t
                    self.get_obj_from_str(inlet["obj"]).scene_add_from_inlet(
                        inlet["center"], inlet["size"], inlet["norm"], inlet["speed"], inlet["volume_frac"],
                        int(inlet["color"], 16), inlet["particle_relaxing_factor"])
                    if inlet["obj"] == "bound":
                        self.bound_dirty = True
                except Exception:
                    warnings.warn("inlet WARNING: refresh_inlets failed for:" + json.dumps(inlet), RuntimeWarning)
                    self.inlets.pop(i)

    def refresh_scenario(self):
        """main.py:103-112: add waiting objects whose time has come, then
        refresh the inlets."""
        for i in range(len(self.scenario_wait_buffer) - 1, -1, -1):
            param = self.scenario_wait_buffer[i]
            try:
                if param["time"] <= self.globalvar.time_count:
                    self.add_to_scenario(param["_part"], param)
                    self.scenario_wait_buffer.pop(i)
            except Exception:
                warnings.warn("scenario WARNING: invalid scenario object:" + json.dumps(param), RuntimeWarning)
                self.scenario_wait_buffer.pop(i)
        self.refresh_inlets()

    def run_step(self):
        """One GUI frame of main.py:339-356: steps (refresh_scenario,
        cfl_condition, sph_step) until the simulated time reaches the next
        frame at gui_fps.  Returns the number of steps taken."""
        g = self.globalvar
        g.time_counter += 1
        steps = 0
        while g.time_count < g.time_counter / self.config.gui_fps:
            self.refresh_scenario()
            st = self.sph_step(stats=True)
            g.time_count += st.dt
            g.step_counter += 1
            g.div_iter_count, g.incom_iter_count = st.div_iter, st.incom_iter
            g.frame_div_iter += st.div_iter
            g.frame_incom_iter += st.incom_iter
            self.last_stats = st
            steps += 1
        return steps

    # ------------------------------------------------------------ fused step
    def bound_needs_sort(self):
        """The boundary's cell order no longer matches its storage: the
        scene added boundary particles (add_to_scenario, an inlet), or the
        caller wrote through Fluid.scene_add_* / push_part_seq /
        from_numpy("pos").  The reference rebuilds its Ngrid every step
        (main.py:178-183), so a changed boundary is always searched; here the
        static boundary is re-sorted before the next step instead."""
        return self.bound_dirty or not self.bound.cell_table_ok

    def _sort_bound(self):
        P = ctypes.byref(self.params)
        ws, wsb = self._ws()
        rc = self.lib.tsph_sort_object(P, ctypes.byref(self.bound.struct()), self.ctrl_ptr(), 0, ws, wsb,
                                       self.stream_ptr())
        _lib.check(rc, "tsph_sort_object(bound)")
        self.bound.adopt_swaps()
        self.bound_dirty = False

    def forced_params(self, iters):
        """A parameter block whose stop rule runs exactly `iters` Jacobi
        iterations (iter_min = iter_max = iters, thresholds +inf)."""
        p = _lib.TsphParams.from_buffer_copy(bytes(self.params))
        p.iter_min = iters
        p.iter_max = iters
        p.divergence_threshold = float("inf")
        p.compression_threshold = float("inf")
        return p

    def _solve(self, solve, tape=None, force=None, frozen=None):
        """Run solve 0 (divergence) or 1 (density) to its stop rule.  With
        `tape` (a list), each launched iteration k records vel_adv entering it
        and its kappa into new tensors appended as (vel_adv, kappa); entries
        past the iterations that ran are dropped.  `force` = n runs exactly n
        iterations.  `frozen` (a list of recorded kappa tensors, one per
        iteration) clamps psi where the recorded kappa is 0 instead of where
        psi is negative (tsph_solver_psi_frozen)."""
        params = self.params if force is None else self.forced_params(force)
        P = ctypes.byref(params)
        ws, wsb = self._ws()
        st = self.stream_ptr()
        f, b = ctypes.byref(self.fluid.struct()), ctypes.byref(self.bound.struct())
        self._token = self._token % 0x7FFFFFFF + 1
        token = self._token
        _lib.check(self.lib.tsph_solver_begin_signal(P, self.ctrl_ptr(), solve,
                                                     self._signal_dev, token, st),
                   "tsph_solver_begin_signal")
        kmax = (self.config.iter_threshold_max if force is None else force) + 1
        if force == 0:
            kmax = 0
        k = 0
        it = 0
        chunk = max(self.pred[solve] if force is None else force, 1)
        cap = self.fluid.t["pos"].shape[0]
        while k < kmax:
            for _ in range(min(chunk, kmax - k)):
                if frozen is not None:
                    self._call("solver_psi", self.lib.tsph_solver_psi_frozen, P, f, b, self.ctrl_ptr(), k,
                               ctypes.c_void_p(frozen[k].data_ptr()), ws, wsb, st)
                elif tape is None:
                    self._call("solver_psi", self.lib.tsph_solver_psi, P, f, b, self.ctrl_ptr(), k, ws, wsb, st)
                else:
                    tv = torch.empty(cap, 4, dtype=self.fluid.dtype, device=self.device)
                    tk = torch.empty(cap, dtype=self.fluid.dtype, device=self.device)
                    tape.append((tv, tk))
                    self._call("solver_psi", self.lib.tsph_solver_psi_taped, P, f, b, self.ctrl_ptr(), k,
                               ctypes.c_void_p(tv.data_ptr()), ctypes.c_void_p(tk.data_ptr()), ws, wsb, st)
                self._call("solver_update", self.lib.tsph_solver_update, P, f, b, self.ctrl_ptr(), k, ws, wsb, st)
                k += 1
            it, stop = self._wait_signal(token, k)
            if stop:
                break
            chunk = 1
        if force is None:
            self.pred[solve] = max(it, 1)
        if tape is not None:
            del tape[it:]
        return it

    def sph_step(self, stats=False):
        """One step of main.py:175-256 (preceded by cfl_condition, main.py:348).
        With no fluid particles it still sorts and updates the boundary, and
        each solve runs its minimum iteration count (the reference's empty
        mean is NaN)."""
        P = ctypes.byref(self.params)
        ws, wsb = self._ws()
        st = self.stream_ptr()
        if self.bound_needs_sort():
            self._sort_bound()
        self._call("sort", self.lib.tsph_sort_object, P, ctypes.byref(self.fluid.struct()), self.ctrl_ptr(),
                   1 if self.cfl else 0, ws, wsb, st)
        self.fluid.adopt_swaps()
        f, b = ctypes.byref(self.fluid.struct()), ctypes.byref(self.bound.struct())
        self._call("density_alpha", self.lib.tsph_density_alpha, P, f, b, self.ctrl_ptr(), ws, wsb, st)
        div_it = self._solve(0)
        self._call("advect", self.lib.tsph_advect, P, f, b, self.ctrl_ptr(), ws, wsb, st)
        inc_it = self._solve(1)
        if self.wcsph:  # main.py:236-240
            self._call("wcsph", self.lib.tsph_wcsph, P, f, b, self.ctrl_ptr(), st)
        if self.fbm:
            rounds = 0
            while rounds < 1000:
                self._call("fbm_round", self.lib.tsph_fbm_round, P, f, self.ctrl_ptr(), ws, wsb, st)
                rounds += 1
                _, _, gflag = self._stop_state()
                if not gflag:
                    break
        self._call("tail", self.lib.tsph_tail, P, f, self.ctrl_ptr(), 1 if self.fbm else 0, ws, wsb, st)
        if stats:
            return StepStats(self.control(), self.precision)
        return div_it, inc_it

    # -------------------------------------------------- reference-order step
    def sph_step_reference(self, wcsph=None, fbm=None):
        """main.py:175-256 kernel by kernel through the sph.py surface, in the
        reference's call forms f(ngrid, obj, nobj) and f(obj)
        (cfl_condition first, as run_step does at main.py:348)."""
        from . import sph as K

        fbm = self.fbm if fbm is None else fbm
        wcsph = self.wcsph if wcsph is None else wcsph
        fluid, bound, ngrid = self.fluid, self.bound, self.ngrid
        fluid.cell_table_ok = False  # this path walks the Ngrid, not the cell order
        cfg = self.config
        if self.cfl:
            K.cfl_condition(fluid)
        ngrid.clear_node()
        ngrid.encode(fluid)
        ngrid.encode(bound)
        ngrid.mem_shift()
        ngrid.fill_node(fluid)
        ngrid.fill_node(bound)
        K.SPH_clean_value(fluid)
        K.SPH_clean_value(bound)
        K.SPH_prepare_attr(ngrid, fluid, fluid)
        K.SPH_prepare_attr(ngrid, fluid, bound)
        K.SPH_prepare_attr(ngrid, bound, bound)
        K.SPH_prepare_attr(ngrid, bound, fluid)
        K.SPH_prepare_alpha_1(ngrid, fluid, fluid)
        K.SPH_prepare_alpha_1(ngrid, fluid, bound)
        K.SPH_prepare_alpha_2(ngrid, fluid, fluid)
        K.SPH_prepare_alpha_2(ngrid, bound, fluid)
        K.SPH_prepare_alpha(fluid)
        K.SPH_prepare_alpha(bound)
        out = {}
        for solve, thr in ((0, cfg.divergence_threshold), (1, cfg.compression_threshold)):
            it = 0
            errs = []
            if solve == 0:
                K.SPH_vel_2_vel_adv(fluid)
            while it < cfg.iter_threshold_min or self.compression() > thr:
                K.IPPE_adv_psi_init(fluid)
                K.IPPE_adv_psi(ngrid, fluid, fluid)
                K.IPPE_adv_psi(ngrid, fluid, bound)
                K.IPPE_psi_adv_non_negative(fluid)
                errs.append(self.compression())
                K.IPPE_update_vel_adv(ngrid, fluid, fluid)
                K.IPPE_update_vel_adv(ngrid, fluid, bound)
                it += 1
                if it > cfg.iter_threshold_max:
                    break
            out[solve] = (it, errs)
            if solve == 0:
                K.SPH_vel_adv_2_vel(fluid)
                K.SPH_advection_gravity_acc(fluid)
                K.SPH_advection_viscosity_acc(ngrid, fluid, fluid)
                K.SPH_advection_surface_tension_acc(ngrid, fluid, fluid)
                K.SPH_advection_update_vel_adv(fluid)
        if wcsph:
            K.WC_pressure_val(fluid)
            K.WC_pressure_acce(ngrid, fluid, fluid)
            K.WC_pressure_acce(ngrid, fluid, bound)
            K.SPH_advection_update_vel_adv(fluid)
        rounds = 0
        if fbm:
            while self.control().general_flag > 0 and rounds < 1000:
                K.SPH_FBM_clean_tmp(fluid)
                K.SPH_FBM_convect(ngrid, fluid, fluid)
                K.SPH_FBM_diffuse(ngrid, fluid, fluid)
                K.SPH_FBM_check_tmp(fluid)
                rounds += 1
            K.SPH_update_volume_frac(fluid)
        K.SPH_update_mass(fluid)
        K.SPH_update_pos(fluid)
        K.SPH_update_energy(fluid)
        return out[0][0], out[1][0], out[0][1], out[1][1], rounds

    def compression(self):
        c = self.control().compression
        return float(np.float32(c)) if self.precision == 4 else float(c)


def _simulation_of_binding(b):
    return Simulation(b.config, fluid=b.objs[0], bound=b.objs[1], **b.kw)


binding._FACTORY = _simulation_of_binding


def use_config(config, **kw):
    """Bind the reference's module-level construction to `config` (its
    global config, sph_config.py): after this, `Ngrid()`,
    `Fluid(max_part_num=...)` (fluid first, then the boundary, main.py:8-10)
    and `NeighborSearch(number_of_uids)` build the objects the sph.py kernels
    take, and `current_simulation()` gives the Simulation that owns them
    (scenario ingestion, the fused step, the writers).  `kw` are the
    Simulation's options (device, precision, fbm, cfl, ...)."""
    return binding.use_config(config, **kw)


def current_simulation():
    """The Simulation of the objects built against the bound Config."""
    return binding.simulation()


def load_simulation(config_file, scenario_file, **kw):
    sim = Simulation(Config.from_files(config_file, scenario_file), **kw)
    sim.init_scenario()
    return sim


def simulation_from_dicts(config_buffer, scenario_buffer, **kw):
    sim = Simulation(Config(json.loads(json.dumps(config_buffer)), json.loads(json.dumps(scenario_buffer))), **kw)
    sim.init_scenario()
    return sim
