"""Differentiable DFSPH rollouts: hand-written HIP adjoint kernels behind
`torch.autograd.Function`.

The reference has no autodiff (no `needs_grad`, `ti.Tape` or `ti.ad` in
sph.py / main.py).  This module differentiates the fused step of
`Simulation.sph_step()` (main.py:175-256) for one-phase DFSPH with a fixed
dt and no FBM.  The derivative is that of the recorded forward step: the
Jacobi iteration counts, the psi clamps (sph.py:219-226), the alpha clamp
(sph.py:103-105) and the neighbour sets are held at their recorded values;
between such events the step is smooth (the cubic spline is C^1 with
W'(h) = W''(h) = 0) and the derivative is exact.

Memory.  The forward pass keeps checkpoints: the persistent state (pos,
vel, mass, rest density, rest volume, volume fractions, pid and the control
block) every `every` steps, and the tapes of the last segment, which the
backward pass reverses first.  With a tape budget (`tape_budget` bytes) it
also keeps the tapes of the most recent earlier segments whose tapes fit in
the budget, dropping the oldest taped segment whenever the total exceeds it.
The backward pass replays every segment whose tapes were not kept from its
checkpoint while recording a tape per step, then reverses the segment step
by step.  A replayed step records exactly the tape of the forward step (the
same kernels on the same state, with the forward's iteration counts), so
the gradients do not depend on the budget; the budget only trades HBM for
the replay work.  A tape holds the sorted positions, masses, rest
densities, pids and cell table, the densities, alpha factors, volumes and
alpha_1 of pass A, vel_adv and kappa of every Jacobi iteration, and the
advection's velocity and normals, and (`DiffRunner.tape_lists`, on by
default) the fluid neighbour lists: a taped step's pass A writes them into
a list buffer of the tape's own (list_cap_f rows of the capacity, 4 bytes
each, with both lists' counts), so they cost no copy; without stored lists
`tsph_relist` recomputes them bit-identically from the recorded positions.
The boundary enters the reverse only through per-particle sums over the
boundary neighbours (`tsph_adj_bound_sums`, recomputed for each reversed
step), so the boundary lists are neither stored nor rebuilt.  The reverse
reads a tape in place, through a copy of the fluid's descriptor.

    runner = DiffRunner(sim)
    x, v = rollout(runner, v0, mu, gamma, steps=200, every=20)
    loss = ((x - target) ** 2).sum(1).mean()
    loss.backward()          # v0.grad, mu.grad, gamma.grad
"""
import ctypes

import torch

from . import _lib


class StepTape:
    """Everything the reverse of one step reads that is not recomputed."""

    __slots__ = ("pos", "mass", "rest_density", "pid", "cell_start", "density", "alpha", "vol",
                 "alpha_1", "vel", "normal", "iters", "calls", "nbr_f", "nbr_nf", "nbr_nb")

    def __init__(self):
        self.iters = ([], [])  # per solve: [(vel_adv entering it, kappa), ...]
        # the fluid lists (list_cap_f slot-major rows of the object's
        # capacity), their counts and the boundary lists' counts, when the
        # runner stores them
        self.nbr_f = self.nbr_nf = self.nbr_nb = None

    def nbytes(self):
        ts = [self.pos, self.mass, self.rest_density, self.pid, self.cell_start, self.density, self.alpha,
              self.vol, self.alpha_1, self.vel, self.normal]
        ts += [t for s in self.iters for pair in s for t in pair]
        ts += [t for t in (self.nbr_f, self.nbr_nf, self.nbr_nb) if t is not None]
        return sum(t.numel() * t.element_size() for t in ts)


class DiffRunner:
    """Drives a `Simulation` forward with optional tapes and backward through
    the adjoint entry points of the C ABI."""

    def __init__(self, sim):
        cfg = sim.config
        if cfg.solver_type != "DFSPH" or cfg.phase_num != 1:
            raise ValueError("the adjoint supports one-phase DFSPH scenes only")
        if sim.fbm:
            raise ValueError("the adjoint does not differentiate FBM rounds")
        if sim.cfl:
            raise ValueError("the adjoint needs a fixed dt: build the Simulation with cfl=False")
        if cfg.phase_dynamic_viscosity is not None:
            raise ValueError("the adjoint differentiates the global viscosity only (no per-phase viscosity)")
        self.sim = sim
        self.lib = sim.lib
        f = sim.fluid
        cap = f.t["pos"].shape[0]
        dt_, dev = f.dtype, sim.device
        self.buf = {name: torch.zeros(cap, 4, dtype=dt_, device=dev)
                    for name in ("xbar", "vbar", "vbar_out", "nbar", "a1bar")}
        for name in ("rhobar", "alphabar"):
            self.buf[name] = torch.zeros(cap, dtype=dt_, device=dev)
        self.buf["bmom"] = torch.zeros(3 * cap, 4, dtype=dt_, device=dev)  # boundary sums
        self.buf["bflag"] = torch.zeros(cap, dtype=torch.int32, device=dev)
        self.grad = torch.zeros(2, dtype=torch.float64, device=dev)
        self.adj = _lib.TsphAdjoint()
        for name, t in self.buf.items():
            setattr(self.adj, name, t.data_ptr())
        self.adj.grad = self.grad.data_ptr()
        # False skips the surface-tension reverse; allowed only while gamma == 0
        self.surface_tension_grad = True
        # True stores each taped step's fluid lists (HBM for the relist's work)
        self.tape_lists = True
        self._own_lists = (f.t["nbr_f"], f.t["nbr_nf"], f.t["nbr_nb"])  # the Simulation's own
        self.force = None     # optional list of (div_iter, incom_iter) per step to replay
        self.step_index = 0
        self.iter_log = []    # (div_iter, incom_iter) of every forward step since the last reset

    @property
    def surface_tension_grad(self):
        return self._st_grad

    @surface_tension_grad.setter
    def surface_tension_grad(self, on):
        """With the surface-tension reverse on, the forward advection records
        what that reverse gathers also at gamma = 0 (TSPH_FLAG_ST_RECORDS)."""
        self._st_grad = bool(on)
        p = self.sim.params
        if self._st_grad:
            p.flags |= _lib.TSPH_FLAG_ST_RECORDS
        else:
            p.flags &= ~_lib.TSPH_FLAG_ST_RECORDS

    # -------------------------------------------------------------- params
    def set_params(self, mu=None, gamma=None):
        if mu is not None:
            self.sim.params.dynamic_viscosity = float(mu)
        if gamma is not None:
            self.sim.params.surface_tension_gamma = float(gamma)

    # --------------------------------------------------------------- state
    def checkpoint(self):
        f = self.sim.fluid
        n = f.part_num
        st = {name: f.t[name][:n].clone() for name in
              ("pos", "vel", "mass", "rest_density", "rest_volume", "volume_frac", "pid")}
        st["ctrl"] = self.sim.ctrl.clone()
        st["step_index"] = self.step_index
        return st

    def restore(self, st):
        f = self.sim.fluid
        n = f.part_num
        for name in ("pos", "vel", "mass", "rest_density", "rest_volume", "volume_frac", "pid"):
            f.t[name][:n].copy_(st[name])
        self.sim.ctrl.copy_(st["ctrl"])
        self.step_index = st["step_index"]

    def set_velocity_ref(self, v):
        """Fluid velocities from an [n, dim] tensor in reference (pid) order."""
        f = self.sim.fluid
        n, d = f.part_num, self.sim.config.dim
        pid = f.t["pid"][:n].long()
        f.t["vel"][:n, :d].copy_(v.to(f.dtype)[pid])

    def positions_ref(self):
        f = self.sim.fluid
        n, d = f.part_num, self.sim.config.dim
        out = torch.empty(n, d, dtype=f.dtype, device=self.sim.device)
        out[f.t["pid"][:n].long()] = f.t["pos"][:n, :d]
        return out

    def velocities_ref(self):
        f = self.sim.fluid
        n, d = f.part_num, self.sim.config.dim
        out = torch.empty(n, d, dtype=f.dtype, device=self.sim.device)
        out[f.t["pid"][:n].long()] = f.t["vel"][:n, :d]
        return out

    # ------------------------------------------------------------- forward
    def step(self, tape=False, frozen=None):
        """One fused step (Simulation.sph_step without CFL / FBM); with
        tape=True returns the StepTape of the step.  With `frozen` (a StepTape
        of this step, recorded earlier) the Jacobi solves clamp psi where the
        recorded iterations clamped it (tsph_solver_psi_frozen): the step then
        follows the recorded path, whose derivative the adjoint takes, also
        under a perturbed input (finite differences along that path)."""
        sim = self.sim
        f, b = sim.fluid, sim.bound
        P = ctypes.byref(sim.params)
        ws, wsb = sim._ws()
        stp = sim.stream_ptr()
        force = self.force[self.step_index] if self.force is not None else (None, None)
        if sim.bound_needs_sort():
            sim._sort_bound()
        # pass A writes the fluid lists into the buffers this step's tape
        # keeps, or into the Simulation's own (never into a kept tape's)
        if tape and self.tape_lists:
            lists = tuple(torch.empty_like(t) for t in self._own_lists)
        else:
            lists = self._own_lists
        self._use_lists(*lists)
        sim._call("sort", sim.lib.tsph_sort_object, P, ctypes.byref(f.struct()), sim.ctrl_ptr(), 0, ws, wsb, stp)
        f.adopt_swaps()
        tp = None
        n = f.part_num
        if tape:
            tp = StepTape()
            if self.tape_lists:
                tp.nbr_f, tp.nbr_nf, tp.nbr_nb = lists
            tp.pos = f.t["pos"][:n].clone()
            tp.mass = f.t["mass"][:n].clone()
            tp.rest_density = f.t["rest_density"][:n].clone()
            tp.pid = f.t["pid"][:n].clone()
            tp.cell_start = f.t["cell_start"].clone()
        fs, bs = ctypes.byref(f.struct()), ctypes.byref(b.struct())
        sim._call("density_alpha", sim.lib.tsph_density_alpha, P, fs, bs, sim.ctrl_ptr(), ws, wsb, stp)
        if tape:
            tp.density = f.t["sph_density"][:n].clone()
            tp.alpha = f.t["alpha"][:n].clone()
            tp.vol = f.t["vol"][:n].clone()
            tp.alpha_1 = f.t["alpha_1"][:n].clone()
        fz = [None, None] if frozen is None else [[kap for _, kap in frozen.iters[q]] for q in (0, 1)]
        if frozen is not None and force[0] is None:
            force = (len(fz[0]), len(fz[1]))
        d_it = sim._solve(0, tape=tp.iters[0] if tape else None, force=force[0], frozen=fz[0])
        sim._call("advect", sim.lib.tsph_advect, P, fs, bs, sim.ctrl_ptr(), ws, wsb, stp)
        if tape:
            tp.vel = f.t["vel"][:n].clone()
            tp.normal = f.t["normal"][:n].clone()
        i_it = sim._solve(1, tape=tp.iters[1] if tape else None, force=force[1], frozen=fz[1])
        sim._call("tail", sim.lib.tsph_tail, P, fs, sim.ctrl_ptr(), 0, ws, wsb, stp)
        if tape:
            # while no later call runs, the Simulation's lists are this step's
            # (what a tape without stored lists relies on)
            tp.calls = sim.calls
        # the Simulation keeps its own list buffers between steps, so that no
        # later step writes into a tape's
        self._use_lists(*self._own_lists)
        self.iter_log.append((d_it, i_it))
        self.step_index += 1
        return tp

    # ------------------------------------------------------------ backward
    def reverse(self, tp, xbar_ref, vbar_ref):
        """Reverse of one recorded step.  xbar_ref / vbar_ref ([n, 4], pid
        order) hold d loss / d (pos, vel) after the step on entry and before
        it on return.  adj.grad accumulates d loss / d (mu, gamma).  The tape
        is consumed (its vel_adv records carry psi_bar afterwards)."""
        sim = self.sim
        f, b = sim.fluid, sim.bound
        P = ctypes.byref(sim.params)
        ws, wsb = sim._ws()
        stp = sim.stream_ptr()
        # the reverse kernels read the recorded fields (and lists) through a
        # copy of the fluid's descriptor that points at the tape: nothing is
        # copied back, and no reverse kernel writes these fields
        fsd = _lib.TsphObject.from_buffer_copy(f.struct())
        for name, t in (("pos", tp.pos), ("mass", tp.mass), ("rest_density", tp.rest_density),
                        ("cell_start", tp.cell_start), ("sph_density", tp.density),
                        ("alpha", tp.alpha), ("vol", tp.vol), ("alpha_1", tp.alpha_1),
                        ("nbr_f", tp.nbr_f), ("nbr_nf", tp.nbr_nf), ("nbr_nb", tp.nbr_nb)):
            if t is not None:
                setattr(fsd, name, t.data_ptr())
        fs, bs = ctypes.byref(fsd), ctypes.byref(b.struct())
        A = ctypes.byref(self.adj)
        vp = ctypes.c_void_p
        def last_kappa(pairs):
            return pairs[-1][1].data_ptr() if pairs else None

        if tp.calls != sim.calls and tp.nbr_f is None:
            sim._call("relist", self.lib.tsph_relist, P, fs, bs, sim.ctrl_ptr(), ws, wsb, stp)
        sim._call("adj_bound_sums", self.lib.tsph_adj_bound_sums, P, fs, bs, A, stp)
        self.adj.kappa_next = last_kappa(tp.iters[1])
        sim._call("adj_begin", self.lib.tsph_adj_begin, P, fs, sim.ctrl_ptr(), A, vp(tp.pid.data_ptr()),
                  vp(xbar_ref.data_ptr()), vp(vbar_ref.data_ptr()), stp)

        def iterations(pairs):
            # each call reads its iteration's kappa from vbar.w and stores the
            # next (earlier) iteration's kappa into vbar_out.w
            for k in range(len(pairs) - 1, -1, -1):
                self.adj.tape_vel_adv = pairs[k][0].data_ptr()
                self.adj.kappa_next = pairs[k - 1][1].data_ptr() if k > 0 else None
                sim._call("adj_solver_iteration", self.lib.tsph_adj_solver_iteration, P, fs, bs,
                          sim.ctrl_ptr(), A, ws, wsb, stp)
                self._swap_v()

        iterations(tp.iters[1])
        self.adj.kappa_next = last_kappa(tp.iters[0])
        self.adj.tape_vel = tp.vel.data_ptr()
        self.adj.tape_normal = tp.normal.data_ptr()
        skip_st = not self.surface_tension_grad and sim.params.surface_tension_gamma == 0.0
        self.adj.nbar = None if skip_st else self.buf["nbar"].data_ptr()
        sim._call("adj_advect", self.lib.tsph_adj_advect, P, fs, sim.ctrl_ptr(), A, ws, wsb, stp)
        self._swap_v()
        iterations(tp.iters[0])
        sim._call("adj_density_alpha", self.lib.tsph_adj_density_alpha, P, fs, bs, A, stp)
        sim._call("adj_end", self.lib.tsph_adj_end, P, fs, A, vp(tp.pid.data_ptr()),
                  vp(xbar_ref.data_ptr()), vp(vbar_ref.data_ptr()), stp)

    def _use_lists(self, nbr_f, nbr_nf, nbr_nb):
        """Points the fluid's list buffers (tensors and descriptor) at these:
        the fluid lists, their counts and the boundary lists' counts."""
        f = self.sim.fluid
        if f.t["nbr_f"] is nbr_f and f.t["nbr_nf"] is nbr_nf and f.t["nbr_nb"] is nbr_nb:
            return
        f.t["nbr_f"], f.t["nbr_nf"], f.t["nbr_nb"] = nbr_f, nbr_nf, nbr_nb
        f._refresh_struct()
        f.layout_epoch += 1  # a SceneBatch holding the old descriptor rebuilds it

    def _swap_v(self):
        self.buf["vbar"], self.buf["vbar_out"] = self.buf["vbar_out"], self.buf["vbar"]
        self.adj.vbar = self.buf["vbar"].data_ptr()
        self.adj.vbar_out = self.buf["vbar_out"].data_ptr()


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v0, mu, gamma, runner, steps, every, budget):
        runner.set_params(mu.item(), gamma.item())
        runner.surface_tension_grad = bool(gamma.requires_grad) or gamma.item() != 0.0
        runner.set_velocity_ref(v0.detach())
        runner.step_index = 0
        runner.iter_log = []
        ckpts = []
        nseg = (steps + every - 1) // every
        kept = {}  # segment -> its tapes, for the last segment and those within the budget
        held = []  # (segment, bytes) of the kept earlier segments, oldest first
        used = 0
        est = None  # bytes of one step's tape, from the first tape recorded
        for t in range(steps):
            s = t // every
            if t % every == 0:
                ckpts.append(runner.checkpoint())
                # the last segment is reversed first: record its tapes now
                # instead of replaying it in the backward pass; an earlier
                # one only if a whole segment of tapes can fit in the budget
                seg_len = min(every, steps - t)
                if s == nseg - 1 or (budget > 0 and (est is None or est * seg_len <= budget)):
                    kept[s] = []
            tp = runner.step(tape=s in kept)
            if tp is None:
                continue
            if est is None:
                est = tp.nbytes()
            kept[s].append(tp)
            if s == nseg - 1:
                continue
            if not held or held[-1][0] != s:
                held.append([s, 0])
            b = tp.nbytes()
            held[-1][1] += b
            used += b
            while held and used > budget:
                s_old, b_old = held.pop(0)
                used -= b_old
                del kept[s_old]
        ctx.runner, ctx.ckpts, ctx.steps, ctx.every = runner, ckpts, steps, every
        ctx.kept = kept
        ctx.mu, ctx.gamma = float(mu.item()), float(gamma.item())
        ctx.v0_dtype = v0.dtype
        return runner.positions_ref(), runner.velocities_ref()

    @staticmethod
    def backward(ctx, gx, gv):
        runner = ctx.runner
        sim = runner.sim
        f = sim.fluid
        n, d = f.part_num, sim.config.dim
        final = runner.checkpoint()
        runner.set_params(ctx.mu, ctx.gamma)
        xbar = torch.zeros(n, 4, dtype=f.dtype, device=sim.device)
        vbar = torch.zeros(n, 4, dtype=f.dtype, device=sim.device)
        if gx is not None:
            xbar[:, :d] = gx.to(f.dtype)
        if gv is not None:
            vbar[:, :d] = gv.to(f.dtype)
        runner.grad.zero_()
        saved_force, saved_log = runner.force, runner.iter_log
        runner.force = list(saved_log)  # replay the forward's iteration counts exactly
        runner.iter_log = []
        for s in reversed(range(len(ctx.ckpts))):
            t0 = s * ctx.every
            t1 = min(t0 + ctx.every, ctx.steps)
            if s in ctx.kept:
                tapes = ctx.kept.pop(s)
            else:
                runner.restore(ctx.ckpts[s])
                tapes = [runner.step(tape=True) for _ in range(t0, t1)]
            for tp in reversed(tapes):
                runner.reverse(tp, xbar, vbar)
            del tapes
        runner.force, runner.iter_log = saved_force, saved_log
        runner.restore(final)
        g = runner.grad.to(torch.float64)
        gmu = g[0].to(ctx.v0_dtype)
        ggam = g[1].to(ctx.v0_dtype)
        return vbar[:, :d].to(ctx.v0_dtype), gmu, ggam, None, None, None, None


def rollout(runner, v0, mu, gamma, steps, every=20, tape_budget=0):
    """Differentiable rollout of `steps` fused DFSPH steps from the runner's
    current state with initial fluid velocity v0 ([n, dim], reference
    order), dynamic viscosity mu and surface-tension coefficient gamma (0-d
    tensors).  Returns final (positions, velocities) in reference order.
    The state is checkpointed every `every` steps; the tapes of the last
    segment, and of the latest earlier segments whose tapes fit in
    `tape_budget` bytes of HBM, are kept from the forward pass instead of
    being replayed in the backward pass."""
    if not torch.is_tensor(mu):
        mu = torch.tensor(float(mu), dtype=v0.dtype, device=v0.device)
    if not torch.is_tensor(gamma):
        gamma = torch.tensor(float(gamma), dtype=v0.dtype, device=v0.device)
    return _Rollout.apply(v0, mu, gamma, runner, int(steps), max(int(every), 1), max(int(tape_budget), 0))
