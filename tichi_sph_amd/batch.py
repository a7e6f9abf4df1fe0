"""Batched scenes: many independent Simulations advanced by grouped launches.

The reference runs one scene per process.  `SceneBatch([sim, ...])`
advances every scene by one `sph_step` (main.py:175-256, fbm off) with ONE
grouped launch per kernel over all scenes: a device-resident scene table
(include/tichi_sph.h, tsph_batch_*) gives each grouped block its scene, and
the same per-block kernel bodies as the single-scene path run on it, so a
batched step is bit-identical to each scene's own `sph_step()`.

Each scene keeps its own stop rule (its own control block).  The host
launches a predicted number of Jacobi iterations for all scenes, then reads
every scene's stop state with one copy and continues while any scene's solve
is still running; scenes that have stopped are predicated off on the device.

`capture()` records the fixed parts of a step in three hipGraphs (sort +
pass A + the predicted divergence iterations; advection + the predicted
density iterations; the tail), replayed by `step()`; extra iterations beyond
the prediction are launched eagerly.  `autotune()` picks, per kernel group,
the number of blocks of the grouped launch (a grid-stride loop over the
tiles) by timing candidates on the live batch and restoring its state.
"""
import ctypes
import sys

import numpy as np
import torch

from . import _lib
from .solver import EventTimer

_SWAPPED = ("pos", "vel", "mass", "rest_density", "rest_volume", "volume_frac", "pid")
GROUPS = ("sort", "density_alpha", "solver_psi", "solver_update", "advect", "tail")


class SceneBatch:
    def __init__(self, sims, cfl=False):
        if not sims:
            raise ValueError("SceneBatch needs at least one Simulation")
        prec = sims[0].precision
        for s in sims:
            if s.precision != prec:
                raise ValueError("all scenes of a batch must share one precision")
            if s.fbm:
                raise ValueError("FBM rounds are not batched")
            if s.wcsph:
                raise ValueError("the WCSPH block (main.py:236-240) is not batched")
            if s.fluid.part_num == 0:
                # the grouped stop rule runs in the scene's first fluid tile
                raise ValueError("every scene of a batch needs fluid particles")
            if s.device != sims[0].device:
                raise ValueError("all scenes of a batch must live on one device")
        for s in sims:
            if s.bound_needs_sort():
                s._sort_bound()
        self.sims = list(sims)
        self.lib = sims[0].lib
        self.device = sims[0].device
        self.cfl = 1 if cfl else 0
        ns = len(sims)
        self.ns = ns
        self._ctrl = (ctypes.c_void_p * ns)(*[s.ctrl.data_ptr() for s in sims])
        self._ws = (ctypes.c_void_p * ns)(*[s.ws.data_ptr() for s in sims])
        self._wsb = (ctypes.c_size_t * ns)(*[s.ws.numel() for s in sims])
        nbytes = int(self.lib.tsph_batch_table_bytes(prec, ns))
        self.table = torch.zeros(nbytes, dtype=torch.uint8, device=self.device)
        self._staging = torch.zeros(nbytes, dtype=torch.uint8).pin_memory()
        self.info = _lib.TsphBatchInfo()
        self._prepare()
        self._stop_dev = torch.zeros(2 * ns, dtype=torch.int32, device=self.device)
        self._stop_host = torch.zeros(2 * ns, dtype=torch.int32).pin_memory()
        # early stop signals of the solves: one uint64 per scene in pinned,
        # device-mapped host memory (tsph_batch_solver_begin_signal)
        self._token = 0
        self._sig_host = None
        if self.device.type == "cuda":
            h, d = ctypes.c_void_p(), ctypes.c_void_p()
            _lib.check(self.lib.tsph_signal_alloc(8 * ns, ctypes.byref(h), ctypes.byref(d)), "tsph_signal_alloc")
            self._sig_host, self._sig_dev = h, d
            self._sig_np = np.ctypeslib.as_array(ctypes.cast(h, ctypes.POINTER(ctypes.c_uint64)), shape=(ns,))
        self.kmax = max(s.config.iter_threshold_max for s in sims) + 1
        self.pred = [max(max(s.config.iter_threshold_min, 1) for s in sims)] * 2
        self.grid = {g: 0 for g in GROUPS}
        self.timer = None
        self.graphs = None
        self.iter_log = []

    # ------------------------------------------------------------ helpers
    def _prepare(self):
        """(Re)build the device scene table from the scenes' descriptors."""
        ns = self.ns
        self._fl = (_lib.TsphObject * ns)(*[s.fluid.struct() for s in self.sims])
        self._layouts = [s.fluid.layout_epoch for s in self.sims]
        self._params = (_lib.TsphParams * ns)(*[s.params for s in self.sims])
        self._bo = (_lib.TsphObject * ns)(*[s.bound.struct() for s in self.sims])
        self._epochs = [s.bound.sums_epoch for s in self.sims]
        _lib.check(self.lib.tsph_batch_prepare(
            self._params, self._fl, self._bo, self._ctrl, self._ws, self._wsb, ns,
            ctypes.c_void_p(self._staging.data_ptr()), ctypes.c_void_p(self.table.data_ptr()),
            self.table.numel(), ctypes.byref(self.info), self._stream()), "tsph_batch_prepare")
        torch.cuda.current_stream(self.device).synchronize()  # staging copied

    def _refresh(self):
        """Rebuild the device scene table when a scene changed on the host
        since it was built:
        * a boundary with new or moved particles is re-sorted first, so no
          batched step searches a stale boundary cell order;
        * a boundary whose kept (b,b) sums were invalidated (written
          positions or masses): the grouped pass A marks the table's copy of
          the flag, which only a rebuild from the host descriptors clears;
        * a fluid sorted on its own (the scene's sph_step) or grown
          (push_part_seq): its front/back buffers or count differ from the
          table's.
        Other host-side changes (a scene's parameters) need refresh()."""
        stale, changed = [], False
        for s, e, le in zip(self.sims, self._epochs, self._layouts):
            if s.bound_needs_sort():
                stale.append(s)
            if s.bound.sums_epoch != e or s.fluid.layout_epoch != le:
                changed = True
        if not stale and not changed:
            return
        if self.graphs is not None:
            raise RuntimeError("a scene changed after capture(): release_graphs() first")
        for s in stale:
            s._sort_bound()
        self._prepare()

    def refresh(self):
        """Rebuild the scene table from the scenes as they are now (after a
        change of a scene's parameters, for instance)."""
        if self.graphs is not None:
            raise RuntimeError("release_graphs() before refresh()")
        for s in self.sims:
            if s.bound_needs_sort():
                s._sort_bound()
        self._prepare()

    def __del__(self):
        h = getattr(self, "_sig_host", None)
        if h is not None and not sys.is_finalizing():
            self.lib.tsph_signal_free(h)
            self._sig_host = None

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _tab(self):
        return ctypes.c_void_p(self.table.data_ptr())

    def _call(self, name, fn, *args):
        if self.timer is None or not self.timer.wants(name):
            _lib.check(fn(*args), name)
            return
        e0, e1 = self.timer.pair()
        e0.record()
        _lib.check(fn(*args), name)
        e1.record()
        self.timer.records.append((name, e0, e1))

    def _adopt(self):
        """The batched sort swapped every scene's front/back state once: each
        Fluid mirrors that (its tensors and descriptor) when next used."""
        for sim in self.sims:
            sim.fluid.note_device_swap()

    def _swap_host(self):
        """The sort's front/back swap applied to the host descriptors."""
        for s in range(self.ns):
            o = self._fl[s]
            for name in _SWAPPED:
                a, b = getattr(o, name), getattr(o, name + "_b")
                setattr(o, name, b)
                setattr(o, name + "_b", a)

    def stop_state(self):
        """(it, stop) of every scene's current solve (one copy to the host)."""
        _lib.check(self.lib.tsph_batch_stop_state(ctypes.byref(self.info), self._tab(),
                                                  ctypes.c_void_p(self._stop_dev.data_ptr()),
                                                  self._stream()), "tsph_batch_stop_state")
        self._stop_host.copy_(self._stop_dev, non_blocking=True)
        torch.cuda.current_stream(self.device).synchronize()
        a = self._stop_host.numpy().reshape(self.ns, 2)
        return a[:, 0].copy(), a[:, 1].copy()

    # -------------------------------------------------------------- phases
    def _sort(self, host=True):
        self._call("sort", self.lib.tsph_batch_sort, ctypes.byref(self.info), self._tab(),
                   self._fl if host else None, self.cfl, self.grid["sort"], self._stream())

    def _density_alpha(self):
        self._call("density_alpha", self.lib.tsph_batch_density_alpha, ctypes.byref(self.info), self._tab(),
                   self.grid["density_alpha"], self._stream())

    def _begin(self, solve, signal=False):
        if signal and self._sig_host is not None:
            self._token = self._token % 0x7FFFFFFF + 1
            _lib.check(self.lib.tsph_batch_solver_begin_signal(ctypes.byref(self.info), self._tab(), solve,
                                                               self._sig_dev, self._token, self._stream()),
                       "tsph_batch_solver_begin_signal")
            return True
        _lib.check(self.lib.tsph_batch_solver_begin(ctypes.byref(self.info), self._tab(), solve,
                                                    self._stream()), "tsph_batch_solver_begin")
        return False

    def _wait_signals(self, launched, token=None):
        """(it, stop) of every scene once each scene's update half-steps have
        decided on the last of the `launched` iterations (or stopped before
        it), read from the scenes' early signals without draining the
        stream (a scene whose solve enters no iteration signals (0, stop) from
        its begin); falls back to the device stop states once the stream has
        drained."""
        a = self._sig_np
        tok = np.uint64(self._token if token is None else token)
        stream = torch.cuda.current_stream(self.device)
        spins = 0
        while True:
            v = a.copy()
            w = v & np.uint64(0xFFFFFFFF)
            stop = (w & np.uint64(1)).astype(np.int64)
            it = (w >> np.uint64(1)).astype(np.int64)
            if np.all(((v >> np.uint64(32)) == tok) & ((stop == 1) | (it >= launched))):
                return it, stop
            spins += 1
            if spins % 64 == 0 and stream.query():
                return self.stop_state()

    def _iteration(self, k):
        bi, st = ctypes.byref(self.info), self._stream()
        self._call("solver_psi", self.lib.tsph_batch_solver_psi, bi, self._tab(), k, self.grid["solver_psi"], st)
        self._call("solver_update", self.lib.tsph_batch_solver_update, bi, self._tab(), k,
                   self.grid["solver_update"], st)

    def _advect(self):
        self._call("advect", self.lib.tsph_batch_advect, ctypes.byref(self.info), self._tab(),
                   self.grid["advect"], self._stream())

    def _tail(self):
        self._call("tail", self.lib.tsph_batch_tail, ctypes.byref(self.info), self._tab(),
                   self.grid["tail"], self._stream())

    def _finish_solve(self, solve, k, signal=False, token=None):
        """Iterations beyond the k launched ones, one at a time, until every
        scene's stop rule has ended its solve (read from the early signals
        when the solve was opened with them, else from the device)."""
        state = (lambda: self._wait_signals(k, token)) if signal else self.stop_state
        its, stops = state()
        while not stops.all() and k < self.kmax:
            self._iteration(k)
            k += 1
            its, stops = state()
        self.pred[solve] = max(int(its.max()), 1)
        return its

    # ----------------------------------------------------------------- step
    def step(self):
        """One sph_step of every scene; returns ([div_iter], [incom_iter])."""
        self._refresh()
        if self.graphs is not None:
            return self._step_graph()
        self._sort()
        self._density_alpha()
        its = []
        for solve in (0, 1):
            sig = self._begin(solve, signal=True)
            p = min(self.pred[solve], self.kmax)
            for k in range(p):
                self._iteration(k)
            its.append(self._finish_solve(solve, p, signal=sig))
            if solve == 0:
                self._advect()
        self._tail()
        self._adopt()
        self.iter_log.append((its[0], its[1]))
        return its[0], its[1]

    def capture(self):
        """Record the step's fixed segments as hipGraphs with the current
        iteration predictions; later steps replay them."""
        p0, p1 = (min(p, self.kmax) for p in self.pred)
        torch.cuda.synchronize(self.device)
        graphs, sig, tokens = [], [], []
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._sort(host=False)
            self._density_alpha()
            sig.append(self._begin(0, signal=True))
            tokens.append(self._token)
            for k in range(p0):
                self._iteration(k)
        graphs.append(g)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._advect()
            sig.append(self._begin(1, signal=True))
            tokens.append(self._token)
            for k in range(p1):
                self._iteration(k)
        graphs.append(g)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._tail()
        graphs.append(g)
        # each replay's solves signal with the two tokens recorded here; every
        # scene's word is rewritten by each solve (by its begin when it
        # enters no iteration), so a word still holding the other solve's
        # token is stale
        self.graphs = (graphs, p0, p1, sig, tokens)

    def release_graphs(self):
        self.graphs = None

    def _step_graph(self):
        (ga, gb, gc), p0, p1, sig, tokens = self.graphs
        ga.replay()
        self._swap_host()
        it0 = self._finish_solve(0, p0, sig[0], tokens[0])
        gb.replay()
        it1 = self._finish_solve(1, p1, sig[1], tokens[1])
        gc.replay()
        self._adopt()
        self.iter_log.append((it0, it1))
        return it0, it1

    # ------------------------------------------------------------- state
    def checkpoint(self):
        out = []
        for sim in self.sims:
            f = sim.fluid
            n = f.part_num
            d = {name: f.t[name][:n].clone() for name in _SWAPPED}
            d["ctrl"] = sim.ctrl.clone()
            out.append(d)
        return out, list(self.pred)

    def restore(self, cp):
        states, pred = cp
        for sim, d in zip(self.sims, states):
            f = sim.fluid
            n = f.part_num
            for name in _SWAPPED:
                f.t[name][:n].copy_(d[name])
            sim.ctrl.copy_(d["ctrl"])
        self.pred = list(pred)

    # ------------------------------------------------------------ autotune
    def autotune(self, candidates=None, reps=2):
        """Per kernel group, the grouped-launch block count with the lowest
        measured time (0 = one block per tile).  The batch's state is
        restored afterwards; results do not depend on the choice."""
        if self.graphs is not None:
            raise RuntimeError("autotune before capture()")
        tiles = self.info.tiles[0]
        if candidates is None:
            candidates = sorted({0} | {c for c in (512, 1024, 2048, 4096, 8192) if c < tiles})
        cp = self.checkpoint()
        report = {}
        for group in GROUPS:
            best = None
            for c in candidates:
                self.grid[group] = c
                self.restore(cp)
                self.timer = EventTimer()
                for _ in range(reps):
                    self.step()
                per = self.timer.per_name()
                self.timer = None
                ms = per.get(group, [0.0, 1])
                t = ms[0] / max(ms[1], 1)
                report.setdefault(group, {})[c] = t
                if best is None or t < best[1]:
                    best = (c, t)
            self.grid[group] = best[0]
        self.restore(cp)
        self.iter_log = []
        torch.cuda.synchronize(self.device)
        self.tuning = report
        return dict(self.grid)


def iteration_counts(batch):
    """[(div_iter, incom_iter)] of the last step, per scene."""
    d, i = batch.iter_log[-1]
    return list(zip(np.asarray(d).tolist(), np.asarray(i).tolist()))
