"""Benchmark: DFSPH particle-steps/s on the 1M-particle 3D dam break.

One step = one fused sph_step (neighbour sort, density/alpha pass, both
Jacobi solves to their stop rule, advection, tail) over the 1,000,000-fluid /
206,464-boundary particle scene (tichi_sph_amd/scenarios/dambreak_3d_1m),
inputs resident in HBM.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
  python bench.py --copy-calibration MIB   # plain copy kernel of known size

Roofline: algorithmic bytes of the dominant kernel per launch (per-particle
figures of the minimal-pass schedule) divided by its average launch duration
from HIP events recorded on the launch stream inside the timed region.  Only
that kernel group's launches, on every 4th timed step, carry events (an event
pair on every launch costs about 90 us per step); the per-group breakdown
(`kernel_ms_per_step`) and the choice of the dominant group come from an
untimed pass of the same steps with every launch timed.
`traffic` is the dominant kernel's HBM bytes per launch from FETCH_SIZE and
WRITE_SIZE counter runs, each scaled by what the same counter reads for the
copy kernel of --copy-calibration; the figures are read from
profiles/traffic_dambreak_3d_1m.json (null when that file is absent).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "DFSPH particle-steps/s @1M particles, fwd vs fwd+adjoint; % HBM peak"
WORKLOAD = "dambreak_3d_1m"

# Algorithmic bytes per particle and launch (FP32, 3D; s = 4 B, v = 16 B with
# the 4th lane counted as the vector's padding-free 12 B + X where it carries X).
# fluid: P reads pos, vel_adv, X, psi, psi0 and writes psi_adv (2v + 4s);
# U reads pos, vel_adv, psi_adv, alpha, X and writes vel_adv (3v + 3s);
# A reads pos, X and writes psi, alpha (v + 3s).  Boundary particles: pos and
# X per pass.  v = 12 B, s = 4 B.
V, S = 12, 4
ALG_BYTES = {
    "solver_psi": lambda nf, nb: nf * (2 * V + 4 * S) + nb * (V + S),
    "solver_update": lambda nf, nb: nf * (3 * V + 3 * S) + nb * (V + S),
    "density_alpha": lambda nf, nb: nf * (V + 3 * S) + nb * (2 * V + 4 * S),
    "advect": lambda nf, nb: nf * (3 * V + 3 * S),
}


def step_bytes(nf, nb, nc, I):
    """B_step of the minimal-pass schedule (3D, FP32)."""
    return nf * ((10 * V + 11 * S) + I * (5 * V + 7 * S)) + nb * (2 * V + 4 * S) * (1 + I) + 16 * nc


def measured_traffic(kernel):
    """HBM bytes per launch of a bench kernel group from the committed counter
    summary (FETCH_SIZE + WRITE_SIZE, KiB per dispatch, copy-calibrated)."""
    path = os.path.join(ROOT, "profiles", "traffic_" + WORKLOAD + ".json")
    if not os.path.exists(path):
        return None
    d = json.load(open(path))
    g = d.get("groups", {}).get(kernel)
    if g is None:
        return None
    return int(g["fetch_kib"] * 1024 * d["fetch_factor"] + g["write_kib"] * 1024 * d["write_factor"])


def copy_calibration(mib, reps=5):
    """A plain elementwise copy (read N bytes, write N bytes) for scaling the
    FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, device="cuda", dtype=torch.float32)
    dst = torch.empty_like(src)
    for _ in range(reps):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 4, "reps": reps,
                      "kernel": "vectorized_elementwise_kernel (MulFunctor)"}), flush=True)


TAPE_BUDGET = 32 << 30  # bytes of HBM (of 288 GB) for kept segment tapes in the differentiable legs


def adjoint_leg(scen, steps, every):
    """fwd vs fwd+adjoint on the same workload (fixed dt, which the CFL
    condition never lowers here): `steps` plain forward steps, then a
    differentiable rollout of `steps` steps (checkpoint every `every`) with
    the backward pass of a position loss; wall time of each, ms per step.
    `ratio` keeps earlier segments' tapes within TAPE_BUDGET bytes;
    `ratio_replay_all` replays every segment but the last from its checkpoint
    (tape budget 0).  Both give bit-identical gradients."""
    import torch
    from tichi_sph_amd import load_simulation
    from tichi_sph_amd.adjoint import DiffRunner, rollout

    sim = load_simulation(os.path.join(scen, WORKLOAD + ".config.json"), os.path.join(scen, WORKLOAD + ".json"),
                          cfl=False)
    r = DiffRunner(sim)
    st0 = r.checkpoint()
    nf, d = sim.fluid.part_num, sim.config.dim
    dev, dt_ = sim.device, sim.fluid.dtype
    x_init = r.positions_ref().clone()
    mu0, gam0 = sim.params.dynamic_viscosity, sim.params.surface_tension_gamma

    def diff_run(k, budget):
        r.restore(st0)
        v = torch.zeros(nf, d, dtype=dt_, device=dev, requires_grad=True)
        mu = torch.tensor(mu0, dtype=dt_, device=dev, requires_grad=True)
        # the plain DFSPH workload has no surface tension (gamma = 0): its
        # gradients are those w.r.t. the initial velocity field and viscosity
        gam = torch.tensor(gam0, dtype=dt_, device=dev, requires_grad=gam0 != 0.0)
        x, _ = rollout(r, v, mu, gam, k, every, tape_budget=budget)
        ((x - x_init) ** 2).sum(1).mean().backward()
        return v, mu, gam

    diff_run(steps, TAPE_BUDGET)  # warm-up: first launches, and the tapes' first allocations
    r.restore(st0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        r.step()
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0

    def timed(budget):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = diff_run(steps, budget)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, out

    t_adj, (v, mu, gam) = timed(TAPE_BUDGET)
    t_rep, (v2, mu2, _) = timed(0)
    same = bool(torch.equal(v.grad, v2.grad) and float(mu.grad) == float(mu2.grad))
    from tichi_sph_amd import EventTimer
    sim.timer = EventTimer()  # per-call breakdown from a separate, untimed rollout
    diff_run(steps, TAPE_BUDGET)
    per = sim.timer.per_name()
    sim.timer = None
    r.restore(st0)
    tp = r.step(tape=True)
    tape_mb = tp.nbytes() / 1e6
    del tp
    ms_f, ms_a, ms_r = t_fwd / steps * 1e3, t_adj / steps * 1e3, t_rep / steps * 1e3
    return {"steps": steps, "checkpoint_every": every, "tape_budget_bytes": TAPE_BUDGET,
            "ms_per_step_fwd": round(ms_f, 4), "ms_per_step_fwd_adj": round(ms_a, 4),
            "ratio": round(ms_a / ms_f, 3), "ms_per_step_fwd_adj_replay_all": round(ms_r, 4),
            "ratio_replay_all": round(ms_r / ms_f, 3), "gradients_identical": same,
            "value_fwd_adj": round(nf * steps / t_adj, 1),
            "grads": ["initial velocity", "dynamic_viscosity"] + (["surface_tension_gamma"] if gam0 else []),
            "grad_finite": bool(torch.isfinite(v.grad).all().item() and math.isfinite(float(mu.grad))),
            "tape_MB_per_step": round(tape_mb, 1),
            "kernel_ms_per_step": {k: round(x[0] / steps, 4) for k, x in sorted(per.items())}}


DIFF_WORKLOAD = "diff_3d_200k"


def differentiable_leg(scen, steps=200, every=20):
    """The differentiable dam break: 200k fluid particles, `steps` steps
    checkpointed every `every` (tapes kept within TAPE_BUDGET, and, for
    `ratio_replay_all`, none kept but the last segment's); loss =
    mean |x - x_target|^2 with the target
    cloud from a forward run at mu = 0.02, gamma = 0.05; gradients w.r.t.
    mu, gamma and the initial velocity field, from mu = 0.01, gamma = 0.07
    and v0 ~ N(0, 0.01^2) per component (numpy default_rng(0))."""
    import numpy as np
    import torch
    from tichi_sph_amd import load_simulation
    from tichi_sph_amd.adjoint import DiffRunner, rollout

    sim = load_simulation(os.path.join(scen, DIFF_WORKLOAD + ".config.json"),
                          os.path.join(scen, DIFF_WORKLOAD + ".json"), cfl=False)
    r = DiffRunner(sim)
    st0 = r.checkpoint()
    nf, d = sim.fluid.part_num, sim.config.dim
    dev, dt_ = sim.device, sim.fluid.dtype
    v0 = torch.as_tensor(np.random.default_rng(0).normal(0.0, 0.01, (nf, d)), dtype=dt_, device=dev)

    def forward(mu, gam):
        r.restore(st0)
        r.set_params(mu, gam)
        r.set_velocity_ref(v0)
        for _ in range(steps):
            r.step()
        return r.positions_ref()

    target = forward(0.02, 0.05).clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    forward(0.01, 0.07)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    def diff_run(budget):
        r.restore(st0)
        v = v0.clone().requires_grad_(True)
        mu = torch.tensor(0.01, dtype=dt_, device=dev, requires_grad=True)
        gam = torch.tensor(0.07, dtype=dt_, device=dev, requires_grad=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x, _ = rollout(r, v, mu, gam, steps, every, tape_budget=budget)
        loss = ((x - target) ** 2).sum(1).mean()
        loss.backward()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, loss, v, mu, gam

    diff_run(TAPE_BUDGET)  # warm-up: the tapes' first allocations are not part of the step cost
    t_adj, loss, v, mu, gam = diff_run(TAPE_BUDGET)
    its = r.iter_log
    t_rep, _, v2, mu2, gam2 = diff_run(0)
    same = bool(torch.equal(v.grad, v2.grad) and float(mu.grad) == float(mu2.grad)
                and float(gam.grad) == float(gam2.grad))
    g_mu, g_gam, g_v = float(mu.grad), float(gam.grad), float(v.grad.double().norm())
    from tichi_sph_amd import EventTimer
    sim.timer = EventTimer()  # per-call breakdown of forward + adjoint from a separate, untimed rollout
    diff_run(TAPE_BUDGET)
    per = sim.timer.per_name()
    sim.timer = None
    return {"workload": DIFF_WORKLOAD, "fluid_particles": nf, "steps": steps, "checkpoint_every": every,
            "tape_budget_bytes": TAPE_BUDGET,
            "ms_per_step_fwd": round(t_fwd / steps * 1e3, 4), "ms_per_step_fwd_adj": round(t_adj / steps * 1e3, 4),
            "ratio": round(t_adj / t_fwd, 3), "ms_per_step_fwd_adj_replay_all": round(t_rep / steps * 1e3, 4),
            "ratio_replay_all": round(t_rep / t_fwd, 3), "gradients_identical": same,
            "loss": float(loss.detach()),
            "grad_mu": g_mu, "grad_gamma": g_gam, "grad_v0_norm": g_v,
            "mean_iterations": [sum(a for a, _ in its) / len(its), sum(b for _, b in its) / len(its)],
            "kernel_ms_per_step": {k: round(x[0] / steps, 4) for k, x in sorted(per.items())}}


def batched_leg(steps=20, n_scenes=64):
    """Batched dam breaks: `n_scenes` independent 3D DFSPH scenes (fluid
    counts log-uniform in [10k, 200k], numpy default_rng(0)), advanced
    sequentially (each Simulation's own sph_step) and as one batch (one
    grouped launch per kernel, autotuned block counts), eager and replayed
    from hipGraphs; fluid particle-steps/s of each."""
    import torch
    from tichi_sph_amd import simulation_from_dicts
    from tichi_sph_amd.batch import SceneBatch
    from tichi_sph_amd.scenarios.make_scenarios import batched_dambreaks

    sims = [simulation_from_dicts(c, sc, cfl=False) for _, sc, c in batched_dambreaks(n_scenes, seed=0)]
    nf = sum(s.fluid.part_num for s in sims)
    names = ("pos", "vel", "mass", "rest_density", "rest_volume", "volume_frac", "pid")
    cp = [({k: s.fluid.t[k][:s.fluid.part_num].clone() for k in names}, s.ctrl.clone()) for s in sims]

    def restore():
        for s, (d, c) in zip(sims, cp):
            for k in names:
                s.fluid.t[k][:s.fluid.part_num].copy_(d[k])
            s.ctrl.copy_(c)

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    for s in sims:  # warm-up (also sorts the boundaries)
        s.sph_step()
    restore()
    t_seq = timed(lambda: [s.sph_step() for s in sims])
    restore()
    batch = SceneBatch(sims)
    grid = batch.autotune()
    batch.step()
    restore()
    t_batch = timed(batch.step)
    restore()
    batch.step()
    batch.capture()
    restore()
    t_graph = timed(batch.step)
    its = batch.iter_log[-1]
    return {"workload": f"batched_dambreaks_{n_scenes}", "scenes": n_scenes, "fluid_particles": nf,
            "steps": steps, "sequential_value": round(nf * steps / t_seq, 1),
            "batched_value": round(nf * steps / t_batch, 1), "graph_value": round(nf * steps / t_graph, 1),
            "ms_per_step_sequential": round(t_seq / steps * 1e3, 3),
            "ms_per_step_batched": round(t_batch / steps * 1e3, 3),
            "ms_per_step_graph": round(t_graph / steps * 1e3, 3),
            "speedup_batched": round(t_seq / t_batch, 3), "speedup_graph": round(t_seq / t_graph, 3),
            "autotuned_grid": grid, "max_iterations_last_step": [int(its[0].max()), int(its[1].max())]}


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or the platform string."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.lower().startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or platform.machine()


def cpu_baseline(steps=5, warmup=2):
    """The oracle (C++ restatement of the reference step, reference schedule:
    atomic bucket grid each step, unsorted storage, separate (obj, nobj)
    sweeps; g++ -O2 -fopenmp, no fast-math) on this host's cores, on a
    bounded sample of the same workload: `warmup` untimed steps, then
    `steps` timed ones (about 10-15 s of host work at 1M particles)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    cfg = json.load(open(os.path.join(ROOT, "tichi_sph_amd", "scenarios", WORKLOAD + ".config.json")))
    scn = json.load(open(os.path.join(ROOT, "tichi_sph_amd", "scenarios", WORKLOAD + ".json")))
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))
    O.lib().oracle_set_threads(cores)
    o = O.Oracle(cfg, scn, 4)
    o.init_scenario()
    for _ in range(warmup):
        o.step(O.STEP_CFL)
    t0 = time.perf_counter()
    its = []
    for _ in range(steps):
        st = o.step(O.STEP_CFL)
        its.append((st.div_iter, st.incom_iter))
    t = time.perf_counter() - t0
    O.lib().oracle_set_threads(1)
    nf = o.n[0]
    return {"value": nf * steps / t, "unit": "particle-steps/s", "cores": cores, "kind": "port",
            "cpu": cpu_model(),
            "sample": f"{WORKLOAD}: {steps} steps after {warmup} warm-up steps ({nf} fluid particles), "
                      f"{t:.1f} s wall, iterations {its}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-timers", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB")
    ap.add_argument("--adjoint-steps", type=int, default=20)
    ap.add_argument("--checkpoint-every", type=int, default=10)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batched-steps", type=int, default=20)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: tichi_sph_amd has no multi-GPU path (--gpus must be 1)",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return

    import torch
    from tichi_sph_amd import load_simulation

    scen = os.path.join(ROOT, "tichi_sph_amd", "scenarios")
    sim = load_simulation(os.path.join(scen, WORKLOAD + ".config.json"), os.path.join(scen, WORKLOAD + ".json"))
    nf, nb, nc = sim.fluid.part_num, sim.bound.part_num, sim.config.node_num
    from tichi_sph_amd import EventTimer
    # warm-up: at least `warmup` steps and 2 s, so that clocks and caches of a
    # fresh device have settled before anything is timed
    t0 = time.perf_counter()
    n = 0
    while n < args.warmup or (time.perf_counter() - t0 < 2.0 and n < 20 * args.warmup + 100):
        sim.sph_step()
        n += 1
    torch.cuda.synchronize()
    breakdown, dom = {}, None
    if not args.no_timers:
        prof_steps = max(1, min(10, args.steps))
        sim.timer = EventTimer()
        for _ in range(prof_steps):
            sim.sph_step()
        breakdown = {k: v[0] / prof_steps for k, v in sim.timer.per_name().items()}
        dom = max((k for k in breakdown if k in ALG_BYTES), key=lambda k: breakdown[k])
    tdom = EventTimer(16 * (args.steps + 2) + 256, only={dom}) if dom else None
    iters = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s_ in range(args.steps):
        # the dominant group's launches of every 4th step carry events
        sim.timer = tdom if s_ % 4 == 0 else None
        iters.append(sim.sph_step())
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall / args.steps * 1e3
    value = nf * args.steps / wall

    per = tdom.per_name() if tdom else {}
    sim.timer = None
    I_mean = sum(a + b for a, b in iters) / len(iters)
    roof = None
    if dom in per:
        avg_ms = per[dom][0] / per[dom][1]
        alg = ALG_BYTES[dom](nf, nb)
        ach = alg / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom), "kernel": dom,
                "avg_launch_us": round(avg_ms * 1e3, 2), "alg_bytes_per_launch": alg}
    B = step_bytes(nf, nb, nc, I_mean)
    step_gbs = B / (ms_per_step * 1e-3) / 1e9
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (lattice dam break generated from the scenario JSON)",
        "config": {"workload": WORKLOAD, "fluid_particles": nf, "boundary_particles": nb, "cells": nc,
                   "solver": "DFSPH", "dt": sim.config.dt, "parallelism": "single GPU"},
        "roofline": roof,
        "step_roofline": {"alg_bytes_per_step": int(B), "achieved_GBs": round(step_gbs, 1),
                          "frac": round(step_gbs / HBM_PEAK_GBS, 4)},
        "iterations": {"mean_div": sum(a for a, _ in iters) / len(iters),
                       "mean_incom": sum(b for _, b in iters) / len(iters),
                       "max_div": max(a for a, _ in iters), "max_incom": max(b for _, b in iters)},
        "kernel_ms_per_step": {k: round(v, 4) for k, v in sorted(breakdown.items())},
    }
    if not args.no_adjoint:
        del sim
        torch.cuda.empty_cache()
        out["adjoint"] = adjoint_leg(scen, args.adjoint_steps, args.checkpoint_every)
        torch.cuda.empty_cache()
        out["adjoint_" + DIFF_WORKLOAD] = differentiable_leg(scen)
    if not args.no_batched:
        torch.cuda.empty_cache()
        out["batched"] = batched_leg(args.batched_steps)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
