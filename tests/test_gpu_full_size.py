"""GPU checks at the benchmark configurations' full sizes.

* The 1M-particle 3D dam break (1,000,000 fluid + 206,464 boundary
  particles): one fused step against the FP32 oracle, which runs the same
  step on the host's cores: cell keys and the sort order bit-exact, W and the
  densities bit-exact, every other field within the single-step bar, and the
  total momentum of the step against the oracle's (the scene's mirror
  symmetry in z is checked after ten steps in test_gpu_parity.py).
* The 64 batched dam breaks (10k-200k fluid particles each, drawn from a
  seed): two batched steps bit-identical, scene by scene, to the scenes' own
  sph_step(), and the three smallest scenes against the oracle.
* The differentiable dam break (200k particles) over its full 200 steps,
  checkpointed every 20: each 20-step segment's vector-Jacobian product
  (w.r.t. the state at the segment's start and w.r.t. the viscosity), taken
  from its checkpoint, against FP64 central differences.  The chain rule
  composes the segments, so this verifies the whole 200-step gradient; a
  direct finite difference over 200 steps cannot, because the rollout is
  chaotic (perturbations grow about 100x every 20 steps).
"""
import os

import numpy as np
import pytest
import torch

import oracle as O
from conftest import SCEN, load_scene
from test_gpu_parity import FIELDS_F, TOL, compare_fields, rel_err

pytestmark = pytest.mark.gpu


def test_dambreak_1m_step_matches_oracle():
    from tichi_sph_amd import load_simulation
    print("[case] dambreak_3d_1m, one step against the oracle", flush=True)
    sim = load_simulation(os.path.join(SCEN, "dambreak_3d_1m.config.json"),
                          os.path.join(SCEN, "dambreak_3d_1m.json"))
    cb, sb = load_scene("dambreak_3d_1m")
    o = O.Oracle(cb, sb, 4)
    o.init_scenario()
    assert sim.fluid.part_num == o.n[0] == 1_000_000 and sim.bound.part_num == o.n[1] == 206_464
    s = sim.sph_step(stats=True)
    st = o.step(O.STEP_CFL)
    print(f"  iterations gpu ({s.div_iter}, {s.incom_iter}) oracle ({st.div_iter}, {st.incom_iter}); "
          f"dt {s.dt!r} / {st.dt!r}", flush=True)
    assert (s.div_iter, s.incom_iter) == (st.div_iter, st.incom_iter)
    assert s.dt == st.dt
    assert s.overflow_f == 0 and s.overflow_b == 0
    f = sim.fluid
    n = f.part_num
    # cell keys of the step's sort, bit-exact, and the canonical order: a
    # stable sort of the reference indices by key
    key_o = o.get(0, "neighb_cell_seq")
    assert np.array_equal(f.to_numpy("neighb_cell_seq"), key_o)
    pid = f.t["pid"][:n].cpu().numpy()
    assert np.array_equal(pid, np.argsort(key_o, kind="stable"))
    for obj, fl in ((0, f), (1, sim.bound)):
        for name in ("W", "sph_density", "sph_compression"):
            assert np.array_equal(fl.to_numpy(name), o.get(obj, name)), (obj, name)
    compare_fields(sim, o, FIELDS_F + ["alpha_1", "alpha_2", "normal"], label="1m")
    compare_fields(sim, o, ["alpha"], obj=1, label="1m")
    # total momentum equal to the oracle's (from rest the first step is
    # gravity alone: the pressure solve's psi clamps to 0)
    m = f.to_numpy("mass").astype(np.float64)
    v = f.to_numpy("vel").astype(np.float64)
    vo = o.get(0, "vel").astype(np.float64)
    P, Po = (m[:, None] * v).sum(0), (m[:, None] * vo).sum(0)
    scale = (m[:, None] * np.abs(vo)).sum(0)
    print(f"  momentum gpu {P} oracle {Po} (sum m|v| {scale})", flush=True)
    assert np.all(np.abs(P - Po) <= 1e-5 * scale.max())


def test_dambreak_1m_moving_step_matches_oracle():
    """The 1M scene with a random velocity field (N(0, 0.05^2) m/s per
    component, seeded), so that both Jacobi solves and the viscous term act
    at full size (from rest the pressure solve's psi clamps to 0).  Iteration
    counts and dt equal, W and density bit-exact, every field within the bar
    of the FP32 oracle or, where the FP32 cancellation in psi - psi_0 shows,
    as close to the FP64 oracle as the FP32 oracle itself is."""
    from tichi_sph_amd import load_simulation
    print("[case] dambreak_3d_1m with a random velocity field, one step against the oracles", flush=True)
    sim = load_simulation(os.path.join(SCEN, "dambreak_3d_1m.config.json"),
                          os.path.join(SCEN, "dambreak_3d_1m.json"))
    cb, sb = load_scene("dambreak_3d_1m")
    n = sim.fluid.part_num
    v0 = np.random.default_rng(5).normal(0, 0.05, (n, 3)).astype(np.float32)
    sim.fluid.from_numpy("vel", v0)
    oracles = []
    for prec in (4, 8):
        o = O.Oracle(cb, sb, prec)
        o.init_scenario()
        o.set(0, "vel", v0)
        oracles.append((o, o.step(O.STEP_CFL)))
    (o, st), (o64, st64) = oracles
    s = sim.sph_step(stats=True)
    print(f"  iterations gpu ({s.div_iter}, {s.incom_iter}) oracle ({st.div_iter}, {st.incom_iter}) "
          f"FP64 oracle ({st64.div_iter}, {st64.incom_iter}); dt {s.dt!r} / {st.dt!r}", flush=True)
    assert (s.div_iter, s.incom_iter) == (st.div_iter, st.incom_iter)
    assert s.dt == st.dt
    f = sim.fluid
    for name in ("W", "sph_density", "sph_compression"):
        assert np.array_equal(f.to_numpy(name), o.get(0, name)), name
    assert float(np.abs(o.get(0, "psi_adv")).max()) > 0  # the solve acted
    for name in FIELDS_F + ["alpha_1", "alpha_2", "normal"]:
        g, r, r64 = f.to_numpy(name), o.get(0, name), o64.get(0, name)
        e32, eg, eo = rel_err(g, r), rel_err(g, r64), rel_err(r, r64)
        print(f"  1m-moving {name:16s} vs oracle {e32:.2e}; vs FP64: gpu {eg:.2e} oracle {eo:.2e}", flush=True)
        assert e32 <= TOL or eg <= 2 * eo + TOL / 10, name


def test_dambreak_1m_300_steps_stay_physical():
    """300 steps of the 1M dam break (30 ms of flow): no list overflow, every
    solve within its iteration bounds, a finite state, constant total mass,
    the centre of mass moving down and along +x as the column collapses, and
    the potential energy falling.  The oracle would take about ten minutes
    for this run, so these size-independent properties stand in for it."""
    from tichi_sph_amd import load_simulation
    print("[case] dambreak_3d_1m, 300 steps", flush=True)
    sim = load_simulation(os.path.join(SCEN, "dambreak_3d_1m.config.json"),
                          os.path.join(SCEN, "dambreak_3d_1m.json"))
    f = sim.fluid
    m0 = f.to_numpy("mass").astype(np.float64)
    x0 = (m0[:, None] * f.to_numpy("pos").astype(np.float64)).sum(0) / m0.sum()
    pe = []
    for step in range(1, 301):
        st = sim.sph_step(stats=True)
        assert st.overflow_f == 0 and st.overflow_b == 0, step
        assert 2 <= st.div_iter <= 101 and 2 <= st.incom_iter <= 101, step
        if step % 50 == 0:
            pe.append(st.gravity_potential_energy)
            assert np.isfinite(f.to_numpy("vel")).all(), step
    m = f.to_numpy("mass").astype(np.float64)
    x = (m[:, None] * f.to_numpy("pos").astype(np.float64)).sum(0) / m.sum()
    print(f"  centre of mass {x0} -> {x}; potential energy every 50 steps {pe}; "
          f"last iterations ({st.div_iter}, {st.incom_iter})", flush=True)
    assert np.array_equal(m, m0)
    assert x[1] < x0[1] and x[0] > x0[0]
    assert all(b < a for a, b in zip(pe, pe[1:]))


def test_batched_64_dambreaks_bit_identical_and_smallest_against_oracle():
    from tichi_sph_amd import simulation_from_dicts
    from tichi_sph_amd.batch import SceneBatch, iteration_counts
    from tichi_sph_amd.scenarios.make_scenarios import batched_dambreaks
    scenes = batched_dambreaks(64, seed=0)
    print(f"[case] 64 batched dam breaks, {sum(1 for _ in scenes)} scenes", flush=True)
    sims = [simulation_from_dicts(c, s, cfl=False) for _, s, c in scenes]
    nf = [sm.fluid.part_num for sm in sims]
    print(f"  fluid particles {min(nf)} .. {max(nf)}, total {sum(nf)}", flush=True)
    assert len(sims) == 64 and min(nf) >= 8_000 and max(nf) <= 220_000
    batch = SceneBatch(sims)
    cp = batch.checkpoint()
    states, iters = [], []
    for _ in range(2):
        batch.step()
        iters.append(iteration_counts(batch))
        states.append([(sm.fluid.to_numpy("pos"), sm.fluid.to_numpy("vel")) for sm in sims])
    torch.cuda.synchronize()
    # the same scenes, each stepped on its own from the same state
    batch.restore(cp)
    for step in range(2):
        for s, sm in enumerate(sims):
            st = sm.sph_step(stats=True)
            assert (st.div_iter, st.incom_iter) == tuple(iters[step][s]), (step, s)
            pos, vel = states[step][s]
            assert np.array_equal(sm.fluid.to_numpy("pos"), pos), (step, s)
            assert np.array_equal(sm.fluid.to_numpy("vel"), vel), (step, s)
    print(f"  iterations of the last step (first 8 scenes): {iters[1][:8]}", flush=True)
    # the three smallest scenes against the oracle (first batched step)
    for s in np.argsort(nf)[:3]:
        name, sb, cb = scenes[s]
        sm = simulation_from_dicts(cb, sb, cfl=False)
        o = O.Oracle(cb, sb, 4)
        o.init_scenario()
        o.scalar("dt", 0, sm.control().dt)
        one = SceneBatch([sm])
        di, ii = one.step()
        stp = o.step(0)
        assert (int(di[0]), int(ii[0])) == (stp.div_iter, stp.incom_iter), name
        assert (int(di[0]), int(ii[0])) == tuple(iters[0][s]), name
        # the batch of 64 gave the same first step as this batch of one
        assert np.array_equal(sm.fluid.to_numpy("pos"), states[0][s][0]), name
        compare_fields(sm, o, FIELDS_F, label=f"batched {name}")


def _seg_fd(r, ck, seg_log, every, wx, wv, dx, dv, mu, gam, eps, dmu, tapes=None):
    """Central differences of the functional wx . x + wv . v after `every`
    steps from checkpoint `ck`, along (dx, dv) and along mu.  With `tapes`
    (the segment's recorded steps) every step replays the recorded psi clamp
    set (tsph_solver_psi_frozen), so the differences follow the recorded path
    whose derivative the adjoint takes; without, they follow the step as it
    runs, whose clamps (sph.py:219-226) flip under the perturbation."""
    f = r.sim.fluid
    n, d = f.part_num, r.sim.config.dim

    def J(e, m):
        r.restore(ck)
        r.set_params(m, gam)
        pid = f.t["pid"][:n].long()
        f.t["pos"][:n, :d] += e * dx[pid]
        f.t["vel"][:n, :d] += e * dv[pid]
        r.force, r.step_index = seg_log, 0
        for t in range(every):
            if tapes is None:
                r.step()
            else:
                r.step(frozen=tapes[t])
                # the recorded clamp set is indexed by storage slot
                assert torch.equal(f.t["pid"][:n], tapes[t].pid)
        r.force = None
        return float((r.positions_ref() * wx).sum() + (r.velocities_ref() * wv).sum())

    fd_state = (J(eps, mu) - J(-eps, mu)) / (2 * eps)
    fd_mu = (J(0.0, mu + dmu) - J(0.0, mu - dmu)) / (2 * dmu)
    r.set_params(mu, gam)
    return fd_state, fd_mu


def test_differentiable_200_steps_segment_vjps_match_fd_fp64():
    """Each 20-step segment's VJP against central differences along the
    recorded path (the psi clamp set of the recorded steps replayed): the
    adjoint is the derivative of that path.  Differences of the free-running
    step are printed beside them: at every scale some of the 200k particles
    sit at the clamp's kink, so those do not converge to any derivative."""
    from tichi_sph_amd import simulation_from_dicts
    from tichi_sph_amd.adjoint import DiffRunner
    steps, every = 200, 20
    mu, gam = 0.01, 0.07
    cb, sb = load_scene("diff_3d_200k")
    print("[case] diff_3d_200k, 200 steps FP64, per-segment VJPs", flush=True)
    r = DiffRunner(simulation_from_dicts(cb, sb, precision=8, cfl=False))
    f = r.sim.fluid
    n, d = f.part_num, r.sim.config.dim
    dev = r.sim.device
    g = torch.Generator().manual_seed(0)
    v0 = (0.01 * torch.randn(n, d, generator=g, dtype=torch.float64)).to(dev)
    r.set_params(mu, gam)
    r.set_velocity_ref(v0)
    r.step_index, r.iter_log, r.force = 0, [], None
    ckpts = []
    for t in range(steps):
        if t % every == 0:
            ckpts.append(r.checkpoint())
        r.step()
    log = list(r.iter_log)
    worst = 0.0
    g = torch.Generator().manual_seed(21)
    for s, ck in enumerate(ckpts):
        seg_log = log[s * every:(s + 1) * every]
        wx, wv, dx, dv = (torch.randn(n, d, generator=g, dtype=torch.float64).to(dev) for _ in range(4))
        dx *= 1e-4  # a position step of dt times the velocity step
        r.restore(ck)
        r.force, r.step_index = seg_log, 0
        tapes = [r.step(tape=True) for _ in range(every)]
        r.force = None
        # differences along the recorded path, before the reverse consumes the tapes
        # steps: along the recorded path the map is smooth, and a scan of the
        # velocity step 1e-10 .. 3e-8 (mu step 1e-5 .. 3e-3 relative) shows the
        # differences converging to the adjoint as the step grows (round-off
        # of the functional, a sum of ~6e5 O(1) terms, dominates below 1e-9)
        fd_state, fd_mu = _seg_fd(r, ck, seg_log, every, wx, wv, dx, dv, mu, gam, 3e-8, 3e-3 * mu, tapes)
        free_state, free_mu = _seg_fd(r, ck, seg_log, every, wx, wv, dx, dv, mu, gam, 1e-10, 1e-4 * mu)
        xbar = torch.zeros(n, 4, dtype=torch.float64, device=dev)
        vbar = torch.zeros_like(xbar)
        xbar[:, :d], vbar[:, :d] = wx, wv
        r.grad.zero_()
        for tp in reversed(tapes):
            r.reverse(tp, xbar, vbar)
        del tapes
        ad_state = float((xbar[:, :d] * dx).sum() + (vbar[:, :d] * dv).sum())
        ad_mu = float(r.grad[0])
        es = abs(ad_state - fd_state) / abs(fd_state)
        em = abs(ad_mu - fd_mu) / abs(fd_mu)
        worst = max(worst, es, em)
        print(f"  segment {s} (steps {s * every}-{(s + 1) * every}): state VJP adjoint {ad_state:+.8e} "
              f"fd {fd_state:+.8e} rel {es:.2e} (free-running fd {free_state:+.4e}); d/d mu adjoint "
              f"{ad_mu:+.8e} fd {fd_mu:+.8e} rel {em:.2e} (free-running fd {free_mu:+.4e})", flush=True)
        assert es <= 1e-3 and em <= 1e-3, s
    print(f"  largest relative error over the 10 segments {worst:.2e}", flush=True)
