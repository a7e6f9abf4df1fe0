"""GPU tests of the hand-written adjoint (tichi_sph_amd/adjoint.py,
csrc/tsph_adjoint.hip) through the C ABI.

The reference has no autodiff, so the gradient is pinned by the mathematics:
FP64 central finite differences of the same rollout with the recorded
Jacobi iteration counts replayed (the derivative taken is that of the
recorded step; a changed iteration count is a discontinuity of the
reference's stop rule, not a derivative).  Bar: relative error <= 1e-3.

FD steps: 1e-4 relative for the scalar parameters (mu, gamma).  For the
per-particle fields the step must stay below the scale at which a psi clamp
(psi = max(psi, 0), sph.py:219-226) of some particle flips.  Measured on the
8k-particle 3D dam break, 3 steps: the FD quotient's distance to the adjoint
is 1.5e-1 / 2.7e-3 / 1.7e-7 / 3.4e-5 for position steps of 1e-9 / 1e-10 /
1e-11 / 1e-12 m (kinks, then round-off), so positions use 1e-11 m and
velocities 1e-8 m/s.
"""
import numpy as np
import pytest
import torch

from conftest import load_scene

pytestmark = pytest.mark.gpu

FD_TOL = 1e-3


def runner_for(name, precision=8, **over):
    from tichi_sph_amd import simulation_from_dicts
    from tichi_sph_amd.adjoint import DiffRunner
    cb, sb = load_scene(name)
    cb = dict(cb, **over)
    print(f"[case] {name} precision={precision} {over}", flush=True)
    sim = simulation_from_dicts(cb, sb, precision=precision, cfl=False)
    return DiffRunner(sim)


def seeded_velocity(runner, seed=0, scale=0.01):
    n, d = runner.sim.fluid.part_num, runner.sim.config.dim
    g = torch.Generator().manual_seed(seed)
    return (torch.randn(n, d, generator=g, dtype=torch.float64) * scale).to(
        runner.sim.device, runner.sim.fluid.dtype)


def forward_only(runner, st0, v0, mu, gamma, steps, force=None):
    runner.restore(st0)
    runner.set_params(mu, gamma)
    runner.set_velocity_ref(v0)
    runner.force = force
    runner.step_index = 0
    runner.iter_log = []
    for _ in range(steps):
        runner.step()
    runner.force = None
    return runner.positions_ref()


def loss_of(x, target):
    return ((x - target) ** 2).sum(1).mean()


def setup_case(name, steps, precision=8, mu=0.01, gamma=0.07, **over):
    from tichi_sph_amd.adjoint import rollout
    r = runner_for(name, precision, **over)
    st0 = r.checkpoint()
    v0 = seeded_velocity(r)
    target = forward_only(r, st0, v0, 0.02, 0.05, steps).detach()
    r.restore(st0)
    return r, st0, v0, target, rollout


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


@pytest.mark.parametrize("name,steps,every", [("dambreak_3d_8k", 6, 2), ("dambreak_2d_8k", 6, 3)])
def test_adjoint_matches_central_fd_fp64(name, steps, every):
    r, st0, v0, target, rollout = setup_case(name, steps)
    mu = torch.tensor(0.01, dtype=torch.float64, device=v0.device, requires_grad=True)
    gam = torch.tensor(0.07, dtype=torch.float64, device=v0.device, requires_grad=True)
    v = v0.clone().requires_grad_(True)
    x, _ = rollout(r, v, mu, gam, steps, every)
    L = loss_of(x, target)
    L.backward()
    base_log = list(r.iter_log)
    print(f"  iterations {base_log}", flush=True)

    def L_at(mu_, gam_, v_):
        xx = forward_only(r, st0, v_, mu_, gam_, steps, force=base_log)
        return float(loss_of(xx, target))

    g = torch.Generator().manual_seed(7)
    dv = torch.randn(v0.shape, generator=g, dtype=torch.float64).to(v0.device)
    checks = []
    e = 1e-4 * 0.01
    checks.append(("mu", float(mu.grad), (L_at(0.01 + e, 0.07, v0) - L_at(0.01 - e, 0.07, v0)) / (2 * e)))
    e = 1e-4 * 0.07
    checks.append(("gamma", float(gam.grad), (L_at(0.01, 0.07 + e, v0) - L_at(0.01, 0.07 - e, v0)) / (2 * e)))
    e = 1e-8
    fd = (L_at(0.01, 0.07, v0 + e * dv) - L_at(0.01, 0.07, v0 - e * dv)) / (2 * e)
    checks.append(("v0 . dv", float((v.grad * dv).sum()), fd))
    bad = []
    for what, ad, fd in checks:
        print(f"  {what:8s} adjoint={ad:+.10e} fd={fd:+.10e} rel={rel(ad, fd):.2e}", flush=True)
        if rel(ad, fd) > FD_TOL:
            bad.append(what)
    assert not bad, bad


def test_adjoint_at_differentiable_dam_break_size_fp64():
    """The 200k-particle differentiable dam break (target from mu = 0.02,
    gamma = 0.05; start mu = 0.01, gamma = 0.07, v0 ~ N(0, 0.01^2)) over 20
    steps checkpointed every 5, against FP64 central differences along the
    recorded path: the replayed steps clamp psi where the recorded iterations
    did (tsph_solver_psi_frozen), which is the function whose derivative the
    adjoint takes.  Differences of the free-running step are printed beside
    them: some of the 200k particles sit at the clamp's kink at every scale,
    so along v0 they reach only ~1e-3 (4.2e-2 / 4.2e-3 / 8.5e-4 / 1.0e-2 for
    steps 1e-8 / 1e-9 / 1e-10 / 1e-11).  Steps 3e-8 (v0) and 3e-3 relative
    (mu, gamma): along the recorded path the differences converge as the
    step grows, round-off dominating below about 1e-9.  Bar 1e-3."""
    steps, every = 20, 5
    r, st0, v0, target, rollout = setup_case("diff_3d_200k", steps)
    mu = torch.tensor(0.01, dtype=torch.float64, device=v0.device, requires_grad=True)
    gam = torch.tensor(0.07, dtype=torch.float64, device=v0.device, requires_grad=True)
    v = v0.clone().requires_grad_(True)
    x, _ = rollout(r, v, mu, gam, steps, every)
    loss_of(x, target).backward()
    base_log = list(r.iter_log)
    print(f"  iterations {base_log}", flush=True)
    # the recorded path's clamp sets
    r.restore(st0)
    r.set_params(0.01, 0.07)
    r.set_velocity_ref(v0)
    r.force, r.step_index = base_log, 0
    tapes = [r.step(tape=True) for _ in range(steps)]
    r.force = None

    def L_at(mu_, gam_, v_, frozen=True):
        r.restore(st0)
        r.set_params(mu_, gam_)
        r.set_velocity_ref(v_)
        r.force, r.step_index = base_log, 0
        for t in range(steps):
            r.step(frozen=tapes[t] if frozen else None)
        r.force = None
        return float(loss_of(r.positions_ref(), target))

    g = torch.Generator().manual_seed(3)
    dv = torch.randn(v0.shape, generator=g, dtype=torch.float64).to(v0.device)
    em, eg, ev = 3e-3 * 0.01, 3e-3 * 0.07, 3e-8
    checks = [("mu", float(mu.grad), (L_at(0.01 + em, 0.07, v0) - L_at(0.01 - em, 0.07, v0)) / (2 * em)),
              ("gamma", float(gam.grad), (L_at(0.01, 0.07 + eg, v0) - L_at(0.01, 0.07 - eg, v0)) / (2 * eg)),
              ("v0 . dv", float((v.grad * dv).sum()),
               (L_at(0.01, 0.07, v0 + ev * dv) - L_at(0.01, 0.07, v0 - ev * dv)) / (2 * ev))]
    free = (L_at(0.01, 0.07, v0 + 1e-10 * dv, False) - L_at(0.01, 0.07, v0 - 1e-10 * dv, False)) / 2e-10
    bad = []
    for what, ad, fd in checks:
        print(f"  {what:8s} adjoint={ad:+.10e} fd={fd:+.10e} rel={rel(ad, fd):.2e}", flush=True)
        if rel(ad, fd) > FD_TOL:
            bad.append(what)
    print(f"  v0 . dv free-running fd (step 1e-10) {free:+.10e} rel {rel(checks[2][1], free):.2e}", flush=True)
    assert not bad, bad


def test_position_adjoint_matches_fd_fp64():
    """d loss / d (initial positions) along a random direction: exercises
    every x-derivative (pass A, both solves, viscosity and normals)."""
    name, steps = "dambreak_3d_8k", 3
    r = runner_for(name, 8)
    st0 = r.checkpoint()
    v0 = seeded_velocity(r)
    target = forward_only(r, st0, v0, 0.02, 0.0, steps).detach()
    f = r.sim.fluid
    n, d = f.part_num, r.sim.config.dim
    # taped forward + manual reverse
    r.restore(st0)
    r.set_params(0.01, 0.0)
    r.set_velocity_ref(v0)
    r.step_index = 0
    r.iter_log = []
    tapes = [r.step(tape=True) for _ in range(steps)]
    base_log = list(r.iter_log)
    x = r.positions_ref().clone().requires_grad_(True)
    L = loss_of(x, target)
    (gx,) = torch.autograd.grad(L, x)
    xbar = torch.zeros(n, 4, dtype=f.dtype, device=v0.device)
    vbar = torch.zeros_like(xbar)
    xbar[:, :d] = gx
    r.grad.zero_()
    for tp in reversed(tapes):
        r.reverse(tp, xbar, vbar)
    g = torch.Generator().manual_seed(11)
    dx = torch.randn(n, d, generator=g, dtype=torch.float64).to(v0.device)
    ad = float((xbar[:, :d] * dx).sum())

    def L_shift(eps):
        r.restore(st0)
        pid_ = f.t["pid"][:n].long()
        f.t["pos"][:n, :d] += eps * dx[pid_]
        st = r.checkpoint()
        xx = forward_only(r, st, v0, 0.01, 0.0, steps, force=base_log)
        return float(loss_of(xx, target))

    e = 1e-11
    fd = (L_shift(e) - L_shift(-e)) / (2 * e)
    print(f"  x0 . dx adjoint={ad:+.10e} fd={fd:+.10e} rel={rel(ad, fd):.2e}", flush=True)
    assert rel(ad, fd) <= FD_TOL


def test_checkpoint_interval_does_not_change_gradient():
    """Replays are deterministic: every = 1 and every = steps, and tape
    budgets that keep none, some or all of the earlier segments' tapes, give
    bit-identical gradients; so do tapes with the step's fluid lists stored
    (the reverse rebuilds only the boundary lists) and without (it rebuilds
    all of them)."""
    name, steps = "dambreak_3d_8k", 6
    r, st0, v0, target, rollout = setup_case(name, steps, precision=4)
    r.restore(st0)
    r.tape_lists = False
    bare = r.step(tape=True)
    r.restore(st0)
    r.tape_lists = True
    full = r.step(tape=True)
    assert bare.nbr_f is None and full.nbr_f is not None
    assert full.nbytes() > bare.nbytes()
    one = full.nbytes()
    out = []
    for tl, every, budget in [(tl, e, b) for tl in (True, False) for e, b in
                              ((1, 0), (steps, 0), (2, 0), (2, int(2.5 * one)), (2, 1 << 40), (1, 1 << 40))]:
        r.tape_lists = tl
        r.restore(st0)
        mu = torch.tensor(0.01, dtype=torch.float32, device=v0.device, requires_grad=True)
        gam = torch.tensor(0.07, dtype=torch.float32, device=v0.device, requires_grad=True)
        v = v0.clone().requires_grad_(True)
        x, _ = rollout(r, v, mu, gam, steps, every, tape_budget=budget)
        loss_of(x, target).backward()
        out.append((float(mu.grad), float(gam.grad), v.grad.clone(), x.detach().clone()))
    for o in out[1:]:
        assert out[0][0] == o[0] and out[0][1] == o[1]
        assert torch.equal(out[0][2], o[2])
        assert torch.equal(out[0][3], o[3])


def test_taped_step_is_bit_identical_to_plain_step():
    r = runner_for("dambreak_3d_8k", 4)
    st0 = r.checkpoint()
    v0 = seeded_velocity(r)
    xa = forward_only(r, st0, v0, 0.01, 0.07, 3)
    r.restore(st0)
    r.set_velocity_ref(v0)
    for _ in range(3):
        r.step(tape=True)
    xb = r.positions_ref()
    assert torch.equal(xa, xb)
    # and the fused step of the Simulation itself
    r.restore(st0)
    r.set_velocity_ref(v0)
    for _ in range(3):
        r.sim.sph_step()
    assert torch.equal(xa, r.positions_ref())


def test_fp32_adjoint_tracks_fp64():
    """FP32 adjoint against FP64 on a well-conditioned loss (distance to the
    initial lattice shifted by 1 cm; a target from a nearby forward run would
    make d loss / d x a difference of nearly equal FP32 positions).  The FP32
    forward carries ~5e-4 relative velocity error from the psi - psi_0
    cancellation, which the pressure solve's sensitivity amplifies: measured
    1.5e-2 (positions) and 6e-3 (velocities) over 3 steps."""
    name, steps = "dambreak_3d_8k", 3
    grads = {}
    for prec in (8, 4):
        from tichi_sph_amd.adjoint import rollout
        r = runner_for(name, prec)
        st0 = r.checkpoint()
        v0 = seeded_velocity(r)
        target = (r.positions_ref().double() + 0.01).to(r.sim.fluid.dtype)
        dt_ = r.sim.fluid.dtype
        mu = torch.tensor(0.01, dtype=dt_, device=v0.device, requires_grad=True)
        gam = torch.tensor(0.07, dtype=dt_, device=v0.device, requires_grad=True)
        v = v0.clone().requires_grad_(True)
        x, _ = rollout(r, v, mu, gam, steps, 2)
        loss_of(x, target).backward()
        grads[prec] = (float(mu.grad), float(gam.grad), v.grad.double().cpu().numpy())
        print(f"  precision {prec}: dmu={grads[prec][0]:+.6e} dgamma={grads[prec][1]:+.6e} "
              f"iters={r.iter_log}", flush=True)
        del st0
    for k in (0, 1):
        print(f"  scalar {k}: rel {rel(grads[4][k], grads[8][k]):.2e}", flush=True)
        assert rel(grads[4][k], grads[8][k]) < 5e-2
    a, b = grads[4][2], grads[8][2]
    e = np.linalg.norm(a - b) / np.linalg.norm(b)
    print(f"  v0 gradient FP32 vs FP64: {e:.2e}", flush=True)
    assert e < 2e-2


@pytest.mark.parametrize("tape_lists", [True, False])
def test_plain_steps_between_taping_and_reverse_leave_the_tapes_intact(tape_lists):
    """Tapes keep what their reverse reads: plain Simulation steps run between
    the taped forward and its reverse (they write the Simulation's own list
    and field buffers, and with tape_lists False force a relist) leave the
    gradient bit-identical."""
    r = runner_for("dambreak_3d_8k", 4)
    r.tape_lists = tape_lists
    st0 = r.checkpoint()
    v0 = seeded_velocity(r)
    n, dev = r.sim.fluid.part_num, r.sim.device
    g = torch.Generator().manual_seed(5)
    wx = torch.randn(n, 4, generator=g).to(dev, r.sim.fluid.dtype)
    wv = torch.randn(n, 4, generator=g).to(dev, r.sim.fluid.dtype)
    wx[:, 3] = 0
    wv[:, 3] = 0
    out = []
    for plain in (0, 3):
        r.restore(st0)
        r.set_params(0.01, 0.07)
        r.set_velocity_ref(v0)
        r.step_index, r.iter_log, r.force = 0, [], None
        tapes = [r.step(tape=True) for _ in range(3)]
        for _ in range(plain):
            r.sim.sph_step()
        xbar, vbar = wx.clone(), wv.clone()
        r.grad.zero_()
        for tp in reversed(tapes):
            r.reverse(tp, xbar, vbar)
        out.append((xbar.cpu(), vbar.cpu(), r.grad.cpu()))
        print(f"  plain steps {plain}: |xbar| {float(xbar.abs().max()):.4e} "
              f"dmu {float(r.grad[0]):+.6e}", flush=True)
    for a, b in zip(out[0], out[1]):
        assert torch.equal(a, b)
