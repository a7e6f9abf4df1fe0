"""The adjoint's whole gradient against torch autograd through the dense FP64
restatement of the step (tests/dense_step.py, pinned to the FP64 oracle by
tests/test_dense_step_cpu.py), on the tiny 2D and 3D dam breaks over three
steps with the recorded iteration counts.

Finite differences (tests/test_gpu_adjoint.py) check one direction at a
time; here every component of d loss / d (x0, v0) and both parameter
derivatives are compared.  The loss is a random linear functional of the
final positions and velocities, so its gradient is the vector-Jacobian
product of the three-step map itself.

The coincident case puts two fluid particles at one position.  Every pair
term of the reference requires r > 0, so that pair contributes nothing; the
surface-tension sums test r > 0 through j != i and check positions only in
the particle's own cell, and the normal-adjoint term of such a pair would not
vanish if they did not.  Autograd through the dense step, which masks r = 0
pairs, pins that handling."""
import numpy as np
import pytest
import torch

from dense_step import DenseDFSPH, tiny_dambreak

pytestmark = pytest.mark.gpu

STEPS = 3
MU, GAMMA = 0.01, 0.07


def rel_vec(a, b):
    return float((a - b).abs().max() / b.abs().max())


@pytest.mark.parametrize("coincident", [False, True], ids=["plain", "coincident"])
@pytest.mark.parametrize("dim", [2, 3])
def test_adjoint_gradient_equals_autograd_of_dense_step(dim, coincident):
    from tichi_sph_amd import simulation_from_dicts
    from tichi_sph_amd.adjoint import DiffRunner
    cb, sb = tiny_dambreak(dim, gamma=GAMMA)
    print(f"[case] tiny {dim}D dam break{' with a coincident pair' if coincident else ''}, {STEPS} steps, "
          f"adjoint vs dense autograd", flush=True)
    sim = simulation_from_dicts(cb, sb, precision=8, cfl=False)
    r = DiffRunner(sim)
    f = sim.fluid
    n, nb = f.part_num, sim.bound.part_num
    dev = sim.device
    if coincident:
        # the particle after the block's middle one moves onto it (the next
        # step's sort puts both in one cell)
        pid = f.t["pid"][:n].long()
        ia, ib = (int((pid == q).nonzero()[0]) for q in (n // 2, n // 2 + 1))
        f.t["pos"][ib, :dim] = f.t["pos"][ia, :dim]
    st0 = r.checkpoint()
    x0 = r.positions_ref().detach().cpu()
    if coincident:
        assert torch.equal(x0[n // 2], x0[n // 2 + 1])
    # the state's constants at the start (the tail's SPH_update_mass rewrites the masses)
    state = {k: np.asarray(f.to_numpy(k), dtype=np.float64) for k in ("mass", "rest_density", "rest_volume")}
    bound = {k: np.asarray(sim.bound.to_numpy(k), dtype=np.float64) for k in ("pos", "mass")}
    g = torch.Generator().manual_seed(dim)
    v0 = 0.01 * torch.randn(n, dim, generator=g, dtype=torch.float64)
    wx = torch.randn(n, dim, generator=g, dtype=torch.float64)
    wv = torch.randn(n, dim, generator=g, dtype=torch.float64)

    # the adjoint: taped forward, then the reverse of every step
    r.restore(st0)
    r.set_params(MU, GAMMA)
    r.set_velocity_ref(v0.to(dev))
    r.step_index, r.iter_log, r.force = 0, [], None
    tapes = [r.step(tape=True) for _ in range(STEPS)]
    log = list(r.iter_log)
    xT, vT = r.positions_ref().cpu(), r.velocities_ref().cpu()
    dt = sim.control().dt
    xbar = torch.zeros(n, 4, dtype=torch.float64, device=dev)
    vbar = torch.zeros_like(xbar)
    xbar[:, :dim], vbar[:, :dim] = wx.to(dev), wv.to(dev)
    r.grad.zero_()
    for tp in reversed(tapes):
        r.reverse(tp, xbar, vbar)
    ad = {"x0": xbar[:, :dim].cpu(), "v0": vbar[:, :dim].cpu(), "mu": r.grad[0].cpu(), "gamma": r.grad[1].cpu()}
    print(f"  {n} fluid, {nb} boundary particles; iterations {log}; dt {dt!r}", flush=True)

    # the dense restatement on the same state, with the same iteration counts
    p = sim.params
    t64 = lambda a: torch.tensor(np.asarray(a, dtype=np.float64))  # noqa: E731
    D = DenseDFSPH(t64(state["mass"]), t64(state["rest_density"]), t64(state["rest_volume"]),
                   t64(bound["pos"]), t64(bound["mass"]), p.kernel_h[1], p.kernel_h[2],
                   p.kernel_sig[dim], t64(list(p.gravity)[:dim]), dt, t64(list(p.lb)[:dim]))
    xs = x0.clone().requires_grad_(True)
    vs = v0.clone().requires_grad_(True)
    mu = torch.tensor(MU, dtype=torch.float64, requires_grad=True)
    gam = torch.tensor(GAMMA, dtype=torch.float64, requires_grad=True)
    x, v = xs, vs
    for d_it, i_it in log:
        x, v = D.step(x, v, mu, gam, d_it, i_it)
    ex, ev = rel_vec(xT, x.detach()), rel_vec(vT, v.detach())
    print(f"  forward: positions {ex:.2e}, velocities {ev:.2e} of the largest", flush=True)
    assert ex <= 1e-12 and ev <= 1e-9
    L = (wx * x).sum() + (wv * v).sum()
    gx, gv, gmu, ggam = torch.autograd.grad(L, [xs, vs, mu, gam])
    ref = {"x0": gx, "v0": gv, "mu": gmu, "gamma": ggam}
    worst = 0.0
    for k in ("x0", "v0", "mu", "gamma"):
        e = rel_vec(ad[k].reshape(-1), ref[k].reshape(-1))
        worst = max(worst, e)
        print(f"  d loss / d {k:5s}: adjoint vs autograd {e:.2e} of the largest component "
              f"(largest {float(ref[k].abs().max()):.3e})", flush=True)
    assert worst <= 1e-8
