"""Pins the dense FP64 torch restatement of the step (tests/dense_step.py),
which the adjoint's gradients are checked against by autograd: over three
steps of the tiny 2D and 3D dam breaks (random initial velocities, mu 0.01,
surface tension 0.07) it must track the FP64 oracle to round-off, with the
oracle's own iteration counts.  It must also differ from the oracle when the
viscosity or the surface tension is dropped, so both terms are exercised."""
import numpy as np
import pytest
import torch

import oracle as O
from dense_step import DenseDFSPH, tiny_dambreak


def f32(a):
    return float(np.float32(a))


def t64(a):
    return torch.tensor(np.asarray(a, dtype=np.float64))


def dense_for_oracle(o, sb, dim, dt):
    env = sb["sim_env"]
    return DenseDFSPH(t64(o.get(0, "mass")), t64(o.get(0, "rest_density")), t64(o.get(0, "rest_volume")),
                      t64(o.get(1, "pos")), t64(o.get(1, "mass")), o.c.kernel_h[1], o.c.kernel_h[2],
                      o.c.kernel_sig[dim], t64([f32(g) for g in env["gravity"]]), dt,
                      t64([f32(v) for v in env["sim_space_lb"]]))


def run(dim, mu, gamma, steps=3):
    cb, sb = tiny_dambreak(dim)
    o = O.Oracle(cb, sb, 8)
    o.init_scenario()
    n = o.n[0]
    v0 = np.random.default_rng(dim).normal(0, 0.01, (n, dim))
    o.set(0, "vel", v0)
    dt = f32(0.02 / 50 if dim == 2 else 0.01 / 100)
    o.scalar("dt", 0, dt)
    D = dense_for_oracle(o, sb, dim, dt)
    x, v = t64(o.get(0, "pos")), t64(v0)
    errs = []
    for _ in range(steps):
        st = o.step(0)
        assert st.div_iter >= 2 and st.incom_iter >= 2
        x, v = D.step(x, v, mu, gamma, st.div_iter, st.incom_iter)
        xo, vo = o.get(0, "pos"), o.get(0, "vel")
        errs.append((np.abs(x.numpy() - xo).max() / np.abs(xo).max(),
                     np.abs(v.numpy() - vo).max() / np.abs(vo).max()))
    return errs


@pytest.mark.parametrize("dim", [2, 3])
def test_dense_step_tracks_fp64_oracle(dim):
    errs = run(dim, f32(0.01), f32(0.07))
    print(f"  dense vs FP64 oracle, {dim}D, (pos, vel) per step: {errs}", flush=True)
    for ex, ev in errs:
        assert ex <= 1e-12 and ev <= 1e-10


@pytest.mark.parametrize("dim", [2, 3])
def test_dense_step_sees_viscosity_and_surface_tension(dim):
    for mu, gamma in ((0.0, f32(0.07)), (f32(0.01), 0.0)):
        ev = max(e[1] for e in run(dim, mu, gamma, steps=1))
        assert ev > 1e-6, (mu, gamma, ev)
