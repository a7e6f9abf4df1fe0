"""GPU parity at the edges of the grid and of the particle set.

- Particles outside the domain, in cell 0 and far away: the reference keeps
  a particle out of the grid when its cell code is not in (0, node_num)
  (sph_obj.py:305, 319) but still lets it walk its own neighbourhood, whose
  codes are formed with wrapping row arithmetic (sph.py:55-57); a particle
  just below the lower corner therefore lands in the previous row's last
  cells.  Two particles at the same position contribute W(0) to each
  other's density and are excluded from every r > 0 term.
- No fluid particles: the reference still updates the boundary, and each
  solve's mean compression is 0/0 = NaN (sph.py:221-226), so the solve runs
  its minimum iteration count.
Both are compared with the oracle on the same inputs, to the bar of the
other parity tests.  The adjoint over the stray particles, with surface
tension on, is compared with central differences of the recorded step.
"""
import copy

import numpy as np
import pytest

import oracle as O
from conftest import load_scene
from test_gpu_parity import FIELDS_F, TOL, compare_fields, make, rel_err

pytestmark = pytest.mark.gpu


def _stray_positions(sim, pos, dtype=np.float32):
    """Moves six interior particles (reference order) to the edge cases."""
    cfg = sim.config
    d = cfg.dim
    h = float(sim.params.kernel_h[1])
    lb = np.array(list(sim.params.lb)[:d])
    rt = np.array(list(sim.params.rt)[:d])
    p = pos.astype(np.float64).copy()
    n = len(p)
    idx = [n // 7, 2 * n // 7, 3 * n // 7, 4 * n // 7, 5 * n // 7, 5 * n // 7 + 1]
    p[idx[0], 0] = lb[0] - 1.5 * h           # below the lower corner: wraps into the previous row
    p[idx[1], d - 1] = rt[d - 1] + 2.5 * h   # beyond the upper corner: code out of range
    p[idx[2]] = lb + 0.5 * h                 # cell 0: never inserted
    p[idx[3]] = 1.0e9                        # cell index beyond int32
    p[idx[5]] = p[idx[4]]                    # two particles at one position
    return p.astype(dtype), idx


@pytest.mark.parametrize("name", ["dambreak_3d_8k", "dambreak_2d_8k"])
def test_stray_particles_fused_step_matches_oracle(name):
    sim, o = make(name)
    p, idx = _stray_positions(sim, o.get(0, "pos"))
    print(f"  moved particles {idx}", flush=True)
    o.set(0, "pos", p)
    sim.fluid.from_numpy("pos", p)
    assert np.array_equal(sim.fluid.to_numpy("pos"), o.get(0, "pos"))
    s = sim.sph_step(stats=True)
    st = o.step(O.STEP_CFL)
    print(f"  iters gpu=({s.div_iter},{s.incom_iter}) oracle=({st.div_iter},{st.incom_iter})", flush=True)
    assert (s.div_iter, s.incom_iter) == (st.div_iter, st.incom_iter)
    assert s.dt == st.dt
    # The holes the moved particles leave make the solve's FP32 rounding
    # visible: a field more than TOL away from the FP32 oracle must be as
    # close to the FP64 oracle as the FP32 oracle itself is.
    cb, sb = load_scene(name)
    o64 = O.Oracle(cb, sb, 8)
    o64.init_scenario()
    o64.set(0, "pos", p)
    o64.step(O.STEP_CFL)
    for f in FIELDS_F + ["alpha_1", "alpha_2", "normal"]:
        g, r, r64 = sim.fluid.to_numpy(f), o.get(0, f), o64.get(0, f)
        e32, eg, eo = rel_err(g, r), rel_err(g, r64), rel_err(r, r64)
        worst = int(np.argmax(np.abs(np.asarray(g, np.float64) - r).reshape(len(g), -1).max(1)))
        print(f"  {f:16s} vs oracle {e32:.2e} (worst particle {worst}{' moved' if worst in idx else ''}); "
              f"vs FP64: gpu {eg:.2e} oracle {eo:.2e}", flush=True)
        assert e32 <= TOL or eg <= 2 * eo + TOL / 10, f
    compare_fields(sim, o, ["sph_density", "alpha", "W"], obj=1, label="stray")
    # pass A's strict sums are bit-exact for the moved particles too
    for f in ("sph_density", "sph_compression", "W"):
        assert np.array_equal(sim.fluid.to_numpy(f)[idx], o.get(0, f)[idx]), f
    # the particles outside the grid see no neighbours, not even themselves
    dens = sim.fluid.to_numpy("sph_density")
    assert dens[idx[1]] == 0 and dens[idx[2]] == 0 and dens[idx[3]] == 0
    # a second step sorts the moved state again
    s = sim.sph_step(stats=True)
    st = o.step(O.STEP_CFL)
    assert (s.div_iter, s.incom_iter) == (st.div_iter, st.incom_iter)
    g, r = sim.fluid.to_numpy("pos")[idx], o.get(0, "pos")[idx]
    assert np.allclose(g, r, rtol=1e-6, atol=1e-9)


def test_stray_particles_reference_surface_step_matches_oracle():
    sim, o = make("dambreak_3d_8k")
    p, _ = _stray_positions(sim, o.get(0, "pos"))
    o.set(0, "pos", p)
    sim.fluid.from_numpy("pos", p)
    di, ii, _, _, _ = sim.sph_step_reference()
    st = o.step(O.STEP_CFL)
    assert (di, ii) == (st.div_iter, st.incom_iter)
    compare_fields(sim, o, FIELDS_F, label="stray ref-surface")


def test_empty_fluid_step_matches_oracle():
    from tichi_sph_amd import simulation_from_dicts
    cb, sb = load_scene("dambreak_3d_8k")
    sb = copy.deepcopy(sb)
    sb["fluid"]["objs"] = []
    print("[case] dambreak_3d_8k without fluid", flush=True)
    sim = simulation_from_dicts(cb, sb, precision=4)
    o = O.Oracle(cb, sb, 4)
    o.init_scenario()
    assert sim.fluid.part_num == 0 and o.n[0] == 0
    for step in range(2):
        s = sim.sph_step(stats=True)
        st = o.step(O.STEP_CFL)
        print(f"  step {step}: iters gpu=({s.div_iter},{s.incom_iter}) "
              f"oracle=({st.div_iter},{st.incom_iter}) dt {s.dt} {st.dt}", flush=True)
        assert (s.div_iter, s.incom_iter) == (st.div_iter, st.incom_iter)
        assert s.div_iter == sim.config.iter_threshold_min
        assert s.dt == st.dt
        assert all(np.isnan(e) for e in s.div_err) and all(np.isnan(e) for e in s.incom_err)
        assert s.kinetic_energy == 0
    worst = compare_fields(sim, o, ["sph_density", "sph_compression", "alpha", "alpha_2", "W"], obj=1,
                           label="no fluid")
    assert max(worst.values()) == 0  # the boundary's sums are strict: bit-exact


def test_adjoint_over_stray_particles_matches_fd_fp64():
    """The reverse over the same stray particles with surface tension on: the
    particles outside the grid have no neighbours and zero density (their
    V = X / psi is infinite but enters no pair term), and two share a
    position.  The gradients are finite and equal central differences."""
    import torch
    from tichi_sph_amd.adjoint import rollout
    from test_gpu_adjoint import FD_TOL, forward_only, loss_of, rel, runner_for, seeded_velocity
    steps, every = 3, 3
    r = runner_for("dambreak_3d_8k", 8)
    f = r.sim.fluid
    p, idx = _stray_positions(r.sim, f.to_numpy("pos"), np.float64)
    print(f"  moved particles {idx}; {steps} steps, gamma = 0.07", flush=True)
    f.from_numpy("pos", p)
    st0 = r.checkpoint()
    v0 = seeded_velocity(r)
    target = forward_only(r, st0, v0, 0.02, 0.05, steps).detach()
    r.restore(st0)
    dev = v0.device
    mu = torch.tensor(0.01, dtype=torch.float64, device=dev, requires_grad=True)
    gam = torch.tensor(0.07, dtype=torch.float64, device=dev, requires_grad=True)
    v = v0.clone().requires_grad_(True)
    x, _ = rollout(r, v, mu, gam, steps, every)
    loss_of(x, target).backward()
    log = list(r.iter_log)
    print(f"  iterations {log}", flush=True)
    assert bool(torch.isfinite(v.grad).all()) and np.isfinite(float(mu.grad)) and np.isfinite(float(gam.grad))

    def L_at(mu_, gam_, v_):
        return float(loss_of(forward_only(r, st0, v_, mu_, gam_, steps, force=log), target))

    g = torch.Generator().manual_seed(7)
    dv = torch.randn(v0.shape, generator=g, dtype=torch.float64).to(dev)
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
        if not rel(ad, fd) <= FD_TOL:
            bad.append(what)
    assert not bad, bad
