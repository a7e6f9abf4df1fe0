"""GPU parity of paths the plain dam-break steps leave unexercised:

* the velocity branch of cfl_condition (sph.py:41-47): fluid moving fast
  enough that gps / |v| * cfl_factor < gps / cs, on the fused path (the CFL
  reduction folded into the sort), on the reference surface and in a batch
  (SceneBatch(cfl=True)), each against the oracle and the closed form;
* a boundary grown after a step through Fluid.scene_add_cube
  (sph_obj.py:110-114, 165-180): the next step must search the new boundary
  particles, as the reference's per-step Ngrid does (main.py:178-183);
* the FBM round loop (sph.py:301-379, main.py:244-250) on a constructed
  three-particle state in which one particle's fraction would go negative;
* the values the output writers take from a step (main.py:151-171):
  energies (sph.py:272-279) and dt after a direct sph_step().
"""
import json

import numpy as np
import pytest
import torch

import oracle as O
from conftest import load_scene
from test_gpu_parity import FIELDS_F, TOL, compare_fields, rel_err

pytestmark = pytest.mark.gpu


def cfl_dicts(name, cs=2):
    """The scene with a small solver_cs, so that moderate speeds (about
    1.5 m/s) already limit dt through the velocity branch."""
    cb, sb = load_scene(name)
    return dict(cb, solver_cs=cs), sb


def fast_velocities(n, d, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(0, 0.05, (n, d))
    v[:, 0] += 1.5
    return v.astype(np.float32)


def cfl_closed_form(v, gps, cs, cfl):
    """min(gps / cs, gps / |v_i| * cfl over |v_i| > 1e-4) in f32 arithmetic,
    |v| = sqrt of the sum of squares in component order."""
    f = np.float32
    s = v[:, 0] * v[:, 0]
    for k in range(1, v.shape[1]):
        s = (s + v[:, k] * v[:, k]).astype(f)
    vn = np.sqrt(s).astype(f)
    dt = f(gps) / f(cs)
    cand = (f(gps) / vn[vn > f(1e-4)] * f(cfl)).astype(f)
    return dt, min(dt, cand.min())


@pytest.mark.parametrize("path", ["fused", "reference"])
@pytest.mark.parametrize("name", ["dambreak_2d_8k", "dambreak_3d_8k"])
def test_cfl_velocity_branch_step_matches_oracle(name, path):
    from tichi_sph_amd import simulation_from_dicts
    cb, sb = cfl_dicts(name)
    print(f"[case] CFL-limited {name} {path}", flush=True)
    sim = simulation_from_dicts(cb, sb)
    o = O.Oracle(cb, sb, 4)
    o.init_scenario()
    n, d = sim.fluid.part_num, sim.config.dim
    assert n > 8 * 256  # the CFL partials span many blocks of the key kernel
    v = fast_velocities(n, d, seed=11)
    sim.fluid.from_numpy("vel", v)
    o.set(0, "vel", v)
    dt_cs, expect = cfl_closed_form(v, o.c.part_size[1], cb["solver_cs"], np.float32(cb["solver_cfl_factor"]))
    assert expect < dt_cs  # the velocity branch decides dt
    if path == "fused":
        s = sim.sph_step(stats=True)
        iters, dt = (s.div_iter, s.incom_iter), s.dt
    else:
        di, ii, _, _, _ = sim.sph_step_reference()
        iters, dt = (di, ii), sim.control().dt
    st = o.step(O.STEP_CFL)
    print(f"  dt gpu {dt!r} oracle {st.dt!r} closed form {float(expect)!r} (gps/cs {float(dt_cs)!r}); "
          f"iterations gpu {iters} oracle {(st.div_iter, st.incom_iter)}", flush=True)
    assert dt == st.dt == float(expect)
    assert iters == (st.div_iter, st.incom_iter)
    # A dt 15x the plain scenes' makes the solver's FP32 cancellation in
    # psi = (rho - rho0) + dt * sum visible: a field more than TOL away from
    # the FP32 oracle must be as close to the FP64 oracle as the FP32 oracle
    # itself is.
    o64 = O.Oracle(cb, sb, 8)
    o64.init_scenario()
    o64.set(0, "vel", v)
    st64 = o64.step(O.STEP_CFL)
    assert (st64.div_iter, st64.incom_iter) == iters
    for f in FIELDS_F:
        g, r, r64 = sim.fluid.to_numpy(f), o.get(0, f), o64.get(0, f)
        e32, eg, eo = rel_err(g, r), rel_err(g, r64), rel_err(r, r64)
        print(f"  cfl-{path} {f:16s} vs oracle {e32:.2e}; vs FP64: gpu {eg:.2e} oracle {eo:.2e}", flush=True)
        assert e32 <= TOL or eg <= 2 * eo + TOL / 10, f


def test_batched_cfl_matches_single_scenes_and_closed_form():
    """SceneBatch(cfl=True): every scene's dt comes from its own CFL
    reduction (kb_* sort + finalize), bit-identical to the scene's own
    sph_step() and to the closed form, and the states stay bit-identical."""
    from tichi_sph_amd import simulation_from_dicts
    from tichi_sph_amd.batch import SceneBatch
    names = ["dambreak_3d_8k", "dambreak_2d_8k", "dambreak_3d_8k"]
    print(f"[case] batched CFL {names}", flush=True)

    def build():
        sims = []
        for s, name in enumerate(names):
            cb, sb = cfl_dicts(name)
            sim = simulation_from_dicts(cb, sb)
            v = fast_velocities(sim.fluid.part_num, sim.config.dim, seed=20 + s)
            v[:, 0] *= 1.0 + 0.2 * s  # a different dt per scene
            sim.fluid.from_numpy("vel", v)
            sims.append((sim, v, cb))
        return sims

    ref = build()
    batch_sims = build()
    batch = SceneBatch([s for s, _, _ in batch_sims], cfl=True)
    for step in range(2):
        di, ii = batch.step()
        for s, (sim, v, cb) in enumerate(ref):
            if step == 0:
                _, expect = cfl_closed_form(v, sim.config.part_size[1], cb["solver_cs"],
                                            np.float32(cb["solver_cfl_factor"]))
            st = sim.sph_step(stats=True)
            bdt = batch.sims[s].control().dt
            assert (int(di[s]), int(ii[s])) == (st.div_iter, st.incom_iter), (step, s)
            assert bdt == st.dt, (step, s, bdt, st.dt)
            if step == 0:
                assert bdt == float(expect), (s, bdt, float(expect))
            for f in ("pos", "vel"):
                assert np.array_equal(batch.sims[s].fluid.to_numpy(f), sim.fluid.to_numpy(f)), (step, s, f)
    dts = [b.control().dt for b in batch.sims]
    print(f"  per-scene dt {dts}", flush=True)
    assert len(set(dts)) == len(dts)


def test_boundary_grown_after_a_step_is_searched():
    """One step, then a plate of 100 boundary particles added through
    Fluid.scene_add_cube just above the fluid's free surface (within h of the
    top layer); the next fused step must match the oracle, which searches the
    new boundary particles (the reference rebuilds its Ngrid every step)."""
    from tichi_sph_amd import simulation_from_dicts
    cb, sb = load_scene("dambreak_3d_8k")
    print("[case] dambreak_3d_8k, boundary plate added after step 1", flush=True)
    nb0 = 9344
    start, end = [0.05, 0.2, 0.05], [0.15, 0.21, 0.15]
    sim = simulation_from_dicts(cb, sb, bound_max=nb0 + 100)
    oracles = [O.Oracle(cb, sb, p, bound_max=nb0 + 100) for p in (4, 8)]
    for o in oracles:
        o.init_scenario()
        assert o.n[1] == nb0
    assert sim.bound.part_num == nb0
    s = sim.sph_step(stats=True)
    sts = [o.step(O.STEP_CFL) for o in oracles]
    assert (s.div_iter, s.incom_iter) == (sts[0].div_iter, sts[0].incom_iter)
    sim.bound.scene_add_cube(start, end, [1.0], [0.0, 0.0, 0.0], 0xAAAAAA, 1.0)
    assert not sim.bound.cell_table_ok and sim.bound_needs_sort()
    o = oracles[0]
    x = O.lattice(3, o.c.part_size[1], 1.0, start, end, 0).astype(np.float32)
    assert len(x) == 100
    nb = nb0 + len(x)
    assert sim.bound.part_num == nb
    assert np.array_equal(sim.bound.to_numpy("pos")[nb0:], x)
    top = sim.fluid.to_numpy("pos")[:, 1].max()
    assert 0 < x[:, 1].min() - top < o.c.kernel_h[1]  # within h of the top layer
    vol = np.float32(o.c.part_size[3])
    add = {"pos": x, "vel": np.zeros((len(x), 3)), "volume_frac": np.ones(len(x)),
           "rest_volume": np.full(len(x), vol), "rest_density": np.full(len(x), 1000.0),
           "mass": np.full(len(x), np.float32(np.float32(1000.0) * vol))}
    for oo in oracles:
        old = {f: oo.get(1, f) for f in add}
        oo.set_part_num(1, nb)
        for f, a in add.items():
            oo.set(1, f, np.concatenate([old[f].reshape(nb0, -1), np.asarray(a).reshape(len(x), -1)]))
    for f in ("pos", "rest_volume", "rest_density", "mass"):
        assert np.array_equal(sim.bound.to_numpy(f).reshape(nb, -1), o.get(1, f).reshape(nb, -1)), f
    s = sim.sph_step(stats=True)
    sts = [oo.step(O.STEP_CFL) for oo in oracles]
    assert not sim.bound_needs_sort()
    assert (s.div_iter, s.incom_iter) == (sts[0].div_iter, sts[0].incom_iter)
    # W and densities are strict sums in walk order: bit-exact next to the
    # plate too.  The other fields: within TOL of the FP32 oracle, or as close
    # to the FP64 oracle as the FP32 oracle is (the solver's residual psi_adv
    # carries the FP32 cancellation of psi = rho - rho0 + dt * sum).
    for f in ("W", "sph_density", "sph_compression"):
        assert np.array_equal(sim.fluid.to_numpy(f), o.get(0, f)), f
    for f in FIELDS_F + ["alpha_1"]:
        g, r, r64 = sim.fluid.to_numpy(f), o.get(0, f), oracles[1].get(0, f)
        e32, eg, eo = rel_err(g, r), rel_err(g, r64), rel_err(r, r64)
        print(f"  grown boundary {f:16s} vs oracle {e32:.2e}; vs FP64: gpu {eg:.2e} oracle {eo:.2e}", flush=True)
        assert e32 <= TOL or eg <= 2 * eo + TOL / 10, f
    compare_fields(sim, o, ["sph_density", "alpha", "W"], obj=1, label="grown boundary")
    # the plate pushes on the top layer: the fluid below it is denser than
    # the same layer elsewhere
    dens = sim.fluid.to_numpy("sph_density")
    p = sim.fluid.to_numpy("pos")
    under = (p[:, 1] > top - 0.005) & (np.abs(p[:, 0] - 0.1) < 0.04) & (np.abs(p[:, 2] - 0.1) < 0.04)
    away = (p[:, 1] > top - 0.005) & (p[:, 0] > 0.17)
    assert dens[under].mean() > dens[away].mean()


def three_particle_fbm_state(sim_or_none=None):
    """Three fluid particles in a row (h = 0.016): particle 0 (phi = [0.02,
    0.98], volume 1e-7) next to particle 1 (phi = [1, 0], volume 1e-5) at
    q = 0.6, particle 2 (phi = [0.5, 0.5]) at q = 0.9 from particle 1 and
    beyond h from particle 0; one boundary particle far away.  With a
    diffusion coefficient of 1, particle 0's phase-1 fraction would go
    negative in the first FBM round."""
    base = np.array([0.05, 0.05, 0.05])
    pos = np.stack([base, base + [0.0096, 0, 0], base + [0.024, 0, 0]]).astype(np.float32)
    phi = np.array([[0.02, 0.98], [1.0, 0.0], [0.5, 0.5]], dtype=np.float32)
    V = np.array([1e-7, 1e-5, 1e-5], dtype=np.float32)
    return pos, phi, V


@pytest.mark.parametrize("path", ["fused", "reference"])
def test_fbm_round_loop_on_constructed_state(path):
    from tichi_sph_amd import simulation_from_dicts
    cb, sb = load_scene("two_phase_3d_4k")
    cb = dict(cb, solver_fbm_diffusion_term=1.0, solver_fbm_convection_term=0.0,
              solver_surface_tension_gamma=0.0)
    sb = json.loads(json.dumps(sb))
    sb["fluid"]["objs"] = []
    sb["bound"]["objs"] = []
    print(f"[case] three-particle FBM round loop, {path}", flush=True)
    sim = simulation_from_dicts(cb, sb, fbm=True, cfl=False, fluid_max=3, bound_max=1)
    pos, phi, V = three_particle_fbm_state()
    sim.fluid.push_part_seq(3, pos, [1.0, 0.0], [0.0, 0.0, 0.0])
    sim.bound.push_part_seq(1, pos[:1] + np.float32(0.1), [1.0, 0.0], [0.0, 0.0, 0.0])
    sim.fluid.from_numpy("volume_frac", phi)
    sim.fluid.from_numpy("rest_volume", V)
    sim.fluid.update_mass_all()
    sim.set_dt(1e-4)
    o = O.Oracle(cb, sb, 4, fluid_max=3, bound_max=1)
    o.set_part_num(0, 3)
    o.set_part_num(1, 1)
    for obj, fl in ((0, sim.fluid), (1, sim.bound)):
        for f in ("pos", "vel", "volume_frac", "rest_volume", "rest_density", "mass"):
            o.set(obj, f, fl.to_numpy(f))
    o.scalar("dt", 0, float(np.float32(1e-4)))
    if path == "fused":
        s = sim.sph_step(stats=True)
        rounds = s.fbm_rounds
    else:
        _, _, _, _, rounds = sim.sph_step_reference()
    st = o.step(O.STEP_FBM)
    flags = sim.fluid.to_numpy("flag")
    got = sim.fluid.to_numpy("volume_frac")
    print(f"  rounds gpu {rounds} oracle {st.fbm_rounds}; flags {flags.tolist()}; fractions {got.tolist()}",
          flush=True)
    assert rounds == st.fbm_rounds == 2
    assert flags.tolist() == o.get(0, "flag").tolist() == [1, 0, 0]
    assert np.array_equal(got[0], phi[0])
    assert rel_err(got, o.get(0, "volume_frac")) <= 1e-5
    assert abs(got[1, 1]) > 1e-3


def test_written_json_energies_and_dt_after_sph_step(tmp_path):
    """write_full_json / write_json after direct sph_step() calls write the
    energies of SPH_update_energy and the dt of cfl_condition of the last
    step (main.py:151-171), checked against the oracle."""
    from tichi_sph_amd import simulation_from_dicts
    from tichi_sph_amd.output import write_full_json, write_json
    cb, sb = load_scene("dambreak_3d_8k")
    print("[case] output JSON after 3 steps", flush=True)
    sim = simulation_from_dicts(cb, sb)
    o = O.Oracle(cb, sb, 4)
    o.init_scenario()
    for _ in range(3):
        sim.sph_step()
        st = o.step(O.STEP_CFL)
    torch.cuda.synchronize()
    data = write_full_json(sim, str(tmp_path / "frame.json"))
    back = json.load(open(tmp_path / "frame.json"))
    assert back == json.loads(json.dumps(data))
    e = back["energy"]
    ke, pe = o.scalar("ke"), o.scalar("pe")
    print(f"  written ke {e['statistics_kinetic_energy']!r} pe {e['statistics_gravity_potential_energy']!r} "
          f"dt {back['timeStep']!r}; oracle ke {ke!r} pe {pe!r} dt {st.dt!r}", flush=True)
    assert back["timeStep"] == st.dt
    assert ke > 0 and pe > 0
    assert e["statistics_kinetic_energy"] == pytest.approx(ke, rel=1e-4)
    assert e["statistics_gravity_potential_energy"] == pytest.approx(pe, rel=1e-6)
    assert e["sum_energy"] == pytest.approx(ke + pe, rel=1e-6)
    d2 = write_json(sim, str(tmp_path / "json"))
    assert d2["energy"] == back["energy"] and d2["timeStep"] == back["timeStep"]
