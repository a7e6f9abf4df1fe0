"""GPU parity of the widened path: map_velocity (sph.py:382-397), the fused
WCSPH block (main.py:236-240) and a scenario that grows mid-run (timed
objects and inlets, main.py:42-112), each against the oracle on the same
inputs with the 1e-5 bar of the parity tests."""
import json

import numpy as np
import pytest

import oracle as O
from conftest import load_scene
from test_gpu_parity import compare_fields, rel_err, FIELDS_F, TOL

pytestmark = pytest.mark.gpu


def sim_and_oracle(name, **kw):
    from tichi_sph_amd import simulation_from_dicts
    cb, sb = load_scene(name)
    print(f"[case] {name} {kw}", flush=True)
    sim = simulation_from_dicts(cb, sb, **kw)
    o = O.Oracle(cb, sb, 4)
    o.init_scenario()
    return sim, o


@pytest.mark.parametrize("name", ["dambreak_2d_8k", "dambreak_3d_8k"])
def test_map_velocity_matches_oracle(name):
    from tichi_sph_amd import sph
    from tichi_sph_amd.grid import Grid
    sim, o = sim_and_oracle(name)
    sim.sph_step()
    o.step(O.STEP_CFL)
    g = Grid(sim)
    sph.map_velocity(sim, g, sim.fluid)
    ref = o.map_velocity(0, list(g.shape), g.dist)
    got = g.vel_numpy().reshape(-1, sim.config.dim)
    e = rel_err(got, ref)
    print(f"  grid {g.shape} nodes={g.size} rel_err={e:.2e}", flush=True)
    assert np.abs(ref).max() > 0
    assert e <= TOL


@pytest.mark.parametrize("name", ["dambreak_2d_8k", "dambreak_3d_8k"])
def test_wcsph_block_matches_oracle(name):
    """One step (the fused path's per-step differences, ~1e-6, compound
    through the stiff (rho/rho0)^7 equation of state over later steps)."""
    sim, o = sim_and_oracle(name, wcsph=True)
    s = sim.sph_step(stats=True)
    st = o.step(O.STEP_CFL | O.STEP_WCSPH)
    assert (s.div_iter, s.incom_iter) == (st.div_iter, st.incom_iter)
    if name == "dambreak_2d_8k":  # the 3D lattice starts below rest density: p clamps to 0
        assert np.abs(o.get(0, "pressure")).max() > 0
    compare_fields(sim, o, ["pos", "vel", "vel_adv", "acce_adv", "pressure", "sph_density"], label="wcsph")


def test_wcsph_reference_surface_matches_oracle():
    sim, o = sim_and_oracle("dambreak_2d_8k", wcsph=True)
    sim.sph_step_reference()
    o.step(O.STEP_CFL | O.STEP_WCSPH)
    compare_fields(sim, o, ["pos", "vel", "vel_adv", "acce_adv", "pressure"], label="wcsph-ref")


def test_growing_scenario_then_step_matches_oracle():
    """A timed cube and an inlet add particles during run_step frames; the
    grown state, handed to the oracle, then steps identically."""
    from tichi_sph_amd import simulation_from_dicts
    cb, sb = load_scene("dambreak_3d_8k")
    sb = json.loads(json.dumps(sb))
    sb["fluid"]["max_part_num"] = 20000
    base = json.loads(json.dumps(sb))
    sb["fluid"]["objs"].append({"type": "cube", "start_pos": [0.25, 0.02, 0.05], "end_pos": [0.3098, 0.0798, 0.1098],
                                "vel": [0, 0, 0], "volume_frac": [1], "color": "0xff0000",
                                "particle_relaxing_factor": 1.0, "time": 0.002})
    sb["fluid"]["objs"].append({"type": "inlet", "center": [0.3, 0.17, 0.1], "size": [0.05, 0.05],
                                "norm": [0, -1, 0], "speed": 1.0, "volume_frac": [1], "color": "0x0000ff",
                                "particle_relaxing_factor": 1.0, "end_time": 0.03})
    cb = dict(cb, gui_fps=500)
    print("[case] growing scenario, 3d_8k + timed cube + inlet", flush=True)
    sim = simulation_from_dicts(cb, sb)
    n0 = sim.fluid.part_num
    counts = []
    for _ in range(6):
        sim.run_step()
        counts.append(sim.fluid.part_num)
    print(f"  part_num per frame {counts} (start {n0}), t={sim.time_count:.4g}", flush=True)
    assert counts[-1] > n0 + 5 ** 3  # the cube and several inlet sheets arrived
    assert all(b >= a for a, b in zip(counts, counts[1:]))
    f = sim.fluid
    n = f.part_num
    pid = f.t["pid"][:n].cpu().numpy()
    assert np.array_equal(np.sort(pid), np.arange(n))
    # hand the grown state to the oracle (built from the scene without the extras)
    o = O.Oracle(cb, base, 4)
    o.init_scenario()
    o.set_part_num(0, n)
    for name in ("pos", "vel", "mass", "rest_density", "rest_volume", "volume_frac"):
        o.set(0, name, f.to_numpy(name))
    s = sim.sph_step(stats=True)
    st = o.step(O.STEP_CFL)
    assert (s.div_iter, s.incom_iter) == (st.div_iter, st.incom_iter)
    compare_fields(sim, o, ["pos", "vel", "sph_density", "alpha"], label="grown")
