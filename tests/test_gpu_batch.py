"""GPU tests of batched scenes (tichi_sph_amd/batch.py, tsph_batch_*):
a grouped launch over a scene table runs the same per-block bodies as the
single-scene kernels, so every scene of a batch must match its own
Simulation.sph_step() bit for bit, in eager mode, under hipGraph replay and
for every grouped-launch block count."""
import numpy as np
import pytest
import torch

from conftest import load_scene

pytestmark = pytest.mark.gpu


def scene_dicts():
    from tichi_sph_amd.scenarios.make_scenarios import batched_dambreaks
    out = [(n, s, c) for n, s, c in batched_dambreaks(n_scenes=4, seed=1, n_min=2e3, n_max=2e4)]
    for name in ("dambreak_3d_8k", "dambreak_2d_8k"):
        cb, sb = load_scene(name)
        out.append((name, sb, cb))
    return out


def make_sims(dicts, precision=4):
    from tichi_sph_amd import simulation_from_dicts
    return [simulation_from_dicts(c, s, precision=precision, cfl=False) for _, s, c in dicts]


def state(sim):
    f = sim.fluid
    return f.to_numpy("pos"), f.to_numpy("vel")


def assert_same(sims_a, sims_b, label):
    for s, (a, b) in enumerate(zip(sims_a, sims_b)):
        pa, va = state(a)
        pb, vb = state(b)
        assert np.array_equal(pa, pb), (label, s, "pos")
        assert np.array_equal(va, vb), (label, s, "vel")


def test_batched_step_bit_identical_to_single_scenes():
    from tichi_sph_amd.batch import SceneBatch
    d = scene_dicts()
    print(f"[case] batch of {len(d)} scenes: {[n for n, _, _ in d]}", flush=True)
    ref = make_sims(d)
    batch = SceneBatch(make_sims(d))
    for step in range(4):
        di, ii = batch.step()
        for s, sim in enumerate(ref):
            st = sim.sph_step(stats=True)
            assert (int(di[s]), int(ii[s])) == (st.div_iter, st.incom_iter), (step, s)
        assert_same(batch.sims, ref, f"step {step}")
    print(f"  iterations per scene: {list(zip(di.tolist(), ii.tolist()))}", flush=True)


def test_graph_replay_matches_eager_batch():
    from tichi_sph_amd.batch import SceneBatch
    d = scene_dicts()
    print(f"[case] graph replay, {len(d)} scenes", flush=True)
    eager = SceneBatch(make_sims(d))
    graphed = SceneBatch(make_sims(d))
    eager.step()
    graphed.step()
    graphed.capture()
    for step in range(4):
        eager.step()
        graphed.step()
        torch.cuda.synchronize()
        assert_same(graphed.sims, eager.sims, f"replay {step}")
        _ = [s.fluid.to_numpy("vel").sum() for s in eager.sims]  # eager work between replays


def test_autotuned_grid_does_not_change_results():
    from tichi_sph_amd.batch import SceneBatch
    d = scene_dicts()
    print(f"[case] autotune, {len(d)} scenes", flush=True)
    ref = SceneBatch(make_sims(d))
    tuned = SceneBatch(make_sims(d))
    grid = tuned.autotune(candidates=[0, 7, 64])
    print(f"  chosen grid {grid}", flush=True)
    for step in range(3):
        ref.step()
        tuned.step()
        assert_same(tuned.sims, ref.sims, f"step {step}")
