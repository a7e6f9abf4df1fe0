"""GPU tests of batched scenes (tichi_sph_amd/batch.py, tsph_batch_*):
a grouped launch over a scene table runs the same per-block bodies as the
single-scene kernels, so every scene of a batch must match its own
Simulation.sph_step() bit for bit, in eager mode, under hipGraph replay and
for every grouped-launch block count.  The batch mixes 2D and 3D DFSPH dam
breaks with a 3D VFSPH two-phase mixture (surface tension, per-phase
viscosity)."""
import numpy as np
import pytest
import torch

from conftest import load_scene

pytestmark = pytest.mark.gpu


def scene_dicts():
    from tichi_sph_amd.scenarios.make_scenarios import batched_dambreaks
    out = [(n, s, c) for n, s, c in batched_dambreaks(n_scenes=4, seed=1, n_min=2e3, n_max=2e4)]
    # a VFSPH two-phase mixture with surface tension and per-phase viscosity
    # (no FBM: not batched) beside the DFSPH dam breaks
    for name in ("dambreak_3d_8k", "two_phase_3d_4k", "dambreak_2d_8k"):
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


def test_batch_follows_a_scene_stepped_on_its_own():
    """A scene of a batch stepped by its own sph_step() between batched steps
    (its fluid is sorted on the host's path: front and back buffers trade
    places): the next batched step rebuilds the scene table and every scene
    stays bit-identical to the same sequence run scene by scene."""
    from tichi_sph_amd.batch import SceneBatch
    d = scene_dicts()
    print(f"[case] batch of {len(d)} scenes with a single-scene step in between", flush=True)
    ref = make_sims(d)
    batch = SceneBatch(make_sims(d))
    batch.step()
    for sim in ref:
        sim.sph_step()
    for s in (0, 3):
        batch.sims[s].sph_step()
        ref[s].sph_step()
    for step in range(2):
        batch.step()
        for sim in ref:
            sim.sph_step()
        assert_same(batch.sims, ref, f"after the single-scene steps, batched step {step}")


def test_batched_fp64_step_bit_identical_to_single_scenes():
    """The grouped kernels' FP64 instantiation (the precision of the
    differentiable runs) against the FP64 single-scene steps."""
    from tichi_sph_amd.batch import SceneBatch
    d = scene_dicts()[-3:]
    print(f"[case] FP64 batch of {len(d)} scenes: {[n for n, _, _ in d]}", flush=True)
    ref = make_sims(d, precision=8)
    batch = SceneBatch(make_sims(d, precision=8))
    for step in range(3):
        di, ii = batch.step()
        for s, sim in enumerate(ref):
            st = sim.sph_step(stats=True)
            assert (int(di[s]), int(ii[s])) == (st.div_iter, st.incom_iter), (step, s)
        assert_same(batch.sims, ref, f"fp64 step {step}")


def test_batched_kept_boundary_sums_follow_boundary_writes():
    """The grouped pass A keeps each static boundary's (b,b) density sums
    after its first step (the table's TSPH_OBJ_STATIC_SUMS).  Writing one
    scene's boundary masses afterwards must make the batch sum them again:
    the boundary densities and the fluid states stay bit-identical to the
    scenes' own steps, which see the same write."""
    from tichi_sph_amd.batch import SceneBatch
    d = scene_dicts()
    print(f"[case] kept boundary sums in a batch of {len(d)} scenes, then a boundary mass write", flush=True)
    ref = make_sims(d)
    batch = SceneBatch(make_sims(d))
    before = None
    for step in range(4):
        if step == 2:
            before = batch.sims[0].bound.to_numpy("sph_density")
            for sim in (batch.sims[0], ref[0]):
                m = sim.bound.to_numpy("mass")
                sim.bound.from_numpy("mass", (m * np.float32(1.01)).astype(m.dtype))
        batch.step()
        for sim in ref:
            sim.sph_step()
        assert_same(batch.sims, ref, f"step {step}")
        for s, (a, b) in enumerate(zip(batch.sims, ref)):
            assert np.array_equal(a.bound.to_numpy("sph_density"), b.bound.to_numpy("sph_density")), (step, s)
    after = batch.sims[0].bound.to_numpy("sph_density")
    print(f"  boundary density of scene 0 before / after the write: {before.mean():.6g} / {after.mean():.6g}",
          flush=True)
    assert not np.array_equal(before, after)


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


def test_grid_stride_launches_match_one_block_per_tile():
    """Every grouped kernel with 7 blocks (each block strides over many tiles,
    across scenes and, in pass A, from the fluid tiles into the boundary
    tiles) gives the one-block-per-tile results bit for bit."""
    from tichi_sph_amd.batch import SceneBatch
    d = scene_dicts()
    print(f"[case] grid 7 against grid 0, {len(d)} scenes", flush=True)
    ref = SceneBatch(make_sims(d))
    strided = SceneBatch(make_sims(d))
    for g in strided.grid:
        strided.grid[g] = 7
    for step in range(3):
        ref.step()
        strided.step()
        assert_same(strided.sims, ref.sims, f"step {step}")


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
