"""Host-side scenario ingestion (main.py:42-112), PLY IO (sph_util.py:83-99),
output writers (main.py:115-171, 371-378), Grid (sph_obj.py:326-344) and the
oracle's map_velocity (sph.py:382-397).  No GPU: Simulations are host-only.

The reference has no tests for these; each is pinned by a property its
formula must satisfy (kinematics of the inlet interval, geometry of the
inlet sheet, PLY round trips, brute force for map_velocity)."""
import json
import os
import struct

import numpy as np
import pytest

import oracle as O
from conftest import load_scene


def host_sim(name="dambreak_3d_8k", extra_objs=(), fluid_max=None):
    from tichi_sph_amd import Config, Simulation
    cb, sb = load_scene(name)
    sb = json.loads(json.dumps(sb))
    for part, obj in extra_objs:
        sb[part]["objs"].append(obj)
    cfg = Config(cb, sb)
    sim = Simulation(cfg, device="cpu", fluid_max=fluid_max)
    sim.init_scenario()
    return sim


# ------------------------------------------------------------------ PLY
def test_ply_ascii_round_trip(tmp_path):
    from tichi_sph_amd.ply import read_ply, write_ply
    rng = np.random.default_rng(0)
    pos = rng.uniform(-1, 1, (57, 3)).astype(np.float32)
    fname = write_ply(tmp_path / "pts", 7, 3, 57, pos)
    assert fname.endswith("pts_7.ply")
    back = read_ply(fname)
    assert back.dtype == np.float32 and np.array_equal(back, pos)
    pos2 = rng.uniform(-1, 1, (9, 2)).astype(np.float32)
    back2 = read_ply(write_ply(tmp_path / "pts2", 0, 2, 9, pos2))
    assert np.array_equal(back2[:, :2], pos2) and np.all(back2[:, 2] == 0)


def test_ply_binary_with_extra_elements(tmp_path):
    from tichi_sph_amd.ply import read_ply
    verts = np.array([[0.5, 1.5, -2.0], [3.25, 0.0, 1e-3], [7.0, 8.0, 9.0]], dtype=np.float32)
    hdr = ("ply\nformat binary_little_endian 1.0\ncomment made by a test\n"
           "element camera 1\nproperty double fov\n"
           "element vertex 3\nproperty uchar label\nproperty float x\nproperty float y\n"
           "property float z\nproperty double w\n"
           "element face 1\nproperty list uchar int vertex_indices\nend_header\n").encode()
    body = struct.pack("<d", 60.0)
    for i, v in enumerate(verts):
        body += struct.pack("<Bfffd", i, *v, 1.0)
    body += struct.pack("<Biii", 3, 0, 1, 2)
    p = tmp_path / "b.ply"
    p.write_bytes(hdr + body)
    assert np.array_equal(read_ply(p), verts)


def test_ply_scenario_object(tmp_path):
    from tichi_sph_amd.ply import write_ply
    pts = np.array([[0.1, 0.1, 0.1], [0.12, 0.1, 0.1], [0.1, 0.12, 0.1]], dtype=np.float32)
    fname = write_ply(tmp_path / "drop", 0, 3, 3, pts)
    obj = {"type": "ply", "file_name": fname, "volume_frac": [1], "vel": [0, -1, 0], "color": "0x00ff00"}
    base = host_sim().fluid.part_num
    sim = host_sim(extra_objs=[("fluid", obj)], fluid_max=base + 3)
    f = sim.fluid
    assert f.part_num == base + 3
    assert np.array_equal(f.to_numpy("pos")[base:], pts)
    assert np.array_equal(f.to_numpy("vel")[base:], np.tile(np.float32([0, -1, 0]), (3, 1)))


# --------------------------------------------------------------- inlets
def test_inlet_interval_is_the_time_to_advance_one_spacing():
    from tichi_sph_amd.solver import inlet_helper_get_dt
    cfg = host_sim().config
    r = cfg.part_size[1]
    for norm, speed in (([0, -1, 0], 0.5), ([1, -1, 0], 1.0), ([0, -1, 0], 0.0)):
        dt = inlet_helper_get_dt(norm, speed, 1.0, cfg)
        n = np.array(norm, dtype=np.float64) / np.linalg.norm(norm)
        a = float(np.dot(n, np.array(cfg.gravity, dtype=np.float32)))
        # s(dt) = v dt + a dt^2 / 2 reaches one spacing
        assert speed * dt + 0.5 * a * dt * dt == pytest.approx(r, rel=1e-9)
    # no acceleration along the normal: r / v, or never
    assert inlet_helper_get_dt([1, 0, 0], 2.0, 1.5, cfg) == pytest.approx(r * 1.5 / 2.0)
    assert inlet_helper_get_dt([1, 0, 0], 0.0, 1.0, cfg) == float("inf")


def test_inlet_sheet_geometry_and_emission_schedule():
    center, size, norm, speed = [0.1, 0.15, 0.1], [0.05, 0.03], [0.3, -1.0, 0.2], 0.7
    inlet = {"type": "inlet", "center": center, "size": size, "norm": norm, "speed": speed,
             "volume_frac": [1], "color": "0x0000ff", "particle_relaxing_factor": 1.0, "end_time": 1.0}
    sim = host_sim(extra_objs=[("fluid", inlet)], fluid_max=20000)
    base = sim.fluid.part_num
    assert len(sim.inlets) == 1 and sim.fluid.part_num == base  # emitted at refresh, not at init
    sim.refresh_scenario()
    f = sim.fluid
    k = f.part_num - base
    spacing = sim.config.part_size[1]
    shape, _ = f.scene_add_help_centering([0, 0], size, spacing)
    assert k == int(shape[0]) * int(shape[1]) > 0
    x = f.to_numpy("pos")[base:].astype(np.float64)
    n = np.array(norm) / np.linalg.norm(norm)
    assert np.abs((x - np.array(center)) @ n).max() < 1e-6  # on the plane through center
    assert np.abs(x.mean(0) - np.array(center)).max() < spacing  # centred
    d = np.linalg.norm(x[1] - x[0])
    assert d == pytest.approx(spacing, rel=1e-5)  # neighbouring sheet points one spacing apart
    assert np.allclose(f.to_numpy("vel")[base:], np.float32(np.array(norm) * speed))
    # schedule: nothing new before the interval has passed, one sheet after it
    sim.refresh_scenario()
    assert f.part_num == base + k
    sim.globalvar.time_count += sim.inlets[0]["dt"]
    sim.refresh_scenario()
    assert f.part_num == base + 2 * k
    sim.globalvar.time_count = 1.0  # end_time reached: the inlet is removed
    sim.refresh_scenario()
    assert sim.inlets == [] and f.part_num == base + 2 * k


def test_timed_object_waits_for_its_time():
    cube = {"type": "cube", "start_pos": [0.25, 0.1, 0.05], "end_pos": [0.3098, 0.1598, 0.1098], "vel": [0, 0, 0],
            "volume_frac": [1], "color": "0xff0000", "particle_relaxing_factor": 1.0, "time": 0.01}
    sim = host_sim(extra_objs=[("fluid", cube)], fluid_max=20000)
    base = sim.fluid.part_num
    assert len(sim.scenario_wait_buffer) == 1
    sim.globalvar.time_count = 0.005
    sim.refresh_scenario()
    assert sim.fluid.part_num == base
    sim.globalvar.time_count = 0.01
    sim.refresh_scenario()
    assert sim.fluid.part_num == base + 5 ** 3 and sim.scenario_wait_buffer == []
    pid = sim.fluid.t["pid"][: sim.fluid.part_num].numpy()
    assert np.array_equal(np.sort(pid), np.arange(sim.fluid.part_num))


def test_host_only_simulation_refuses_compute():
    from tichi_sph_amd._lib import TsphError
    sim = host_sim()
    with pytest.raises(TsphError, match="host-only"):
        sim.sph_step()


# -------------------------------------------------------------- writers
def test_writers_follow_the_reference_schema(tmp_path):
    from tichi_sph_amd.grid import Grid
    from tichi_sph_amd.output import write_files, write_full_json, write_json, write_scene_data
    sim = host_sim()
    sim.globalvar.time_counter, sim.globalvar.step_counter = 3, 17
    sim.globalvar.frame_div_iter, sim.globalvar.frame_incom_iter = 5, 6
    d = write_full_json(sim, tmp_path / "f.json")
    assert set(d) == {"step", "frame", "timeInSimulation", "timeStep", "fps", "iteration", "energy"}
    assert d["iteration"] == {"divergenceFree_iteration": 5, "incompressible_iteration": 6, "sum_iteration": 11}
    assert set(d["energy"]) == {"statistics_kinetic_energy", "statistics_gravity_potential_energy", "sum_energy"}
    assert json.load(open(tmp_path / "f.json")) == d
    write_json(sim, str(tmp_path / "js"))
    assert os.path.exists(tmp_path / "js" / "DFSPH17.json")
    g = Grid(sim)
    s = write_scene_data(sim, g, tmp_path / "scene.json")
    assert s["grid_count"] == g.size and s["fluid_part_count"] == sim.fluid.part_num
    write_files(sim, str(tmp_path))
    pos = np.load(tmp_path / "DFSPH" / "part_data" / "pos_3.npy")
    assert np.array_equal(pos, sim.fluid.to_numpy("pos"))
    from tichi_sph_amd.ply import read_ply
    assert np.array_equal(read_ply(tmp_path / "DFSPH" / "ply" / "fluid_pos_3.ply"), pos)


def test_grid_shape_intended_and_literal():
    from tichi_sph_amd.grid import Grid
    sim = host_sim()
    cfg = sim.config
    lb = np.array(cfg.sim_space_lb, dtype=np.float32)
    rt = np.array(cfg.sim_space_rt, dtype=np.float32)
    g = Grid(sim)
    # (rt - lb) / dist nodes per axis, dist = one neighbour cell
    assert g.dist == pytest.approx(cfg.kernel_h[1])
    assert g.shape == tuple(int(v) for v in ((rt - lb) / np.float32(g.dist)).astype(np.int32))
    assert np.all(g.pos_numpy()[-1, -1, -1] <= rt)
    lit = Grid(sim, literal_shape=True)  # sph_obj.py:326 as written
    assert lit.shape == tuple(int(v) for v in (rt - lb / cfg.part_size[1] * 2).astype(np.int32))


# ------------------------------------------------ map_velocity (oracle pin)
def test_oracle_map_velocity_equals_brute_force():
    cb, sb = load_scene("dambreak_2d_8k")
    o = O.Oracle(cb, sb, 8)
    o.init_scenario()
    n = o.n[0]
    rng = np.random.default_rng(1)
    v = rng.normal(0, 0.1, (n, 2))
    o.set(0, "vel", v)
    for k in ("clear_node",):
        o.kernel(k, 0)
    o.kernel("encode", 0); o.kernel("encode", 1); o.kernel("mem_shift", 0)
    o.kernel("fill_node", 0); o.kernel("fill_node", 1)
    o.kernel("SPH_clean_value", 0)
    o.kernel("SPH_prepare_attr", 0, 0)
    o.kernel("SPH_prepare_attr", 0, 1)
    h = o.c.kernel_h[1]
    lb = np.array(sb["sim_env"]["sim_space_lb"], dtype=np.float32).astype(np.float64)
    shape = [20, 12]
    g = o.map_velocity(0, shape, h)
    x = o.get(0, "pos")
    X = o.get(0, "mass")
    rho = o.get(0, "sph_density")
    I = np.stack(np.meshgrid(*[np.arange(s) for s in shape], indexing="ij"), -1).reshape(-1, 2)
    gp = lb + I * h
    sig = o.c.kernel_sig[2]
    ref = np.zeros_like(g)
    for a, p in enumerate(gp):
        r = np.linalg.norm(p - x, axis=1)
        near = np.where(r < h)[0]
        w = np.array([O.lib().oracle_kernel_fn(0, 8, h, sig, float(rr)) for rr in r[near]])
        ref[a] = ((X[near] / rho[near])[:, None] * v[near] * w[:, None]).sum(0)
    assert np.abs(ref).max() > 1e-3  # the grid covers fluid
    assert np.allclose(g, ref, rtol=1e-12, atol=1e-15)
