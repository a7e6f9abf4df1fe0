"""Writes the benchmark scenes as scenario / solver-config JSON files in the
reference's format (scenario/*.json, config/config.json).

Layout rules (all in units of the particle spacing s = 2r):
  * the fluid block and the wall lattice share one lattice;
  * `cube` objects are given an extent of n*s + 0.98*s, so the reference's
    centring (sph_obj.py:102-107) puts the first centre 0.49*s from the
    object's start and places exactly n particles per axis;
  * walls (`box`, `layers` thick) sit outside the interior;
  * the grid has one cell (h = 2s) of padding beyond the walls, so no particle
    is in cell 0 or on a grid face, and every centre is >= 0.245 cells away
    from a cell boundary (no key depends on the last ulp).
Run:  python -m tichi_sph_amd.scenarios.make_scenarios
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def solver_config(solver_type, cs, div_thr=1e-3, comp_thr=1e-3, it_min=2, it_max=100,
                  fbm_d=0.0, fbm_c=50, gamma=0.0):
    return {
        "arch": "gpu", "device_memory_GB": 6,
        "solver_type": solver_type, "solver_wc_gamma": 7,
        "solver_divergence_threshold": div_thr, "solver_compression_threshold": comp_thr,
        "solver_iter_threshold_min": it_min, "solver_iter_threshold_max": it_max,
        "solver_if_cfl": True, "solver_cs": cs, "solver_cfl_factor": 0.5,
        "solver_neighb_range": 1, "solver_neighb_grid_size_TO_global_part_size": 2,
        "solver_fbm_diffusion_term": fbm_d, "solver_fbm_convection_term": fbm_c,
        "solver_surface_tension_gamma": gamma,
        "gui_res": [1080, 1080], "gui_part_zoomer": 1.0, "gui_fps": 30,
        "gui_canvas_bgcolor": [0.2, 0.2, 0.6], "gui_camera_pos": [6.0, 1.0, 0.0],
        "gui_camera_lookat": [0.0, 0.0, 0.0],
    }


def _r(x):
    return round(x, 12)


def tank(gps, interior, layers, fluid_blocks, phase_rest_density, gravity, mu, max_extra=0,
         phase_mu=None):
    """interior: lattice points per axis inside the walls; fluid_blocks: list
    of (start_idx, count, volume_frac) per block, in lattice units; phase_mu:
    per-phase dynamic viscosities (the phase_dynamic_viscosity extension)."""
    s = gps
    d = len(interior)
    fluid_objs = []
    nf = 0
    for start_idx, count, vf in fluid_blocks:
        st = [_r(start_idx[a] * s) for a in range(d)]
        en = [_r(start_idx[a] * s + count[a] * s + 0.98 * s) for a in range(d)]
        fluid_objs.append({"type": "cube", "vel": [0] * d, "particle_relaxing_factor": 1.0,
                           "start_pos": st, "end_pos": en, "volume_frac": vf, "color": "0x6F7DBC"})
        n = 1
        for a in range(d):
            n *= count[a]
        nf += n
    nbox = [interior[a] + 2 * layers for a in range(d)]
    bst = [_r(-layers * s) for a in range(d)]
    ben = [_r(-layers * s + nbox[a] * s + 0.98 * s) for a in range(d)]
    nb = 1
    ni = 1
    for a in range(d):
        nb *= nbox[a]
        ni *= interior[a]
    nb -= ni
    lb = [_r(-layers * s - 2 * s) for a in range(d)]
    rt = [_r(ben[a] + 2 * s) for a in range(d)]
    P = len(phase_rest_density)
    env = {
        "global_part_size": gps, "global_dynamic_viscosity": mu, "sim_dim": d,
        "sim_space_lb": lb, "sim_space_rt": rt, "gravity": gravity, "phase_num": P,
        "phase_rest_density": phase_rest_density,
        "phase_color_hex": ["0x6F7DBC", "0xFA8012", "0xFA8072"][:P],
    }
    if phase_mu is not None:
        env["phase_dynamic_viscosity"] = phase_mu
    return {
        "sim_env": env,
        "fluid": {"max_part_num": nf + max_extra, "objs": fluid_objs},
        "bound": {"max_part_num": nb, "objs": [
            {"type": "box", "vel": [0] * d, "particle_relaxing_factor": 1.0, "start_pos": bst,
             "end_pos": ben, "layers": layers, "volume_frac": [1] + [0] * (P - 1), "color": "0xAAAAAA"}]},
    }


def scenes():
    out = {}
    # 2D dam break: 128 x 64 fluid in a 4.0 x 2.0 m tank, 3 wall layers, r = 0.01
    out["dambreak_2d_8k"] = (tank(0.02, [200, 100], 3, [([0, 0], [128, 64], [1])], [1000], [0, -9.81], 0.01),
                             solver_config("DFSPH", cs=50))
    # 3D dam break: 100^3 fluid in a 2.0 (x) x 1.0 (y, up) x 1.0 (z) m tank, 2 wall layers, r = 0.005
    out["dambreak_3d_1m"] = (tank(0.01, [200, 100, 100], 2, [([0, 0, 0], [100, 100, 100], [1])], [1000],
                                  [0, -9.81, 0], 0.01), solver_config("DFSPH", cs=100))
    # Small 3D dam break for parity tests: 20^3 fluid in a 40 x 20 x 20 tank
    out["dambreak_3d_8k"] = (tank(0.01, [40, 20, 20], 2, [([0, 0, 0], [20, 20, 20], [1])], [1000],
                                  [0, -9.81, 0], 0.01), solver_config("DFSPH", cs=100))
    # 3D two-phase VFSPH: 125^3 in a 1 m box, phase A (1000, viscosity 0.01) below, phase B
    # (500, viscosity 0.05) above, r = 0.004
    out["two_phase_3d_2m"] = (tank(0.008, [125, 125, 125], 2,
                                   [([0, 0, 0], [125, 62, 125], [1, 0]), ([0, 62, 0], [125, 63, 125], [0, 1])],
                                   [1000, 500], [0, -9.81, 0], 0.01, phase_mu=[0.01, 0.05]),
                              solver_config("VFSPH", cs=80, fbm_d=1e-4, fbm_c=50, gamma=0.07))
    # Small two-phase scene for parity tests: 16^3
    out["two_phase_3d_4k"] = (tank(0.008, [16, 16, 16], 2,
                                   [([0, 0, 0], [16, 8, 16], [1, 0]), ([0, 8, 0], [16, 8, 16], [0, 1])],
                                   [1000, 500], [0, -9.81, 0], 0.01, phase_mu=[0.01, 0.05]),
                              solver_config("VFSPH", cs=80, fbm_d=1e-4, fbm_c=50, gamma=0.07))
    # Differentiable dam break: 50 x 80 x 50 = 200,000 fluid in a 1.0 x 1.0 x 0.5 m tank, r = 0.005,
    # starting from mu = 0.01, gamma = 0.07 (the gradients are taken w.r.t. both)
    out["diff_3d_200k"] = (tank(0.01, [100, 100, 50], 2, [([0, 0, 0], [50, 80, 50], [1])], [1000],
                                [0, -9.81, 0], 0.01), solver_config("DFSPH", cs=100, gamma=0.07))
    return out


def batched_dambreaks(n_scenes=64, seed=0, n_min=1e4, n_max=2e5):
    """Independent 3D DFSPH dam breaks for a batch: fluid counts log-uniform
    in [n_min, n_max], block aspect ratios (x/y, z/y) uniform in [0.5, 2],
    the tank's interior twice the block per axis, viscosity uniform in
    [0.005, 0.05], r = 0.005, dt = 1e-4 (cs = 100); numpy default_rng(seed).
    Returns [(name, scenario, config)]."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_scenes):
        n = float(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))
        ax, az = rng.uniform(0.5, 2.0, 2)
        mu = float(rng.uniform(0.005, 0.05))
        ny = max(2, int(round((n / (ax * az)) ** (1.0 / 3.0))))
        nx, nz = max(2, int(round(ax * ny))), max(2, int(round(az * ny)))
        scn = tank(0.01, [2 * nx, 2 * ny, 2 * nz], 2, [([0, 0, 0], [nx, ny, nz], [1])], [1000],
                   [0, -9.81, 0], _r(mu))
        out.append((f"batch{seed}_{i:02d}_{nx}x{ny}x{nz}", scn, solver_config("DFSPH", cs=100)))
    return out


def main():
    for name, (scn, cfg) in scenes().items():
        with open(os.path.join(HERE, name + ".json"), "w") as f:
            json.dump(scn, f, indent=1)
        with open(os.path.join(HERE, name + ".config.json"), "w") as f:
            json.dump(cfg, f, indent=1)


if __name__ == "__main__":
    main()
