"""Output files of the reference's run loop (main.py:115-171, 371-378).

write_full_json / write_json: per-frame / per-step statistics with the
reference's schema (step, frame, simulated time, dt, fps, iteration counts,
energies).  write_scene_data: the grid and scene summary.  write_files: the
PLY positions, the frame JSON and the npy velocity / position arrays, in the
reference's directory layout under `root` (`<solver>/ply`, `<solver>/json`,
`<solver>/part_data`; the reference hard-codes Windows separators, here the
platform's).  The GUI image of write_files is not produced.  Arrays are in
the reference's particle order (creation order).
"""
import json
import os

import numpy as np

from .ply import write_ply


def _control(sim):
    """The device control block, where the step (SPH_update_energy,
    sph.py:272-279) and cfl_condition (sph.py:41-47) leave the energies and
    dt, or None for a host-only Simulation (nothing has been computed)."""
    if getattr(sim, "device", None) is not None and sim.device.type == "cuda":
        return sim.control()
    return None


def _energy(sim):
    c = _control(sim)
    return (float(c.ke), float(c.pe)) if c is not None else (0.0, 0.0)


def _dt(sim):
    c = _control(sim)
    return float(c.dt) if c is not None else sim.config.dt


def write_full_json(sim, fname):
    """main.py:150-171 (frame-accumulated iteration counts)."""
    g = sim.globalvar
    ke, pe = _energy(sim)
    data = {
        "step": g.step_counter,
        "frame": g.time_counter,
        "timeInSimulation": g.time_count,
        "timeStep": _dt(sim),
        "fps": sim.config.gui_fps,
        "iteration": {
            "divergenceFree_iteration": g.frame_div_iter,
            "incompressible_iteration": g.frame_incom_iter,
            "sum_iteration": g.frame_div_iter + g.frame_incom_iter,
        },
        "energy": {
            "statistics_kinetic_energy": ke,
            "statistics_gravity_potential_energy": pe,
            "sum_energy": ke + pe,
        },
    }
    with open(fname, "w") as f:
        f.write(json.dumps(data))
    return data


def write_json(sim, directory="json"):
    """main.py:127-147: the last step's iteration counts, to
    <directory>/<solver><step>.json."""
    g = sim.globalvar
    ke, pe = _energy(sim)
    data = {
        "step": g.step_counter,
        "frame": g.time_counter,
        "timeInSimulation": g.time_count,
        "timeStep": _dt(sim),
        "fps": sim.config.gui_fps,
        "iteration": {
            "divergenceFree_iteration": g.div_iter_count,
            "incompressible_iteration": g.incom_iter_count,
            "sum_iteration": g.div_iter_count + g.incom_iter_count,
        },
        "energy": {
            "statistics_kinetic_energy": ke,
            "statistics_gravity_potential_energy": pe,
            "sum_energy": ke + pe,
        },
    }
    os.makedirs(directory, exist_ok=True)
    with open(os.path.join(directory, f"{sim.config.solver_type}{g.step_counter}.json"), "w") as f:
        f.write(json.dumps(data))
    return data


def write_scene_data(sim, grid, fname):
    """main.py:115-125."""
    data = {
        "grid_count": int(grid.size),
        "grid_lb": [float(v) for v in grid.lb],
        "init_part_size": sim.config.part_size[1],
        "fluid_part_count": sim.fluid.part_num,
        "bound_part_count": sim.bound.part_num,
        "dt": _dt(sim),
        "frame_rate": sim.config.gui_fps,
    }
    with open(fname, "w") as f:
        json.dump(data, f)
    return data


def write_files(sim, root="."):
    """main.py:371-378 without the window image: PLY of the fluid positions,
    the frame JSON, vel / pos npy arrays."""
    solver = sim.config.solver_type
    frame = sim.globalvar.time_counter
    for sub in ("ply", "json", "part_data"):
        os.makedirs(os.path.join(root, solver, sub), exist_ok=True)
    f = sim.fluid
    pos = f.to_numpy("pos")
    vel = f.to_numpy("vel")
    write_ply(os.path.join(root, solver, "ply", "fluid_pos"), frame, sim.config.dim, f.part_num, pos)
    write_full_json(sim, os.path.join(root, solver, "json", f"frame{frame}.json"))
    np.save(os.path.join(root, solver, "part_data", f"vel_{frame}"), vel)
    np.save(os.path.join(root, solver, "part_data", f"pos_{frame}"), pos)
