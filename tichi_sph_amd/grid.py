"""Particle-to-grid export: `Grid` (sph_obj.py:326-344) for map_velocity
(sph.py:382-397, see tichi_sph_amd.sph.map_velocity).

Node I sits at sim_space_lb + I * dist with dist = global_part_size *
neighb_grid_size_TO_global_part_size (one neighbour cell).  The reference's
shape expression, `rt - lb / part_size * ratio` (sph_obj.py:326), divides
only lb by the spacing; the intended node count per axis is
(rt - lb) / dist, which is the default here; `literal_shape=True` keeps the
expression as written.
"""
import numpy as np
import torch


class Grid:
    """`Grid()` (the reference's form, main.py:11) is built on the Simulation
    of the Config bound by tichi_sph_amd.use_config; `Grid(sim)` on `sim`."""

    def __init__(self, sim=None, literal_shape=False):
        if sim is None:
            from . import binding
            sim = binding.simulation()
        cfg = sim.config
        d = cfg.dim
        lb = np.array(cfg.sim_space_lb[:d], dtype=np.float32)
        rt = np.array(cfg.sim_space_rt[:d], dtype=np.float32)
        ps = cfg.part_size[1]
        ratio = cfg.neighb_grid_size_TO_global_part_size
        if literal_shape:
            shape = (rt - lb / ps * ratio).astype(np.int32)
        else:
            shape = ((rt - lb) / np.float32(ps * ratio)).astype(np.int32)
        self.shape = tuple(int(s) for s in shape)
        self.lb = lb
        self.dist = ps * ratio
        self.size = int(np.prod(self.shape))
        self.dim = d
        self.vel = torch.zeros(max(self.size, 1), 4, dtype=sim.fluid.dtype, device=sim.device)

    def pos_numpy(self):
        """Node positions lb + I * dist (float32), shape + (dim,)."""
        idx = np.stack(np.meshgrid(*[np.arange(s) for s in self.shape], indexing="ij"), axis=-1)
        return self.lb + idx.astype(np.float32) * np.float32(self.dist)

    def vel_numpy(self):
        """Node velocities, shape + (dim,) (the reference field's to_numpy())."""
        return self.vel[: self.size, : self.dim].cpu().numpy().reshape(self.shape + (self.dim,))
