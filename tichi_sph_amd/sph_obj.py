"""Particle storage (`Fluid`) and the reference-layout neighbour grid (`Ngrid`).

Mirrors sph_obj.py:8-323 of the reference.  Storage lives in device tensors
owned by this object; the HIP library only receives raw pointers through a
`TsphObject` descriptor.  Vector fields are [capacity, 4] (z = 0 in 2D), and
pos[:, 3] carries the solver alias X (mass for DFSPH, rest_volume for VFSPH).

The fused step keeps particles in cell order; `pid` maps each storage slot
to the reference particle index (creation order).  Use `to_numpy(name)` /
`from_numpy(name, array)` to read or write a field in reference order.
"""
import ctypes

import numpy as np
import torch

from . import _lib


def _np_normalize(v):
    """np_normalize, sph_util.py:74-75."""
    return v / np.linalg.norm(v)


VEC_FIELDS = ("pos", "vel", "vel_adv", "acce_adv", "acce", "normal", "alpha_1")
SCALAR_FIELDS = ("mass", "rest_density", "rest_volume", "W", "sph_compression", "sph_density",
                 "psi_adv", "alpha", "alpha_2", "kappa", "vol", "fbm_zeta", "pressure")
INT_FIELDS = ("pid", "flag", "cell", "in_cell", "nbr_nf", "nbr_nb")


class _Scalar0D:
    """A 0-D device value read as `obj.x[None]`, as the reference reads its
    0-D fields (e.g. fluid.compression[None], main.py:201, 222)."""

    def __init__(self, read):
        self._read = read

    def __getitem__(self, k):
        if k is not None:
            raise TypeError("a 0-D field is indexed with [None]")
        return self._read()


class Fluid:
    """A particle object (sph_obj.py:8-93).  Two forms:
    * `Fluid(max_part_num=n)`, the reference's (main.py:9-10): the object is
      built against the Config bound by tichi_sph_amd.use_config and takes the
      next uid in creation order (sph_obj.py:13-14), fluid first (uid 0),
      then the boundary (uid 1);
    * `Fluid(config, max_part_num, uid, device, precision)`, as a Simulation
      builds its own."""

    def __init__(self, config=None, max_part_num=None, uid=None, device=None, precision=None,
                 list_cap_f=64, list_cap_b=48):
        if not hasattr(config, "phase_num"):  # the reference's form: Fluid(max_part_num)
            from . import binding
            if max_part_num is None:
                max_part_num = config
            b = binding.current()
            if b.sim is not None:
                raise RuntimeError("the bound Config's Simulation already exists: create the objects first")
            config = b.config
            uid = len(b.objs) if uid is None else uid
            device = b.device if device is None else device
            precision = b.precision if precision is None else precision
            b.objs.append(self)
        if max_part_num is None:
            raise TypeError("Fluid needs max_part_num")
        self.config = config
        self.max_part_num = int(max_part_num)
        self.uid = 0 if uid is None else uid
        self.device = torch.device("cuda" if device is None else device)
        self.precision = 4 if precision is None else precision
        precision = self.precision
        device = self.device
        uid = self.uid
        self.sim = None  # the Simulation that owns the object (set by it)
        self.dtype = torch.float32 if precision == 4 else torch.float64
        self.np_dtype = np.float32 if precision == 4 else np.float64
        self.part_num = 0
        self.color = []
        self.list_cap_f = list_cap_f
        self.list_cap_b = list_cap_b
        cap = max(self.max_part_num, 1)
        P = config.phase_num
        z = lambda *s: torch.zeros(*s, dtype=self.dtype, device=self.device)  # noqa: E731
        zi = lambda *s: torch.zeros(*s, dtype=torch.int32, device=self.device)  # noqa: E731
        self._t = {}
        self._pending_swaps = 0  # front/back swaps done on the device, not yet mirrored here
        self.sums_epoch = 0  # invalidations of the kept boundary sums (invalidate_static_sums)
        self.layout_epoch = 0  # host-side storage changes: single-object sorts, added particles
        for name in VEC_FIELDS:
            self.t[name] = z(cap, 4)
        for name in SCALAR_FIELDS:
            self.t[name] = z(cap)
        self.t["volume_frac"] = z(cap, P)
        self.t["volume_frac_tmp"] = z(cap, P)
        self.t["drift_vel"] = z(cap, P, 4)
        for name in INT_FIELDS:
            self.t[name] = zi(cap)
        self.t["cell_structured"] = zi(cap, 4)
        self.t["cell_count"] = zi(config.node_num + 2)
        self.t["cell_start"] = zi(config.node_num + 2)
        self.t["nbr_f"] = zi(list_cap_f * cap)
        self.t["nbr_b"] = zi(list_cap_b * cap)
        # back buffers of the persistent state (swapped by the counting sort)
        for name in ("pos", "vel"):
            self.t[name + "_b"] = z(cap, 4)
        for name in ("mass", "rest_density", "rest_volume"):
            self.t[name + "_b"] = z(cap)
        self.t["volume_frac_b"] = z(cap, P)
        self.t["pid_b"] = zi(cap)
        # the boundary keeps its own-object density sums between steps
        self.t["static_sums"] = z(cap, 4) if self.uid == 1 else None
        self._struct = _lib.TsphObject()
        self._refresh_struct()
        self.cell_table_ok = False  # cell_start matches the storage (set by the sort)

    # ------------------------------------------------------------------ layout
    _SWAPPED = ("pos", "vel", "mass", "rest_density", "rest_volume", "volume_frac", "pid")

    @property
    def t(self):
        """The object's device tensors by field name (front buffers current)."""
        if self._pending_swaps:
            self._resolve_swaps()
        return self._t

    def note_device_swap(self):
        """A batched sort swapped this object's front/back state on the
        device (and in the batch's table): mirror it here when next used."""
        self._pending_swaps += 1
        self.cell_table_ok = True

    def _resolve_swaps(self):
        if self._pending_swaps & 1:
            t, s = self._t, self._struct
            for name in self._SWAPPED:
                nb = name + "_b"
                t[name], t[nb] = t[nb], t[name]
                a, b = getattr(s, name), getattr(s, nb)
                setattr(s, name, b)
                setattr(s, nb, a)
        self._pending_swaps = 0

    def _refresh_struct(self):
        s = self._struct
        s.n = self.part_num
        s.capacity = max(self.max_part_num, 1)
        s.uid = self.uid
        for name in _lib._OBJ_PTRS:
            t = self.t[name]
            setattr(s, name, t.data_ptr() if t is not None else None)

    def struct(self):
        if self._pending_swaps:
            self._resolve_swaps()
        self._struct.n = self.part_num
        return self._struct

    def adopt_swaps(self):
        """After the library sorted the object and swapped front/back pointers
        inside the descriptor, swap the tensors so that self.t matches the
        descriptor again."""
        self.cell_table_ok = True
        self.layout_epoch += 1  # a SceneBatch holding this object's descriptor rebuilds it
        s = self._struct
        for name in ("pos", "vel", "mass", "rest_density", "rest_volume", "volume_frac", "pid"):
            if getattr(s, name) != self.t[name].data_ptr():
                self.t[name], self.t[name + "_b"] = self.t[name + "_b"], self.t[name]
        self._refresh_struct()

    @property
    def compression(self):
        """fluid.compression (sph_obj.py), read as compression[None]: the
        average clamped density error of the last Jacobi iteration."""
        return _Scalar0D(lambda: self.sim.compression())

    # DFSPH/VFSPH aliases, sph_obj.py:65-73
    @property
    def X_name(self):
        return "rest_volume" if self.config.solver_type == "VFSPH" else "mass"

    @property
    def sph_psi_name(self):
        return "sph_compression" if self.config.solver_type == "VFSPH" else "sph_density"

    # --------------------------------------------------------------- field I/O
    def _storage(self, name):
        d = self.config.dim
        n = self.part_num
        if name in VEC_FIELDS:
            return self.t[name][:n, :d]
        if name == "drift_vel":
            return self.t[name][:n, :, :d]
        if name == "cell_structured" or name == "neighb_cell_structured_seq":
            return self.t["cell_structured"][:n, :d]
        alias = {"neighb_cell_seq": "cell", "neighb_in_cell_seq": "in_cell", "X": self.X_name,
                 "sph_psi": self.sph_psi_name}
        return self.t[alias.get(name, name)][:n]

    def to_numpy(self, name):
        """Field `name` (reference attribute name) as numpy, in reference
        particle order."""
        if name == "rest_psi":
            if self.config.solver_type == "VFSPH":
                return np.ones(self.part_num, dtype=self.np_dtype)
            name = "rest_density"
        a = self._storage(name).detach().cpu().numpy()
        pid = self.t["pid"][: self.part_num].cpu().numpy()
        out = np.empty_like(a)
        out[pid] = a
        return out

    def invalidate_static_sums(self):
        """The kept own-object density sums no longer match the particles:
        the next pass A sums them again.  `sums_epoch` counts these events,
        so a SceneBatch holding a device copy of the flag rebuilds it."""
        self._struct.flags &= ~_lib.TSPH_OBJ_STATIC_SUMS
        self.sums_epoch += 1

    def from_numpy(self, name, arr):
        """Write field `name` from an array in reference particle order.
        Writing positions leaves the storage out of cell order: the next
        step sorts the object again (for the static boundary too)."""
        self.invalidate_static_sums()
        if name in ("pos", "neighb_cell_seq", "cell"):
            self.cell_table_ok = False
        pid = self.t["pid"][: self.part_num].cpu().numpy()
        a = np.asarray(arr)[pid]
        dst = self._storage(name)
        dst.copy_(torch.as_tensor(a, dtype=dst.dtype).to(self.device))
        if name == self.X_name:
            self.t["pos"][: self.part_num, 3] = self.t[self.X_name][: self.part_num]

    # ---------------------------------------------------------- scene builders
    def scene_add_help_centering(self, start_pos, end_pos, spacing):
        """sph_obj.py:102-107, with the same numpy dtypes (f32 lattice count,
        f64 padding)."""
        end_pos = np.array(end_pos, dtype=np.float32)
        start_pos = np.array(start_pos, dtype=np.float32)
        matrix_shape = ((end_pos - start_pos + np.float32(1e-7)) / np.float32(spacing)).astype(np.int32)
        padding = ((end_pos - start_pos).astype(np.float64) - matrix_shape.astype(np.float64) * spacing) / 2
        return matrix_shape, padding

    def scene_add_cube(self, start_pos, end_pos, volume_frac, vel, color, relaxing_factor):
        """sph_obj.py:110-114."""
        spacing = self.config.part_size[1] * relaxing_factor
        matrix_shape, padding = self.scene_add_help_centering(start_pos, end_pos, spacing)
        self.push_matrix(np.ones(matrix_shape, dtype=np.bool_),
                         np.asarray(start_pos, dtype=np.float64) + padding, spacing,
                         volume_frac, vel, color)

    def scene_add_box(self, start_pos, end_pos, layers, volume_frac, vel, color, relaxing_factor):
        """sph_obj.py:140-151: hollow box with `layers` layers."""
        spacing = self.config.part_size[1] * relaxing_factor
        matrix_shape, padding = self.scene_add_help_centering(start_pos, end_pos, spacing)
        box = np.ones(matrix_shape, dtype=np.bool_)
        if len(matrix_shape) == 2:
            box[layers: matrix_shape[0] - layers, layers: matrix_shape[1] - layers] = False
        elif len(matrix_shape) == 3:
            box[layers: matrix_shape[0] - layers, layers: matrix_shape[1] - layers,
                layers: matrix_shape[2] - layers] = False
        else:
            raise Exception("scenario error: can only add 2D or 3D boxes")
        self.push_matrix(box, np.asarray(start_pos, dtype=np.float64) + padding, spacing,
                         volume_frac, vel, color)

    def scene_add_from_inlet(self, center, size, norm, speed, volume_frac, color, relaxing_factor):
        """sph_obj.py:117-137: one sheet of particles on the inlet's plane.
        u lies in the x-z plane orthogonal to `norm`, v = normalize(u x norm);
        the sheet is size[0] along u by size[1] along v, centred on `center`,
        moving with velocity norm * speed.  3D only."""
        spacing = self.config.part_size[1] * relaxing_factor
        matrix_shape, padding = self.scene_add_help_centering([0] * len(size), size, spacing)
        if len(matrix_shape) != 2:
            raise Exception("scenario ERROR: can only add 3D inlets.")
        norm = list(norm)
        if norm[0] == 0:
            u = np.array([1, 0, 0])
        else:
            u = _np_normalize(np.array([-norm[2] / norm[0], 0, 1]))
        v = _np_normalize(np.cross(u, norm))
        size = np.array(size) - padding
        start = np.array(center) - (size[0] * u + size[1] * v) / 2
        seq = [start + (float(i) * u + float(j) * v) * spacing
               for i in range(matrix_shape[0]) for j in range(matrix_shape[1])]
        pos_seq = np.array(seq).reshape(-1, 3)
        vel = np.array(norm) * speed
        self.push_part_seq(len(pos_seq), pos_seq, volume_frac, vel, color)

    def push_part_from_ply(self, p_sum, pos_seq, volume_frac, vel, color):
        """sph_obj.py:153-154."""
        self.push_part_seq(p_sum, pos_seq, volume_frac, vel, color)

    def push_matrix(self, matrix, start_position, spacing, volume_frac, vel, color):
        """sph_obj.py:158-163: np.where order (last axis fastest), f64 positions."""
        if len(matrix.shape) != self.config.dim:
            raise Exception("scenario error: wrong object dimension")
        index = np.where(matrix == True)  # noqa: E712
        pos_seq = np.stack(index, axis=1) * spacing + start_position
        self.push_part_seq(len(pos_seq), pos_seq, volume_frac, vel, color)

    def push_part_seq(self, pushed_part_num, pos_seq, volume_frac, vel, color=0):
        """sph_obj.py:165-180: append particles, then recompute rest_density
        and mass of every particle of the object."""
        n0 = self.part_num
        n1 = n0 + int(pushed_part_num)
        if n1 > self.max_part_num:
            raise Exception(f"particle capacity exceeded: {n1} > max_part_num {self.max_part_num}")
        d = self.config.dim
        P = self.config.phase_num
        f = self.np_dtype
        pos = np.zeros((pushed_part_num, 4), dtype=f)
        pos[:, :d] = np.asarray(pos_seq, dtype=np.float64)[:, :d].astype(np.float32).astype(f)
        vf = np.asarray(volume_frac, dtype=np.float64).astype(np.float32).astype(f)
        v = np.zeros(4, dtype=f)
        v[:d] = np.asarray(vel, dtype=np.float64).astype(np.float32)
        if len(vf) != P:
            raise Exception("volume_frac must have phase_num entries")
        sl = slice(n0, n1)
        self.t["pos"][sl] = torch.as_tensor(pos, device=self.device)
        self.t["vel"][sl] = torch.as_tensor(np.tile(v, (pushed_part_num, 1)), device=self.device)
        self.t["volume_frac"][sl] = torch.as_tensor(np.tile(vf, (pushed_part_num, 1)), device=self.device)
        self.t["rest_volume"][sl] = float(np.float32(self.config.part_size[d]))
        # new particles take the next reference indices
        self.t["pid"][sl] = torch.arange(n0, n1, dtype=torch.int32, device=self.device)
        self.color += [color] * int(pushed_part_num)
        self.part_num = n1
        self.layout_epoch += 1
        self.cell_table_ok = False  # the next step sorts the object again
        self.invalidate_static_sums()
        self.update_mass_all()

    def update_mass_all(self):
        """rest_density = phase_rest_density . volume_frac (sequential sum),
        mass = rest_density * rest_volume (sph_obj.py:178-180)."""
        n = self.part_num
        P = self.config.phase_num
        f32 = torch.float32
        vf = self.t["volume_frac"][:n].to(f32)
        rho = [torch.tensor(r, dtype=f32, device=self.device) for r in self.config.phase_rest_density]
        rd = rho[0] * vf[:, 0]
        for k in range(1, P):
            rd = rd + rho[k] * vf[:, k]
        m = rd * self.t["rest_volume"][:n].to(f32)
        self.t["rest_density"][:n] = rd.to(self.dtype)
        self.t["mass"][:n] = m.to(self.dtype)
        self.t["pos"][:n, 3] = self.t[self.X_name][:n]


class _BoundNgrid:
    """`Ngrid()` in the reference's form (main.py:8): the grid of the
    Simulation of the bound Config's objects, resolved on first use (the
    objects are created after the grid in main.py:8-10)."""

    @property
    def sim(self):
        from . import binding
        return binding.simulation()

    def __getattr__(self, name):
        return getattr(self.sim.ngrid, name)


class Ngrid:
    """The reference's shared uniform grid (sph_obj.py:280-323) over the
    objects' current storage order.  clear_node / encode / mem_shift /
    fill_node keep the reference call sequence; the grid is materialised by
    one library call when the last encoded object has been filled.
    `Ngrid()` (the reference's form) gives the grid of the Config bound by
    tichi_sph_amd.use_config."""

    def __new__(cls, sim=None):
        if sim is None:
            return _BoundNgrid()
        return super().__new__(cls)

    def __init__(self, sim):
        self.sim = sim
        cfg = sim.config
        dev = sim.device
        cap = max(sim.fluid.max_part_num, 1) + max(sim.bound.max_part_num, 1)
        self.node_part_count = torch.zeros(cfg.node_num, dtype=torch.int32, device=dev)
        self.node_part_shift = torch.zeros(cfg.node_num, dtype=torch.int32, device=dev)
        self.node_part_shift_count = torch.zeros(cfg.node_num, dtype=torch.int32, device=dev)
        self.part_pid_in_node = torch.zeros(cap, dtype=torch.int32, device=dev)
        self.part_uid_in_node = torch.zeros(cap, dtype=torch.int32, device=dev)
        self._encoded = []
        self._filled = []

    def clear_node(self):
        self._encoded = []
        self._filled = []

    def encode(self, obj):
        self._encoded.append(obj)

    def mem_shift(self):
        pass

    def fill_node(self, obj):
        self._filled.append(obj)
        if len(self._filled) == len(self._encoded):
            self.build(*self._encoded)

    def build(self, *objs):
        objs = sorted(objs, key=lambda o: o.uid)
        sim = self.sim
        arr = (ctypes.POINTER(_lib.TsphObject) * len(objs))(*[ctypes.pointer(o.struct()) for o in objs])
        rc = sim.lib.tsph_ngrid_build(
            ctypes.byref(sim.params), arr, len(objs), self.node_part_count.data_ptr(),
            self.node_part_shift.data_ptr(), self.node_part_shift_count.data_ptr(),
            self.part_pid_in_node.data_ptr(), self.part_uid_in_node.data_ptr(),
            sim.ws.data_ptr(), sim.ws.numel(), sim.stream_ptr())
        _lib.check(rc, "tsph_ngrid_build")

    def ptrs(self):
        return (self.node_part_count.data_ptr(), self.node_part_shift.data_ptr(),
                self.part_pid_in_node.data_ptr(), self.part_uid_in_node.data_ptr())
