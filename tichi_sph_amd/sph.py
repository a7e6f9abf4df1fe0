"""The reference's kernel surface (sph.py): one function per Taichi kernel,
with the reference's signatures, each enqueuing the matching HIP kernel
through tsph_ref_kernel:

    f(ngrid, obj, nobj)    the neighbour kernels (sph.py:50-97, 114-193, ...)
    f(obj)                 the per-particle kernels (sph.py:22-47, 100-111, ...)

`ngrid` is the reference-layout grid of the objects' Simulation (build it with
ngrid.clear_node / encode / mem_shift / fill_node first, as main.py:178-183
does); the Simulation itself is accepted in its place, and as a leading
argument of the per-particle kernels, f(sim, obj).
"""
import ctypes

from . import _lib


def _sim_of(x):
    """The Simulation behind an Ngrid (or a Simulation given directly)."""
    return x if hasattr(x, "ngrid") and hasattr(x, "params") else x.sim


def _run(sim, name, obj, nobj=None):
    obj.invalidate_static_sums()  # a reference kernel may write obj's positions or masses
    P = ctypes.byref(sim.params)
    ws = ctypes.c_void_p(sim.ws.data_ptr())
    npc, nps, pin, uin = sim.ngrid.ptrs()
    rc = sim.lib.tsph_ref_kernel(P, _lib.KERNEL_IDS[name], ctypes.byref(obj.struct()),
                                 ctypes.byref(nobj.struct()) if nobj is not None else None,
                                 npc, nps, pin, uin, sim.ctrl_ptr(), ws, sim.ws.numel(),
                                 sim.stream_ptr())
    _lib.check(rc, name)


def _make_obj_kernel(name):
    def fn(*args):
        if len(args) == 1:  # f(obj), the reference's form
            obj = args[0]
            sim = obj.sim
        elif len(args) == 2:  # f(sim, obj)
            sim, obj = _sim_of(args[0]), args[1]
        else:
            raise TypeError(f"{name}(obj) takes one particle object")
        _run(sim, name, obj)
    fn.__name__ = name
    fn.__doc__ = f"{name}(obj): reference kernel of the same name (sph.py)."
    return fn


def _make_pair_kernel(name):
    def fn(ngrid, obj, nobj):
        _run(_sim_of(ngrid), name, obj, nobj)
    fn.__name__ = name
    fn.__doc__ = f"{name}(ngrid, obj, nobj): reference kernel of the same name (sph.py)."
    return fn


_OBJ = ["SPH_clean_value", "cfl_condition", "SPH_prepare_alpha", "SPH_advection_gravity_acc",
        "WC_pressure_val", "IPPE_adv_psi_init", "IPPE_psi_adv_non_negative",
        "SPH_advection_update_vel_adv", "SPH_vel_2_vel_adv", "SPH_vel_adv_2_vel", "SPH_update_pos",
        "SPH_update_energy", "SPH_update_mass", "SPH_FBM_clean_tmp", "SPH_FBM_check_tmp",
        "SPH_update_volume_frac"]
_PAIR = ["SPH_prepare_attr", "SPH_prepare_alpha_1", "SPH_prepare_alpha_2",
         "SPH_advection_viscosity_acc", "SPH_advection_surface_tension_acc", "WC_pressure_acce",
         "IPPE_adv_psi", "IPPE_update_vel_adv", "SPH_FBM_diffuse", "SPH_FBM_convect"]

for _n in _OBJ:
    globals()[_n] = _make_obj_kernel(_n)
for _n in _PAIR:
    globals()[_n] = _make_pair_kernel(_n)

__all__ = _OBJ + _PAIR


def map_velocity(ngrid, grid, nobj):
    """map_velocity(ngrid, grid, nobj), sph.py:382-397: SPH interpolation of
    nobj's velocity onto the grid's nodes (grid.vel).  nobj's cell table must
    be this step's (after Simulation.sph_step), like the reference's Ngrid."""
    sim = _sim_of(ngrid)
    if not getattr(nobj, "cell_table_ok", False):
        raise _lib.TsphError("map_velocity: the object's cell table is stale (run Simulation.sph_step first)")
    shape = (ctypes.c_int32 * 3)(*(list(grid.shape) + [1] * (3 - len(grid.shape))))
    sim._call("map_velocity", sim.lib.tsph_map_velocity, ctypes.byref(sim.params), ctypes.byref(nobj.struct()),
              shape, float(grid.dist), ctypes.c_void_p(grid.vel.data_ptr()), sim.stream_ptr())


__all__.append("map_velocity")

