"""tsph_sort_object_static (the sort without a workspace, its scratch in a
file-scope device array) against tsph_sort_object (the same sort on the
caller's workspace), through the C ABI: two calls in a row on one device,
the second on a smaller object, each bit-identical in every output (the
permuted state, pid, cell keys, the cell table and the CFL dt).  Sizes
beyond the static scratch are refused before anything is launched."""
import ctypes

import numpy as np
import pytest
import torch

from conftest import load_scene

pytestmark = pytest.mark.gpu

FIELDS = ("pos", "vel", "mass", "rest_density", "rest_volume", "volume_frac", "pid", "cell")


def _sim(name):
    from tichi_sph_amd import simulation_from_dicts
    cb, sb = load_scene(name)
    return simulation_from_dicts(cb, sb)


def _sort(sim, obj, cfl, static):
    from tichi_sph_amd import _lib
    P = ctypes.byref(sim.params)
    if static:
        rc = sim.lib.tsph_sort_object_static(P, ctypes.byref(obj.struct()), sim.ctrl_ptr(), cfl,
                                             sim.stream_ptr())
    else:
        ws, wsb = sim._ws()
        rc = sim.lib.tsph_sort_object(P, ctypes.byref(obj.struct()), sim.ctrl_ptr(), cfl, ws, wsb,
                                      sim.stream_ptr())
    _lib.check(rc, "static sort" if static else "sort")
    obj.adopt_swaps()


def _same(a, b, label):
    n = a.part_num
    assert n == b.part_num
    for f in FIELDS:
        x, y = a.t[f][:n].cpu().numpy(), b.t[f][:n].cpu().numpy()
        assert np.array_equal(x, y), (label, f)
    for f in ("cell_start", "cell_count"):
        assert np.array_equal(a.t[f].cpu().numpy(), b.t[f].cpu().numpy()), (label, f)


def test_static_sort_matches_workspace_sort_on_two_calls():
    from tichi_sph_amd import _lib
    lib = _lib.load()
    print(f"[case] static sort: scratch {lib.tsph_sort_static_scratch_bytes()} bytes, "
          f"max {lib.tsph_sort_static_max_particles()} particles / {lib.tsph_sort_static_max_cells()} cells",
          flush=True)
    # first the larger object (8000 fluid particles, CFL on), then a smaller
    # one (4096) that reuses the scratch the first call left behind
    for call, (name, cfl) in enumerate((("dambreak_3d_8k", 1), ("two_phase_3d_4k", 1))):
        a, b = _sim(name), _sim(name)
        n, d = a.fluid.part_num, a.config.dim
        rng = np.random.default_rng(call)
        v = rng.normal(0, 2.0, (n, d)).astype(np.float32)
        for s in (a, b):
            s.fluid.from_numpy("vel", v)
            s.set_dt(0.0)
        print(f"  call {call}: {name}, {n} fluid + {a.bound.part_num} boundary particles", flush=True)
        for obj_a, obj_b, c in ((a.bound, b.bound, 0), (a.fluid, b.fluid, cfl)):
            _sort(a, obj_a, c, static=False)
            _sort(b, obj_b, c, static=True)
        torch.cuda.synchronize()
        _same(a.bound, b.bound, f"call {call} boundary")
        _same(a.fluid, b.fluid, f"call {call} fluid")
        assert a.control().dt == b.control().dt > 0, call
        # sorting the sorted object again moves nothing
        _sort(b, b.fluid, 0, static=True)
        torch.cuda.synchronize()
        _same(a.fluid, b.fluid, f"call {call} fluid re-sorted")


def test_static_sort_refuses_oversized_objects_without_launching():
    from tichi_sph_amd import _lib
    lib = _lib.load()
    sim = _sim("dambreak_2d_8k")
    big = _lib.TsphObject()
    big.n = big.capacity = lib.tsph_sort_static_max_particles() + 1
    rc = lib.tsph_sort_object_static(ctypes.byref(sim.params), ctypes.byref(big), sim.ctrl_ptr(), 0,
                                     sim.stream_ptr())
    assert rc == -2  # TSPH_ECAPACITY
    assert b"maximum" in lib.tsph_last_error()
