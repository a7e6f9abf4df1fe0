"""The fused step's neighbour lists, slot by slot.  After the sort and pass A
(tsph_density_alpha), every fluid particle's fluid and boundary lists must
hold its in-range neighbours (r^2 <= r2_max and r > 0, the FP32 test of the
kernel) in the order of the reference's walk: the 3^d cells around it in
template order (x fastest, sph_config.py:219-224), each cell's particles in
storage order (sph.py:9-19, NeighborSearch.py:94-132).  The expected lists
are rebuilt here with numpy in FP32.

Pass A stages the first list slots of each particle in LDS and writes them
row by row at the end of its walk, appending later slots directly, so an
entry in the wrong slot, a lost or a repeated entry fails here.  The test
runs at the default capacities (64 / 48) and at 24 / 4, where slots past the
stage are appended directly and some lists overflow (the stored count is
then capacity + 1 and the first `capacity` slots are kept)."""
import ctypes

import numpy as np
import pytest
import torch

from conftest import load_scene

pytestmark = pytest.mark.gpu


def _i32(v):
    return ((v + 2 ** 31) % 2 ** 32) - 2 ** 31


def expected_lists(sim, fpos, opos, ostart):
    """Per fluid particle, the in-range particles of the object with
    positions `opos` (cell table `ostart`), in walk order."""
    p = sim.params
    d = sim.config.dim
    f32 = np.float32
    lb = [f32(p.lb[k]) for k in range(3)]
    h, r2max = f32(p.kernel_h[1]), f32(p.r2_max)
    coder = [int(p.coder[k]) for k in range(3)]
    R, node_num = int(p.neighb_range), int(p.node_num)
    zr = R if d == 3 else 0
    out = []
    for i in range(len(fpos)):
        x = fpos[i]
        c = [int(np.floor((x[k] - lb[k]) / h)) if k < d else 0 for k in range(3)]
        lst = []
        for dz in range(-zr, zr + 1):
            for dy in range(-R, R + 1):
                c0 = coder[0] * (c[0] - R) + coder[1] * (c[1] + dy)
                if d == 3:
                    c0 += coder[2] * (c[2] + dz)
                c0 = _i32(c0)
                lo, hi = max(c0, 1), min(c0 + 2 * R, node_num - 1)
                if lo > hi:
                    continue
                js = np.arange(ostart[lo], ostart[hi + 1])
                q = opos[js]
                dx, dyy, dzz = x[0] - q[:, 0], x[1] - q[:, 1], x[2] - q[:, 2]
                r2 = (dx * dx + dyy * dyy) + dzz * dzz  # FP32, no contraction
                lst.extend(js[(r2 <= r2max) & (r2 > f32(0))].tolist())
        out.append(lst)
    return out


def _check(lists, exp, counts, capacity, what):
    bad = 0
    for i, e in enumerate(exp):
        stored = len(e) if len(e) <= capacity else capacity + 1
        m = min(len(e), capacity)
        if counts[i] != stored or not np.array_equal(lists[:m, i], np.asarray(e[:m], np.int32)):
            bad += 1
    assert bad == 0, f"{what}: {bad} particles' lists differ"


@pytest.mark.parametrize("caps", [(64, 48), (24, 4)], ids=["default", "small"])
@pytest.mark.parametrize("name", ["dambreak_3d_8k", "dambreak_2d_8k"])
def test_pass_a_lists_match_the_reference_walk(name, caps):
    from tichi_sph_amd import _lib, simulation_from_dicts
    capf, capb = caps
    cb, sb = load_scene(name)
    print(f"[case] {name} list capacities {capf} / {capb}", flush=True)
    sim = simulation_from_dicts(cb, sb, precision=4, list_cap_f=capf, list_cap_b=capb)
    P = ctypes.byref(sim.params)
    ws, wsb = sim._ws()
    st = sim.stream_ptr()
    if sim.bound_needs_sort():
        sim._sort_bound()
    _lib.check(sim.lib.tsph_sort_object(P, ctypes.byref(sim.fluid.struct()), sim.ctrl_ptr(), 0,
                                        ws, wsb, st), "tsph_sort_object")
    sim.fluid.adopt_swaps()
    f, b = sim.fluid, sim.bound
    _lib.check(sim.lib.tsph_density_alpha(P, ctypes.byref(f.struct()), ctypes.byref(b.struct()),
                                          sim.ctrl_ptr(), ws, wsb, st), "tsph_density_alpha")
    torch.cuda.synchronize()
    n, nb = f.part_num, b.part_num
    fpos = f.t["pos"][:n, :3].cpu().numpy().astype(np.float32)
    bpos = b.t["pos"][:nb, :3].cpu().numpy().astype(np.float32)
    fstart = f.t["cell_start"].cpu().numpy()
    bstart = b.t["cell_start"].cpu().numpy()
    cap = f.t["nbr_f"].numel() // capf
    lf = f.t["nbr_f"].view(capf, cap)[:, :n].cpu().numpy()
    lb_ = f.t["nbr_b"].view(capb, cap)[:, :n].cpu().numpy()
    nf = f.t["nbr_nf"][:n].cpu().numpy()
    nbb = f.t["nbr_nb"][:n].cpu().numpy()
    ef = expected_lists(sim, fpos, fpos, fstart)
    eb = expected_lists(sim, fpos, bpos, bstart)
    lens = np.array([len(e) for e in ef])
    print(f"  {n} fluid particles; fluid neighbours {lens.min()}-{lens.max()} (mean {lens.mean():.1f}); "
          f"over capacity: {int((lens > capf).sum())} fluid, "
          f"{int((np.array([len(e) for e in eb]) > capb).sum())} boundary lists", flush=True)
    if capf == 24 and sim.config.dim == 3:
        assert (lens > 22).any() and (lens > capf).any()  # direct appends and overflow both occur
    _check(lf, ef, nf, capf, "fluid lists")
    _check(lb_, eb, nbb, capb, "boundary lists")
