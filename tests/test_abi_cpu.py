"""C-ABI checks that need no GPU: the HIP library loads, exports every entry
point include/tichi_sph.h declares, and the ctypes mirrors have the header's
exact layout.  Also host-side config behaviour (missing keys raise like
sph_config.read_param)."""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT, load_scene

HDR = os.path.join(ROOT, "include", "tichi_sph.h")


def header_functions():
    txt = open(HDR).read()
    return sorted(set(re.findall(r"\b(tsph_[a-z_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from tichi_sph_amd import _lib
    lib = _lib.load()
    decl = header_functions()
    assert set(decl) == set(_lib.EXPORTED)
    for name in decl:
        assert hasattr(lib, name), name
    assert lib.tsph_abi_version() == _lib.ABI_VERSION == 5


def test_struct_layout_matches_header(tmp_path):
    from tichi_sph_amd import _lib
    fields = {
        "TsphParams": ["dim", "list_cap_f", "grid_size", "cs", "list_cap_b", "lb", "phase_rest_density",
                       "part_size", "kernel_sig", "r2_max", "cfl_factor", "compression_threshold",
                       "phase_viscosity", "mu_per_phase"],
        "TsphObject": ["n", "flags", "pos", "pressure", "pid", "cell_structured", "cell_start", "nbr_nb", "pid_b",
                       "static_sums"],
        "TsphControl": ["dt", "it", "stop", "general_flag", "ke", "run_u", "err_div", "err_incom", "signal", "token",
                        "overflow_f", "counter2"],
        "TsphAdjoint": ["xbar", "vbar_out", "nbar", "a1bar", "kappa_next", "tape_normal", "grad", "bmom",
                        "bflag"],
    }
    src = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HDR}"', "int main(void){"]
    for s, fl in fields.items():
        src.append(f'printf("{s} %zu\\n", sizeof({s}));')
        for f in fl:
            src.append(f'printf("{s}.{f} %zu\\n", offsetof({s}, {f}));')
    src.append("return 0;}")
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-std=c11", str(c), "-o", str(exe)])
    out = dict(line.split() for line in subprocess.check_output([str(exe)]).decode().splitlines())
    for s, fl in fields.items():
        cls = getattr(_lib, s)
        assert int(out[s]) == ctypes.sizeof(cls), s
        for f in fl:
            assert int(out[f"{s}.{f}"]) == getattr(cls, f).offset, (s, f)


def test_missing_key_raises_like_reference():
    from tichi_sph_amd.config import Config
    cb, sb = load_scene("dambreak_2d_8k")
    cb = dict(cb)
    cb["solver_cs"] = None
    with pytest.raises(Exception, match="solver_cs is missing"):
        Config(cb, sb)


def test_reference_shipped_config_and_scenarios_parse():
    """The reference's own config/scenario files drop in (keys and derived
    constants).  Their copies live in tests/golden as input fixtures."""
    import json
    import numpy as np
    from tichi_sph_amd.config import Config
    g = os.path.join(ROOT, "tests", "golden")
    cb = json.load(open(os.path.join(g, "ref_config.json")))
    for name in ("ref_2d_dambreak.json", "ref_3d_dambreak.json", "ref_2d_two_phases.json",
                 "ref_3d_two_phases.json"):
        sb = json.load(open(os.path.join(g, name)))
        c = Config(cb, sb)
        dims = [int(np.ceil((np.float32(sb["sim_env"]["sim_space_rt"][d]) - np.float32(sb["sim_env"]["sim_space_lb"][d]))
                            / np.float32(c.kernel_h[1]))) for d in range(c.dim)]
        assert c.node_num == int(np.prod(dims))
        assert c.kernel_h[1] == float(__import__("numpy").float32(sb["sim_env"]["global_part_size"] * 2))


def test_build_recipe_flags():
    txt = open(os.path.join(ROOT, "tichi_sph_amd", "csrc", "build.sh")).read()
    assert "--offload-arch=gfx950" in txt
    assert "-packed-fp32-ops" in txt
    assert "fast-math" not in txt


def test_batch_table_sizing_and_argument_checks():
    import ctypes
    from tichi_sph_amd import _lib
    lib = _lib.load()
    b1, b64 = lib.tsph_batch_table_bytes(4, 1), lib.tsph_batch_table_bytes(4, 64)
    assert 0 < b1 < b64 and lib.tsph_batch_table_bytes(8, 64) > b64
    assert lib.tsph_batch_table_bytes(4, 0) == 0
    info = _lib.TsphBatchInfo()
    rc = lib.tsph_batch_prepare(None, None, None, None, None, None, 0, None, None, 0,
                                ctypes.byref(info), None)
    assert rc == -1 and b"null argument" in lib.tsph_last_error()
    rc = lib.tsph_batch_density_alpha(ctypes.byref(info), None, 0, None)
    assert rc == -1



def test_adjoint_boundary_sums_argument_checks():
    """The entries that read or write the boundary sums refuse an adjoint
    without their buffers before any device work (host-only scene)."""
    import ctypes
    from conftest import load_scene
    from tichi_sph_amd import _lib, simulation_from_dicts
    lib = _lib.load()
    cb, sb = load_scene("dambreak_3d_8k")
    sim = simulation_from_dicts(cb, sb, device="cpu", cfl=False)
    P = ctypes.byref(sim.params)
    fs, bs = ctypes.byref(sim.fluid.struct()), ctypes.byref(sim.bound.struct())
    adj = _lib.TsphAdjoint()  # every pointer NULL
    for name, args in (("tsph_adj_bound_sums", (P, fs, bs, ctypes.byref(adj), None)),
                       ("tsph_adj_density_alpha", (P, fs, bs, ctypes.byref(adj), None)),
                       ("tsph_adj_solver_iteration", (P, fs, bs, None, ctypes.byref(adj), None, 0, None))):
        rc = getattr(lib, name)(*args)
        assert rc == -1, name  # TSPH_EINVAL
        assert b"null boundary sums" in lib.tsph_last_error(), name
