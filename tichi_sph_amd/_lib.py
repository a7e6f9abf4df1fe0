"""ctypes binding of libtichi_sph.so (the C ABI in include/tichi_sph.h).

The library is loaded from the package directory.  There is no fallback:
if the shared object is missing, every solver call raises, so a GPU run can
never silently compute on the CPU.
"""
import ctypes
import os

TSPH_MAX_PHASE = 8
TSPH_MAX_ITER_LOG = 256
ABI_VERSION = 5  # TSPH_ABI_VERSION of include/tichi_sph.h
TSPH_FLAG_FAST_QUOTIENT = 1  # TsphParams.flags
TSPH_FLAG_ST_RECORDS = 2
TSPH_OBJ_STATIC_SUMS = 1  # TsphObject.flags
TSPH_DFSPH = 0
TSPH_VFSPH = 1

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtichi_sph.so")

_i32 = ctypes.c_int32
_f64 = ctypes.c_double
_vp = ctypes.c_void_p
_ip = ctypes.POINTER(ctypes.c_int32)


class TsphParams(ctypes.Structure):
    _fields_ = [
        ("dim", _i32), ("phase_num", _i32), ("solver_type", _i32), ("precision", _i32),
        ("neighb_range", _i32), ("ntemp", _i32), ("node_num", _i32), ("list_cap_f", _i32),
        ("grid_size", _i32 * 4), ("coder", _i32 * 4),
        ("cs", _i32), ("wc_gamma", _i32), ("iter_min", _i32), ("iter_max", _i32),
        ("list_cap_b", _i32), ("flags", _i32), ("pad0", _i32), ("pad1", _i32),
        ("lb", _f64 * 4), ("rt", _f64 * 4), ("gravity", _f64 * 4),
        ("phase_rest_density", _f64 * TSPH_MAX_PHASE),
        ("part_size", _f64 * 5), ("kernel_h", _f64 * 5), ("kernel_sig", _f64 * 4),
        ("r2_max", _f64),
        ("dynamic_viscosity", _f64), ("surface_tension_gamma", _f64), ("fbm_diffusion", _f64),
        ("fbm_convection", _f64), ("cfl_factor", _f64),
        ("divergence_threshold", _f64), ("compression_threshold", _f64),
        ("phase_viscosity", _f64 * TSPH_MAX_PHASE), ("mu_per_phase", _i32), ("pad2", _i32),
    ]


_OBJ_PTRS = [
    "pos", "vel", "vel_adv", "acce_adv", "acce", "normal", "alpha_1",
    "mass", "rest_density", "rest_volume",
    "volume_frac", "volume_frac_tmp", "drift_vel",
    "W", "sph_compression", "sph_density", "psi_adv", "alpha", "alpha_2",
    "kappa", "vol", "fbm_zeta", "pressure",
    "pid", "flag", "cell", "cell_structured", "in_cell", "cell_count", "cell_start",
    "nbr_f", "nbr_b", "nbr_nf", "nbr_nb",
    "pos_b", "vel_b", "mass_b", "rest_density_b", "rest_volume_b", "volume_frac_b", "pid_b",
    "static_sums",
]


class TsphObject(ctypes.Structure):
    _fields_ = [("n", _i32), ("capacity", _i32), ("uid", _i32), ("flags", _i32)] + [
        (name, _vp) for name in _OBJ_PTRS
    ]


class TsphControl(ctypes.Structure):
    _fields_ = [
        ("dt", _f64), ("compression", _f64),
        ("it", _i32), ("stop", _i32), ("counter", _i32), ("general_flag", _i32),
        ("div_iter", _i32), ("incom_iter", _i32), ("fbm_rounds", _i32), ("solve", _i32),
        ("ke", _f64), ("pe", _f64),
        ("run_u", _i32 * TSPH_MAX_ITER_LOG),
        ("err_div", _f64 * TSPH_MAX_ITER_LOG),
        ("err_incom", _f64 * TSPH_MAX_ITER_LOG),
        ("overflow_f", _i32), ("overflow_b", _i32), ("counter2", _i32), ("pad", _i32),
        ("signal", ctypes.c_uint64), ("token", _i32), ("pad2", _i32),
    ]


# Reference kernel ids (sph.py), include/tichi_sph.h
_ADJ_PTRS = ["xbar", "vbar", "vbar_out", "rhobar", "alphabar", "nbar", "a1bar",
             "tape_vel_adv", "kappa_next", "tape_vel", "tape_normal", "grad", "bmom", "bflag"]


class TsphAdjoint(ctypes.Structure):
    _fields_ = [(name, _vp) for name in _ADJ_PTRS]


class TsphBatchInfo(ctypes.Structure):
    _fields_ = [("nscenes", _i32), ("precision", _i32), ("tiles", _i32 * 4)]


KERNEL_IDS = {
    "SPH_clean_value": 0, "cfl_condition": 1, "SPH_prepare_attr": 2,
    "SPH_prepare_alpha_1": 3, "SPH_prepare_alpha_2": 4, "SPH_prepare_alpha": 5,
    "SPH_advection_gravity_acc": 6, "SPH_advection_viscosity_acc": 7,
    "SPH_advection_surface_tension_acc": 8, "WC_pressure_val": 9, "WC_pressure_acce": 10,
    "IPPE_adv_psi_init": 11, "IPPE_adv_psi": 12, "IPPE_psi_adv_non_negative": 13,
    "IPPE_update_vel_adv": 14, "SPH_advection_update_vel_adv": 15, "SPH_vel_2_vel_adv": 16,
    "SPH_vel_adv_2_vel": 17, "SPH_update_pos": 18, "SPH_update_energy": 19,
    "SPH_update_mass": 20, "SPH_FBM_clean_tmp": 21, "SPH_FBM_diffuse": 22,
    "SPH_FBM_convect": 23, "SPH_FBM_check_tmp": 24, "SPH_update_volume_frac": 25,
}

EXPORTED = [
    "tsph_last_error", "tsph_abi_version", "tsph_control_bytes", "tsph_workspace_bytes",
    "tsph_sort_object", "tsph_ngrid_build", "tsph_neighbor_lists", "tsph_density_alpha",
    "tsph_solver_begin", "tsph_solver_begin_signal", "tsph_signal_alloc", "tsph_signal_free",
    "tsph_verify_fast_quotient",
    "tsph_solver_iteration", "tsph_solver_psi", "tsph_solver_update", "tsph_advect",
    "tsph_fbm_round", "tsph_tail",
    "tsph_ref_kernel", "tsph_solver_psi_taped", "tsph_solver_psi_frozen", "tsph_relist", "tsph_adj_bound_sums", "tsph_adj_begin",
    "tsph_adj_solver_iteration", "tsph_adj_advect", "tsph_adj_density_alpha", "tsph_adj_end",
    "tsph_batch_table_bytes", "tsph_batch_prepare", "tsph_batch_sort", "tsph_batch_density_alpha",
    "tsph_batch_solver_begin", "tsph_batch_solver_begin_signal", "tsph_batch_solver_psi",
    "tsph_batch_solver_update",
    "tsph_batch_advect", "tsph_batch_tail", "tsph_batch_stop_state", "tsph_map_velocity",
    "tsph_wcsph", "tsph_sort_object_static", "tsph_sort_static_scratch_bytes",
    "tsph_sort_static_max_particles", "tsph_sort_static_max_cells",
]

_lib = None


class TsphError(RuntimeError):
    pass


def load(path=None):
    """Load the HIP library (torch must be imported first so that the HIP
    runtime bundled with torch is the one the library binds to)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    import torch  # noqa: F401  (binds the process to torch's HIP runtime)

    p = path or LIB_PATH
    if not os.path.exists(p):
        raise TsphError(
            f"{p} is missing: build it with tichi_sph_amd/csrc/build.sh "
            "(or __graft_entry__.build()); there is no CPU fallback")
    lib = ctypes.CDLL(p)
    lib.tsph_last_error.restype = ctypes.c_char_p
    lib.tsph_abi_version.restype = ctypes.c_int
    lib.tsph_control_bytes.restype = ctypes.c_size_t
    lib.tsph_workspace_bytes.restype = ctypes.c_size_t
    lib.tsph_workspace_bytes.argtypes = [ctypes.POINTER(TsphParams), _i32]
    P, O, C = ctypes.POINTER(TsphParams), ctypes.POINTER(TsphObject), _vp
    A = ctypes.POINTER(TsphAdjoint)
    BI = ctypes.POINTER(TsphBatchInfo)
    lib.tsph_batch_table_bytes.restype = ctypes.c_size_t
    lib.tsph_batch_table_bytes.argtypes = [_i32, _i32]
    lib.tsph_sort_static_scratch_bytes.restype = ctypes.c_size_t
    lib.tsph_sort_static_max_particles.restype = _i32
    lib.tsph_sort_static_max_cells.restype = _i32
    sz = ctypes.c_size_t
    sigs = {
        "tsph_sort_object": [P, O, C, _i32, _vp, sz, _vp],
        "tsph_sort_object_static": [P, O, C, _i32, _vp],
        "tsph_ngrid_build": [P, ctypes.POINTER(O), _i32, _vp, _vp, _vp, _vp, _vp, _vp, sz, _vp],
        "tsph_neighbor_lists": [P, O, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
        "tsph_density_alpha": [P, O, O, C, _vp, sz, _vp],
        "tsph_solver_begin": [P, C, _i32, _vp],
        "tsph_solver_begin_signal": [P, C, _i32, _vp, _i32, _vp],
        "tsph_signal_alloc": [ctypes.c_size_t, ctypes.POINTER(_vp), ctypes.POINTER(_vp)],
        "tsph_signal_free": [_vp],
        "tsph_verify_fast_quotient": [P, _vp, _vp],
        "tsph_solver_iteration": [P, O, O, C, _i32, _vp, sz, _vp],
        "tsph_solver_psi": [P, O, O, C, _i32, _vp, sz, _vp],
        "tsph_solver_update": [P, O, O, C, _i32, _vp, sz, _vp],
        "tsph_advect": [P, O, O, C, _vp, sz, _vp],
        "tsph_fbm_round": [P, O, C, _vp, sz, _vp],
        "tsph_tail": [P, O, C, _i32, _vp, sz, _vp],
        "tsph_ref_kernel": [P, _i32, O, O, _vp, _vp, _vp, _vp, C, _vp, sz, _vp],
        "tsph_solver_psi_taped": [P, O, O, C, _i32, _vp, _vp, _vp, sz, _vp],
        "tsph_solver_psi_frozen": [P, O, O, C, _i32, _vp, _vp, sz, _vp],
        "tsph_relist": [P, O, O, C, _vp, sz, _vp],
        "tsph_adj_begin": [P, O, C, A, _vp, _vp, _vp, _vp],
        "tsph_adj_solver_iteration": [P, O, O, C, A, _vp, sz, _vp],
        "tsph_adj_advect": [P, O, C, A, _vp, sz, _vp],
        "tsph_adj_density_alpha": [P, O, O, A, _vp],
        "tsph_adj_bound_sums": [P, O, O, A, _vp],
        "tsph_adj_end": [P, O, A, _vp, _vp, _vp, _vp],
        "tsph_batch_prepare": [P, O, O, _vp, _vp, _vp, _i32, _vp, _vp, sz, BI, _vp],
        "tsph_batch_sort": [BI, _vp, O, _i32, _i32, _vp],
        "tsph_batch_density_alpha": [BI, _vp, _i32, _vp],
        "tsph_batch_solver_begin": [BI, _vp, _i32, _vp],
        "tsph_batch_solver_begin_signal": [BI, _vp, _i32, _vp, _i32, _vp],
        "tsph_batch_solver_psi": [BI, _vp, _i32, _i32, _vp],
        "tsph_batch_solver_update": [BI, _vp, _i32, _i32, _vp],
        "tsph_batch_advect": [BI, _vp, _i32, _vp],
        "tsph_batch_tail": [BI, _vp, _i32, _vp],
        "tsph_batch_stop_state": [BI, _vp, _vp, _vp],
        "tsph_map_velocity": [P, O, _vp, _f64, _vp, _vp],
        "tsph_wcsph": [P, O, O, C, _vp],
    }
    for name, args in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    if lib.tsph_abi_version() != ABI_VERSION:
        raise TsphError("libtichi_sph.so ABI version mismatch")
    if lib.tsph_control_bytes() != ctypes.sizeof(TsphControl):
        raise TsphError("TsphControl layout mismatch between Python and the library")
    if path is None:
        _lib = lib
    return lib


def check(rc, what=""):
    if rc != 0:
        msg = _lib.tsph_last_error().decode() if _lib is not None else ""
        raise TsphError(f"{what} failed ({rc}): {msg}")
