"""ctypes mirror of include/huntian_abi.h and the loader of libhuntian_hip.so.

The struct layouts here must match the C header field for field; tests check the sizes
against the library (hnt_abi_version plus a struct-size probe is not exported, so the test
builds a tiny C program instead when a compiler is present).
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhuntian_hip.so")

HNT_OK = 0
HNT_ERR_INVALID = 1
HNT_ERR_DEVICE = 2
HNT_ERR_CAPACITY = 3
HNT_ERR_UNSUPPORTED = 4
HNT_ERR_ALLOC = 5

STAGES = ("pf_mark", "pf_compact", "sa_prep", "sa_kernel", "sa_cross", "pf_smooth", "sa_mask")

HNT_VARIANT_A = 0
HNT_VARIANT_B = 1
HNT_VARIANT_C = 2
VARIANTS = {"A": HNT_VARIANT_A, "B": HNT_VARIANT_B, "C": HNT_VARIANT_C}

_i32p = C.POINTER(C.c_int32)
_f32p = C.POINTER(C.c_float)
_f64p = C.POINTER(C.c_double)
_i8p = C.POINTER(C.c_int8)


class ShellTables(C.Structure):
    _fields_ = [
        ("rmax", C.c_int32),
        ("nshells", C.c_int32),
        ("ncells", C.c_int64),
        ("di", _i32p),
        ("dj", _i32p),
        ("dk", _i32p),
        ("shell_start", _i32p),
        ("shell_count", _i32p),
        ("shell_r2", _i32p),
    ]


class CTab(C.Structure):
    _fields_ = [
        ("nx", C.c_int32),
        ("ny", C.c_int32),
        ("nz", C.c_int32),
        ("x1", C.c_double),
        ("x2", C.c_double),
        ("y1", C.c_double),
        ("y2", C.c_double),
        ("z1", C.c_double),
        ("z2", C.c_double),
        ("out_val", C.c_double),
        ("table", _f32p),
    ]


class Peaks(C.Structure):
    _fields_ = [
        ("capacity", C.c_int64),
        ("i", _i32p),
        ("j", _i32p),
        ("k", _i32p),
        ("filt", _i32p),
        ("delta", _f32p),
        ("count", C.c_int64),
        ("total", C.c_int64),
    ]


class Fields(C.Structure):
    _fields_ = [
        ("n1", C.c_int32),
        ("n2", C.c_int32),
        ("n3", C.c_int32),
        ("delta", _f32p),
        ("etax", _f32p),
        ("etay", _f32p),
        ("etaz", _f32p),
        ("eta2x", _f32p),
        ("eta2y", _f32p),
        ("eta2z", _f32p),
        ("lapd", _f32p),
    ]


class Batch(C.Structure):
    _fields_ = [
        ("npeaks", C.c_int64),
        ("pi", _i32p),
        ("pj", _i32p),
        ("pk", _i32p),
        ("Rf", _f64p),
        ("ir2min", _i32p),
        ("ZZon", _f64p),
        ("fcrit", _f64p),
        ("ZZon_scalar", C.c_double),
        ("fcrit_scalar", C.c_double),
    ]


class FilterBank(C.Structure):
    _fields_ = [
        ("nfilt", C.c_int32),
        ("Rf", _f64p),
        ("fcrit", _f32p),
        ("wsmooth", C.c_int32),
        ("alatt", C.c_double),
    ]


class PkTable(C.Structure):
    _fields_ = [
        ("sqrt_pk", _f32p),
        ("n", C.c_int32),
        ("log_k_min", C.c_float),
        ("d_log_k", C.c_float),
    ]


class Transfer(C.Structure):
    _fields_ = [("id", C.c_int32), ("dim1", C.c_int32), ("dim2", C.c_int32)]


class CosmoTables(C.Structure):
    _fields_ = [
        ("ntab", C.c_int32),
        ("atab", _f64p),
        ("D_atab", _f64p),
        ("nchi", C.c_int32),
        ("chi2z_z", _f64p),
        ("chi2z_chi", _f64p),
        ("chi2z_zmax", C.c_double),
        ("dlnD_atab", _f64p),
    ]


class PackParams(C.Structure):
    _fields_ = [
        ("alatt", C.c_double),
        ("n", C.c_int32),
        ("xbx", C.c_double),
        ("ybx", C.c_double),
        ("zbx", C.c_double),
        ("ZZon", C.c_double),
        ("ievol", C.c_int32),
        ("obs", C.c_double * 3),
        ("z_max", C.c_double),
        ("ilpt", C.c_int32),
        ("Om", C.c_double),
        ("OL", C.c_double),
        ("ioutshear", C.c_int32),
    ]


MLP_INPUTS, MLP_HIDDEN, MLP_OUTPUTS = 18, 128, 3
MLP_NPARAM = 18 * 128 + 128 + 2 * (128 * 128 + 128) + 128 * 3 + 3
HNT_ORDER_CROSSING, HNT_ORDER_PLAIN, HNT_ORDER_LEARNED, HNT_ORDER_GIVEN = 0, 1, 2, 3
ORDERS = {"crossing": HNT_ORDER_CROSSING, "plain": HNT_ORDER_PLAIN,
          "learned": HNT_ORDER_LEARNED, "given": HNT_ORDER_GIVEN}
PRECISIONS = {"fp64": 0, "fp32_kahan": 1}
TUNING = {"brick_log2": 0, "waves_per_cu": 1, "gang": 2, "fft": 3, "fft_direct": 4,
          "tricubic_tab": 5}


class Mlp(C.Structure):
    _fields_ = [
        ("params", _f32p),
        ("norm", _f32p),
        ("adam_m", _f32p),
        ("adam_v", _f32p),
        ("step", C.c_int64),
    ]


class Params(C.Structure):
    _fields_ = [
        ("alatt", C.c_double),
        ("nbuff", C.c_int32),
        ("rmax2rs", C.c_double),
        ("variant", C.c_int32),
        ("early_exit", C.c_int32),
        ("order", C.c_int32),
        ("precision", C.c_int32),
        ("bracket", C.POINTER(Mlp)),
        ("extent", _i32p),
    ]


RESULT_SCALARS = ("RTHL", "Fbarx", "e_v", "p_v", "Srb", "d2F", "zvir_half")
RESULT_VEC3 = ("eigs", "Sbar", "Sbar2", "gradpk", "gradpkf", "gradpkrf")
RESULT_INT = ("nshells_used", "crossing_shell", "bracket")
# every Int32 field of hnt_results (schedule_extent is allocated only on request)
RESULT_I32 = RESULT_INT + ("schedule_extent",)


class Results(C.Structure):
    _fields_ = [
        ("RTHL", _f64p),
        ("Fbarx", _f64p),
        ("e_v", _f64p),
        ("p_v", _f64p),
        ("strain", _f64p),
        ("eigs", _f64p),
        ("Srb", _f64p),
        ("Sbar", _f64p),
        ("Sbar2", _f64p),
        ("gradpk", _f64p),
        ("gradpkf", _f64p),
        ("gradpkrf", _f64p),
        ("d2F", _f64p),
        ("zvir_half", _f64p),
        ("nshells_used", _i32p),
        ("crossing_shell", _i32p),
        ("bracket", _i32p),
        ("schedule_extent", _i32p),
    ]


def as_ptr(addr: int, ctype):
    """Raw address (device or host) -> typed ctypes pointer (NULL for 0/None)."""
    if not addr:
        return C.cast(None, C.POINTER(ctype))
    return C.cast(C.c_void_p(int(addr)), C.POINTER(ctype))


# Symbols the header declares, in header order (the loader test checks every one).
EXPORTED = (
    "hnt_ctx_create",
    "hnt_ctx_destroy",
    "hnt_ctx_release_scratch",
    "hnt_ctx_set_stream",
    "hnt_last_error",
    "hnt_abi_version",
    "hnt_ctx_set_tuning",
    "hnt_ctx_set_profiling",
    "hnt_ctx_get_profile",
    "hnt_shell_table_size",
    "hnt_build_shell_tables",
    "hnt_fsc_of_z",
    "hnt_peak_find",
    "hnt_peak_find_tile",
    "hnt_shell_analyse",
    "hnt_shell_analyse_noalloc",
    "hnt_shell_analyse_tiles",
    "hnt_plan_create",
    "hnt_plan_launch",
    "hnt_plan_destroy",
    "hnt_pack_records",
    "hnt_finalize_eulerian",
    "hnt_merge_catalog",
    "hnt_isolated_convolve",
    "hnt_compute_2lpt",
    "hnt_compute_laplacian",
    "hnt_isolated_poisson_psi",
    "hnt_isolated_short_pot",
    "hnt_interpolate_to_tile",
    "hnt_tile_noise",
    "hnt_downsample_noise",
    "hnt_extended_residual",
    "hnt_shell_gather_full",
    "hnt_eig3_symmetric_batch",
    "hnt_interp3_trilinear",
    "hnt_kernel_strain",
    "hnt_bracket_features",
    "hnt_mlp_forward",
    "hnt_mlp_gradient",
    "hnt_mlp_train_step",
)

_lib = None


class NativeLibraryMissing(RuntimeError):
    pass


def load(path: str | None = None):
    """Load libhuntian_hip.so (built in-tree); raise loudly if it is not there."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise NativeLibraryMissing(
            f"{p} is not built; run __graft_entry__.build() (hipcc --offload-arch=gfx950)")
    # One HIP runtime per process: PyTorch ships its own libamdhip64, and the library must
    # bind to that copy (same soname) rather than pull in the system ROCm one beside it.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(p)
    vp = C.c_void_p
    lib.hnt_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
    lib.hnt_ctx_destroy.argtypes = [vp]
    lib.hnt_ctx_release_scratch.argtypes = [vp]
    lib.hnt_ctx_set_stream.argtypes = [vp, vp]
    lib.hnt_last_error.argtypes = [C.c_char_p, C.c_size_t]
    lib.hnt_last_error.restype = C.c_size_t
    lib.hnt_abi_version.restype = C.c_int
    lib.hnt_ctx_set_tuning.argtypes = [vp, C.c_int32, C.c_int32]
    lib.hnt_ctx_set_profiling.argtypes = [vp, C.c_int]
    lib.hnt_ctx_get_profile.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int64),
                                        C.c_int32]
    lib.hnt_shell_table_size.argtypes = [C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_int32)]
    lib.hnt_build_shell_tables.argtypes = [C.c_int32] + [_i32p] * 6
    lib.hnt_fsc_of_z.argtypes = [C.POINTER(CTab), C.c_double]
    lib.hnt_fsc_of_z.restype = C.c_double
    lib.hnt_peak_find.argtypes = [vp, C.POINTER(vp), _f32p, C.c_int32, C.c_int32, C.c_int32,
                                  _i8p, C.POINTER(Peaks)]
    lib.hnt_peak_find_tile.argtypes = [vp, _f32p, C.c_int32, C.c_int32, C.POINTER(FilterBank),
                                       _i8p, C.POINTER(Peaks), _f32p, C.POINTER(vp)]
    lib.hnt_pack_records.argtypes = [vp, C.POINTER(Batch), C.POINTER(Results), _f32p, _f32p,
                                     C.POINTER(CosmoTables), C.POINTER(PackParams), _f32p,
                                     C.c_int64, C.POINTER(C.c_int64)]
    lib.hnt_finalize_eulerian.argtypes = [vp, _f32p, C.c_int64, C.c_int32,
                                          C.POINTER(CosmoTables), C.c_double, C.c_double,
                                          _f64p, C.c_int32, C.c_double, _f32p]
    lib.hnt_merge_catalog.argtypes = [vp, _f32p, C.c_int64, C.c_int32, C.c_int32, C.c_int32,
                                      _f32p, C.POINTER(C.c_int64)]
    fpp = C.POINTER(_f32p)
    lib.hnt_isolated_convolve.argtypes = [vp, _f32p, C.c_int32, C.c_int32, C.c_double,
                                          C.POINTER(PkTable), C.POINTER(Transfer), C.c_int32,
                                          C.c_int32, fpp]
    lib.hnt_compute_2lpt.argtypes = [vp, _f32p, C.c_int32, C.c_double, fpp]
    lib.hnt_compute_laplacian.argtypes = [vp, _f32p, C.c_int32, C.c_double, _f32p]
    lib.hnt_isolated_poisson_psi.argtypes = [vp, _f32p, C.c_int32, C.c_double, fpp]
    lib.hnt_isolated_short_pot.argtypes = [vp, _f32p, C.c_int32, C.c_double, C.c_double, _f32p]
    lib.hnt_interpolate_to_tile.argtypes = [vp, _f32p] + [C.c_int32] * 7 + [_f32p]
    lib.hnt_tile_noise.argtypes = [vp] + [C.c_int32] * 5 + [C.c_int64, C.c_uint64, _f32p]
    lib.hnt_downsample_noise.argtypes = [vp, C.c_int64, C.c_int32, C.c_uint64, _f32p]
    lib.hnt_extended_residual.argtypes = ([vp] + [C.c_int32] * 5 +
                                          [C.c_int64, C.c_uint64, _f32p, C.c_int32, C.c_int32,
                                           _f32p])
    lib.hnt_shell_analyse.argtypes = [vp, C.POINTER(Fields), C.POINTER(Batch),
                                      C.POINTER(ShellTables), C.POINTER(CTab),
                                      C.POINTER(Params), C.POINTER(Results), _i8p]
    lib.hnt_shell_analyse_noalloc.argtypes = lib.hnt_shell_analyse.argtypes
    lib.hnt_shell_gather_full.argtypes = [vp, C.POINTER(Fields), C.c_int64, _i32p, _i32p, _i32p,
                                          C.POINTER(ShellTables), _f64p, _i32p, _f64p, _f64p,
                                          _f64p, _f64p, _f64p, _f64p]
    lib.hnt_eig3_symmetric_batch.argtypes = [vp, _f64p, C.c_int64, _f64p]
    lib.hnt_interp3_trilinear.argtypes = [vp, C.POINTER(CTab), _f64p, _f64p, _f64p, C.c_int64,
                                          _f64p]
    lib.hnt_kernel_strain.argtypes = [vp, _f64p, _f64p, _f64p, _f64p, C.c_int32, C.c_int32,
                                      C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double,
                                      C.c_double, _f64p]
    lib.hnt_shell_analyse_tiles.argtypes = [vp, C.c_int32, C.POINTER(Fields), C.POINTER(Batch),
                                            C.POINTER(ShellTables), C.POINTER(CTab),
                                            C.POINTER(Params), C.POINTER(Results),
                                            C.POINTER(C.POINTER(C.c_int8))]
    lib.hnt_plan_create.argtypes = [vp, C.c_int32, C.POINTER(Fields), C.POINTER(Batch),
                                    C.POINTER(ShellTables), C.POINTER(CTab), C.POINTER(Params),
                                    C.POINTER(Results), C.POINTER(C.POINTER(C.c_int8)),
                                    C.POINTER(vp)]
    lib.hnt_plan_launch.argtypes = [vp]
    lib.hnt_plan_destroy.argtypes = [vp]
    lib.hnt_bracket_features.argtypes = [vp, C.POINTER(Fields), C.POINTER(Batch),
                                         C.POINTER(ShellTables), C.c_double, _f32p]
    lib.hnt_mlp_forward.argtypes = [vp, C.POINTER(Mlp), _f32p, C.c_int64, _f32p]
    lib.hnt_mlp_gradient.argtypes = [vp, C.POINTER(Mlp), _f32p, _f32p, C.c_int64, _f32p,
                                     C.POINTER(C.c_double)]
    lib.hnt_mlp_train_step.argtypes = [vp, C.POINTER(Mlp), _f32p, _f32p, C.c_int64, C.c_float,
                                       C.POINTER(C.c_double)]
    for name in EXPORTED:
        if name not in ("hnt_last_error", "hnt_abi_version", "hnt_fsc_of_z"):
            getattr(lib, name).restype = C.c_int
    if path is None:
        _lib = lib
    return lib


def last_error(lib) -> str:
    buf = C.create_string_buffer(1024)
    lib.hnt_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


class HuntianError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"huntian status {status}: {msg}")
        self.status = status


def check(lib, status: int):
    if status != HNT_OK:
        raise HuntianError(status, last_error(lib))
