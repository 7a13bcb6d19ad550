"""ctypes binding of lib/libtbk.so (the C ABI declared in include/tbk.h).

Loading fails loudly when the HIP library is missing: there is no CPU
fallback anywhere in the product path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libtbk.so")

TBK_OK, TBK_ERR_ARG, TBK_ERR_PAIRS, TBK_ERR_OVERLAP = 0, 1, 2, 3
TBK_ERR_CONVERGE, TBK_ERR_ZFILL, TBK_ERR_DEGEN = 4, 5, 6
TBK_ERR_HIP, TBK_ERR_NOMEM, TBK_ERR_NODEVICE = 10, 11, 12
TBK_FP64, TBK_MIXED = 0, 1
TBK_END = -1

# Every symbol include/tbk.h declares (checked by the CPU test suite).
EXPORTS = [
    "tbk_search2", "tbk_setpar", "tbk_setparv", "tbk_setpar_spin", "tbk_setparv_spin", "tbk_setpar_noncol",
    "tbk_reciprocal_lattice", "tbk_space_group",
    "tbk_space_group_read", "tbk_kptgen", "tbk_kmesh_weights", "tbk_klist", "tbk_kpath",
    "tbk_plan_create", "tbk_plan_run",
    "tbk_plan_launch", "tbk_plan_wait", "tbk_plan_eigenvalues", "tbk_plan_eigenvectors",
    "tbk_plan_hs", "tbk_plan_projections", "tbk_plan_nv", "tbk_plan_npairs", "tbk_plan_destroy",
    "tbk_plan_stage_times", "tbk_plan_set_timing", "tbk_plan_autotune", "tbk_gemm_tile",
    "tbk_plan_create_group", "tbk_plan_result", "tbk_plan_structure_eigenvalues",
    "tbk_plan_create_spin", "tbk_plan_magnetization", "tbk_plan_create_noncol",
    "tbk_plan_set_spin_directions", "tbk_plan_spin_directions",
    "tbk_plan_set_graph", "tbk_plan_set_window", "tbk_plan_set_fused_build", "tbk_plan_set_tridiag", "tbk_plan_order", "tbk_plan_nstruct", "tbk_plan_device_bytes",
    "tbk_multi_create", "tbk_multi_launch", "tbk_multi_wait", "tbk_multi_run", "tbk_multi_result",
    "tbk_multi_eigenvalues", "tbk_multi_ngroups", "tbk_multi_group_order", "tbk_multi_set_graph",
    "tbk_multi_autotune", "tbk_multi_device_bytes", "tbk_multi_destroy", "tbk_plan_query",
    "tbk_server_create", "tbk_server_submit", "tbk_server_result", "tbk_server_release",
    "tbk_server_stats", "tbk_server_timing", "tbk_server_destroy",
    "tbk_zhegv", "tbk_zhegv_noalloc", "tbk_perturb", "tbk_fermi",
    "tbk_read_params", "tbk_skin_open", "tbk_skin_next", "tbk_skin_close", "tbk_skeng_line", "tbk_static_run",
    "tbk_error_string", "tbk_debug_stages", "tbk_debug_zgemm", "tbk_debug_copy",
    "tbk_debug_gemm_time",
]


class TbkError(RuntimeError):
    def __init__(self, code: int, what: str = ""):
        self.code = code
        msg = _lib.tbk_error_string(code).decode() if _lib is not None else str(code)
        super().__init__(f"{what}: {msg} (code {code})" if what else f"{msg} (code {code})")


class StructureDesc(C.Structure):
    _fields_ = [
        ("natoms", C.c_int), ("lat", C.POINTER(C.c_double)), ("posn", C.POINTER(C.c_double)),
        ("posq", C.POINTER(C.c_double)), ("kkind", C.POINTER(C.c_int)), ("nd", C.c_int * 3),
        ("kinds", C.c_int), ("kbas", C.POINTER(C.c_int)), ("valence", C.POINTER(C.c_double)),
        ("dnn", C.POINTER(C.c_double)), ("screenl", C.POINTER(C.c_double)),
        ("param", C.POINTER(C.c_double)), ("nparam", C.c_int), ("nltype", C.c_int),
        ("jsover", C.c_int), ("thisvol", C.c_double), ("zfill", C.c_double),
    ]


TBK_MAXKINDS = 8
_MINT = TBK_MAXKINDS * (TBK_MAXKINDS + 1) // 2


class ParamSet(C.Structure):
    """tbk_param_set: one NRL parameter file's header (input.f:181-359)."""
    _fields_ = [
        ("jsover", C.c_int), ("jspins", C.c_int), ("nltype", C.c_int), ("kinds", C.c_int),
        ("nparam", C.c_int), ("kbas", C.c_int * TBK_MAXKINDS), ("valence", C.c_double * TBK_MAXKINDS),
        ("dnn", C.c_double * _MINT), ("screenl", C.c_double * _MINT), ("title", C.c_char * 80),
    ]


class SkinStructure(C.Structure):
    """tbk_skin_structure: what the reference program reads of a structure."""
    _fields_ = [
        ("title", C.c_char * 72), ("label", C.c_char * 24), ("zcharge", C.c_double), ("zfill", C.c_double),
        ("thisvol", C.c_double), ("lat", C.c_double * 9), ("latnum", C.c_int), ("strain", C.c_int),
        ("natoms", C.c_int), ("nd", C.c_int * 3), ("lphonon", C.c_int), ("spinflip", C.c_int),
        ("lnoncol", C.c_int), ("ltorque", C.c_int), ("lspdir", C.c_int), ("nk", C.c_int), ("inpstat", C.c_int),
    ]


_lib = None


def lib():
    """The loaded library; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        d = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        i = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        ip = C.POINTER(C.c_int)
        dp = C.POINTER(C.c_double)
        vp = C.c_void_p
        L.tbk_search2.argtypes = [C.c_int, d, d, i, C.c_int, i, d, d, C.c_int, ip, d, d, d, d,
                                  d, i, i, dp]
        L.tbk_setpar.argtypes = [C.c_int, i, C.c_int, d, C.c_int, C.c_int, C.c_int, d, d, i, i,
                                 d, d, d, d]
        L.tbk_setparv.argtypes = [C.c_int, i, C.c_int, d, C.c_int, C.c_int, C.c_double, C.c_int,
                                  d, d, d, i, i, d, d, d, d]
        L.tbk_setpar_spin.argtypes = [C.c_int, i, C.c_int, d, C.c_int, C.c_int, vp, C.c_int, C.c_int,
                                      C.c_int, d, d, i, i, d, d, d, d]
        L.tbk_setparv_spin.argtypes = [C.c_int, i, C.c_int, d, C.c_int, C.c_int, vp, C.c_int, C.c_int,
                                       C.c_double, C.c_int, d, d, d, i, i, d, d, d, d]
        L.tbk_setpar_noncol.argtypes = [C.c_int, i, C.c_int, d, C.c_int, C.c_int, C.c_int, C.c_double,
                                        C.c_int, d, d, d, i, i, d, d, d, d, d, d, d, d, d]
        L.tbk_reciprocal_lattice.argtypes = [d, d, d, d, dp]
        L.tbk_space_group.argtypes = [C.c_int, C.c_int, d, d, C.c_int, d, d, C.c_void_p, d, d, d]
        L.tbk_space_group_read.argtypes = [C.c_char_p, C.c_int, C.c_int, d, d, C.c_int, ip, d, d,
                                           d, C.c_char_p]
        L.tbk_kptgen.argtypes = [C.c_int, d, d, d, C.c_double, C.c_int, d, i, C.c_int, C.c_int,
                                 C.c_int, ip, d, d]
        L.tbk_kmesh_weights.argtypes = [C.c_int, d, d, C.c_int, C.c_int, ip]
        L.tbk_klist.argtypes = [C.c_int, d, C.c_double, d, d]
        L.tbk_kpath.argtypes = [C.c_int, i, d, C.c_int, ip, d, d]
        L.tbk_plan_create_group.argtypes = [C.c_int, C.c_int, C.POINTER(StructureDesc), i, d, d,
                                             C.c_int, C.c_int, C.POINTER(C.c_void_p)]
        L.tbk_plan_create_spin.argtypes = [C.c_int, C.POINTER(StructureDesc), d, vp, C.c_int, d, d,
                                            C.c_int, C.c_int, C.POINTER(C.c_void_p)]
        L.tbk_plan_magnetization.argtypes = [vp, dp]
        L.tbk_plan_set_spin_directions.argtypes = [vp, C.c_int]
        L.tbk_plan_spin_directions.argtypes = [vp, C.c_int, dp]
        L.tbk_plan_create_noncol.argtypes = [C.c_int, C.POINTER(StructureDesc), d, d, C.c_int, d, d,
                                              C.c_int, C.c_int, C.POINTER(C.c_void_p)]
        L.tbk_plan_result.argtypes = [vp, C.c_int, dp, dp, dp]
        L.tbk_plan_structure_eigenvalues.argtypes = [vp, C.c_int, vp, vp]
        L.tbk_plan_set_graph.argtypes = [vp, C.c_int]
        L.tbk_plan_set_tridiag.argtypes = [vp, C.c_int]
        L.tbk_plan_set_window.argtypes = [vp, C.c_int]
        L.tbk_plan_set_fused_build.argtypes = [vp, C.c_int]
        L.tbk_plan_order.argtypes = [vp]
        L.tbk_plan_nstruct.argtypes = [vp]
        L.tbk_plan_device_bytes.argtypes = [vp]
        L.tbk_plan_device_bytes.restype = C.c_longlong
        L.tbk_multi_create.argtypes = [C.c_int, C.c_int, C.POINTER(StructureDesc), i, d, d, C.c_int,
                                        C.c_int, C.c_longlong, C.POINTER(C.c_void_p)]
        L.tbk_multi_launch.argtypes = [vp, C.c_int, C.c_double, C.c_double]
        L.tbk_multi_wait.argtypes = [vp]
        L.tbk_multi_run.argtypes = [vp, C.c_int, C.c_double, C.c_double]
        L.tbk_multi_result.argtypes = [vp, C.c_int, dp, dp, dp]
        L.tbk_multi_eigenvalues.argtypes = [vp, C.c_int, vp, vp]
        L.tbk_multi_ngroups.argtypes = [vp]
        L.tbk_multi_group_order.argtypes = [vp, C.c_int]
        L.tbk_multi_set_graph.argtypes = [vp, C.c_int]
        L.tbk_multi_autotune.argtypes = [vp, C.c_int]
        L.tbk_multi_device_bytes.argtypes = [vp]
        L.tbk_multi_device_bytes.restype = C.c_longlong
        L.tbk_multi_destroy.argtypes = [vp]
        L.tbk_plan_query.argtypes = [vp]
        L.tbk_server_create.argtypes = [C.c_int, C.c_int, C.c_longlong, C.c_int, C.c_int,
                                         C.POINTER(C.c_void_p)]
        L.tbk_server_submit.argtypes = [vp, C.POINTER(StructureDesc), C.c_int, d, d, C.c_int,
                                         C.c_double, C.c_double, C.POINTER(C.c_longlong)]
        L.tbk_server_result.argtypes = [vp, C.c_longlong, C.c_int, ip, dp, dp, dp, dp, dp]
        L.tbk_server_release.argtypes = [vp, C.c_longlong]
        L.tbk_server_stats.argtypes = [vp, C.POINTER(C.c_longlong)]
        L.tbk_server_timing.argtypes = [vp, dp]
        L.tbk_server_destroy.argtypes = [vp]
        L.tbk_debug_copy.argtypes = [C.c_int, C.c_longlong]
        L.tbk_debug_gemm_time.argtypes = [C.c_int] * 11 + [dp]
        L.tbk_plan_autotune.argtypes = [vp, C.c_int, ip]
        L.tbk_gemm_tile.argtypes = [C.c_int] * 8
        L.tbk_plan_create.argtypes = [C.c_int, C.POINTER(StructureDesc), C.c_int, d, d, C.c_int,
                                      C.c_int, C.POINTER(vp)]
        L.tbk_plan_run.argtypes = [vp, C.c_int, C.c_double, C.c_double, dp, dp, dp]
        L.tbk_plan_launch.argtypes = [vp, C.c_int, C.c_double, C.c_double]
        L.tbk_plan_wait.argtypes = [vp, dp, dp, dp]
        L.tbk_plan_eigenvalues.argtypes = [vp, vp, vp]
        L.tbk_plan_eigenvectors.argtypes = [vp, C.c_int, vp]
        L.tbk_plan_hs.argtypes = [vp, C.c_int, C.c_int, vp, vp]
        L.tbk_plan_projections.argtypes = [vp, C.c_int, C.c_double, ip, vp, vp]
        L.tbk_plan_nv.argtypes = [vp]
        L.tbk_plan_npairs.argtypes = [vp]
        L.tbk_plan_destroy.argtypes = [vp]
        L.tbk_plan_stage_times.argtypes = [vp, C.c_int, C.POINTER(C.c_char_p), dp, ip]
        L.tbk_plan_set_timing.argtypes = [vp, C.c_int]
        L.tbk_zhegv.argtypes = [C.c_int, C.c_int, vp, vp, vp, vp]
        L.tbk_zhegv_noalloc.argtypes = [C.c_int, C.c_int, vp, vp, vp, vp]
        L.tbk_perturb.argtypes = [C.c_int, C.c_int, vp, vp, vp, vp, C.c_double, ip, vp]
        L.tbk_fermi.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, d, d, C.c_double,
                                C.c_double, dp]
        L.tbk_debug_stages.argtypes = [C.c_int, C.c_int] + [vp] * 10
        L.tbk_debug_zgemm.argtypes = [C.c_int] * 9 + [vp, C.c_int, vp, C.c_int, vp, C.c_int]
        L.tbk_read_params.argtypes = [C.c_char_p, C.POINTER(ParamSet), C.c_int, d]
        L.tbk_skin_open.argtypes = [C.c_char_p, C.POINTER(vp), ip, dp, dp, C.c_char_p, C.c_int]
        L.tbk_skin_next.argtypes = [vp, C.POINTER(ParamSet), C.c_int, C.c_int, C.POINTER(SkinStructure), i, d, d,
                                    i, d, d, d]
        L.tbk_skin_close.argtypes = [vp]
        L.tbk_skeng_line.argtypes = [C.c_char_p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.c_int, C.c_int, C.c_char_p, C.c_int]
        L.tbk_static_run.argtypes = [C.c_char_p, C.c_int, C.c_int, ip]
        L.tbk_error_string.argtypes = [C.c_int]
        L.tbk_error_string.restype = C.c_char_p
        _lib = L
    return _lib


def check(rc: int, what: str = ""):
    if rc != TBK_OK:
        raise TbkError(rc, what)
