"""ctypes binding of ``libsktopt_amd.so`` (the C ABI in include/sktopt_amd.h).

The library is mandatory: importing a compute entry point without the built
``.so`` raises ImportError -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_HERE = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("SKTOPT_AMD_LIB", _HERE / "_lib" / "libsktopt_amd.so"))

SKT_OK, SKT_EINVAL, SKT_EHIP, SKT_ENOCONV, SKT_ENOTIMPL = 0, 1, 2, 3, 4
F64, F32 = 0, 1
PC_JACOBI, PC_MG, PC_AUTO = 0, 1, 2
SIMP, RAMP, RAMP_EXACT = 0, 1, 2
ELASTIC, HEAT = 0, 1
RED_SUM, RED_DOT, RED_MAXABS, RED_MAXABSDIFF = 0, 1, 2, 3


class ConvergenceError(RuntimeError):
    """The iterative state solve did not reach its tolerance."""


_vp, _i64, _i32, _d, _u8p = C.c_void_p, C.c_int64, C.c_int, C.c_double, C.c_void_p
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)
_pp = C.POINTER(C.c_void_p)

# name -> argtypes (all return int status except skt_last_error/skt_abi_version)
SIGNATURES = {
    "skt_heaviside_project": [_i64, _vp, _d, _d, _vp, _vp],
    "skt_heaviside_derivative": [_i64, _vp, _d, _d, _vp, _vp],
    "skt_interpolate": [_i32, _i64, _vp, _d, _d, _d, _vp, _i32, _vp],
    "skt_dC_drho": [_i32, _i64, _vp, _vp, _i64, _d, _d, _d, _vp, _vp],
    "skt_sensitivity_chain": [_i32, _i64, _i32, _vp, _vp, _vp, _d, _d, _d, _d, _d, _d, _vp, _vp],
    "skt_sensitivity_chain_cols": [_i32, _i64, _i32, _vp, _vp, _vp, _d, _d, _d, _d, _d, _vp, _vp],
    "skt_sensitivity_chain_extra": [_i32, _i64, _i32, _vp, _vp, _vp, _vp, _d, _d, _d, _d, _d, _d,
                                    _i32, _vp, _vp],
    "skt_ewise": [_i32, _i64, _d, _vp, _d, _vp, _vp, _vp],
    "skt_reduce": [_i32, _i64, _vp, _vp, _vp, _vp],
    "skt_reduce_cols": [_i32, _i64, _i32, _vp, _vp, _vp, _vp],
    "skt_percentile_abs": [_i64, _vp, _d, _vp, _vp],
    "skt_percentile_abs_noalloc": [_i64, _vp, _d, _vp, _vp],
    "skt_stats4": [_i64, _vp, _vp, _vp],
    "skt_row_mean": [_i64, _i32, _vp, _vp, _vp],
    "skt_hex8_create": [_i32, _i64, _i64, _i64, _d, _d, _d, _d, _i32, _vp, _pp],
    "skt_tet4_create": [_i32, _i64, _i64, _vp, _vp, _d, _vp, _pp],
    "skt_tet4_create_scalar": [_i64, _i64, _vp, _vp, _d, _d, _vp, _pp],
    "skt_tet4_set_grid": [_vp, _i64, _i64, _i64],
    "skt_tet4_elem_to_node": [_vp, _i32, _vp, _vp, _d, _vp, _vp, _vp],
    "skt_tet4_node_to_elem": [_vp, _i32, _vp, _i32, _i32, _vp, _vp],
    "skt_operator_destroy": [_vp],
    "skt_operator_info": [_vp, _i64p, _i64p, C.POINTER(C.c_int)],
    "skt_operator_element_matrix": [_vp, _dp],
    "skt_operator_apply": [_vp, _i32, _i32, _vp, _vp, _vp, _vp],
    "skt_operator_diag": [_vp, _i32, _vp, _vp, _vp],
    "skt_operator_energy": [_vp, _i32, _i32, _vp, _vp, _vp, _vp],
    "skt_solver_create": [_vp, _i32, _pp],
    "skt_solver_destroy": [_vp],
    "skt_solver_configure": [_vp, _i32, _i32, _i32],
    "skt_hex8_create_scalar": [_i64, _i64, _i64, _d, _d, _d, _d, _d, _vp, _pp],
    "skt_hex8_elem_to_node": [_i64, _i64, _i64, _i32, _vp, _vp, _d, _vp, _vp, _vp],
    "skt_hex8_node_to_elem": [_i64, _i64, _i64, _i32, _vp, _i32, _i32, _vp, _vp],
    "skt_solver_info": [_vp, _i32p, _i32p],
    "skt_solver_set_timing": [_vp, _i32],
    "skt_solver_get_timing": [_vp, _dp, _i64p, _i32],
    "skt_solver_get_vcycle_timing": [_vp, _dp, _i64p, _i32],
    "skt_solver_graph_stats": [_vp, _i64p, _i64p],
    "skt_solver_solve": [_vp, _i32, _i32, _vp, _vp, _vp, _vp, _d, _i32, _d, _i64p, _dp, _vp],
    "skt_filter_create": [_i64, _vp, _vp, _d, _d, _pp],
    "skt_filter_destroy": [_vp],
    "skt_filter_info": [_vp, _i64p, _i64p],
    "skt_filter_export": [_vp, _vp, _vp, _vp, _vp, _vp],
    "skt_filter_forward": [_vp, _vp, _vp, _vp],
    "skt_filter_gradient": [_vp, _vp, _vp, _vp],
    "skt_filter_gather_design": [_vp, _vp, _vp, _vp],
    "skt_filter_scatter_design": [_vp, _vp, _vp, _vp],
    "skt_filter_set_grid": [_vp, _i64, _i64, _i64, _d, _d, _d, _i32p],
    "skt_design_gather": [_i64, _vp, _vp, _vp, _vp],
    "skt_hex8_nodal_average_pairing": [_i64, _i64, _i64, _vp, _vp, _vp, _vp],
    "skt_hex8_robin_virtual": [_i64, _i64, _i64, _d, _d, _d, _i32, _vp, _d, _d, _d, _d, _vp,
                               _vp, _vp],
    "skt_hex8_heat_exchange": [_i64, _i64, _i64, _d, _d, _d, _i32, _vp, _i32, _vp, _d, _d, _d,
                               _d, _vp, _vp, _vp, _vp],
    "skt_hex8_robin_grad": [_i64, _i64, _i64, _d, _d, _d, _i32, _vp, _vp, _i32, _vp, _d, _d,
                            _d, _d, _vp, _vp, _vp],
    "skt_operator_set_extra": [_vp, _vp, _vp],
    "skt_stencil27_apply": [_i64, _i64, _i64, _i32, _vp, _vp, _vp, _vp],
    "skt_tet4_nodal_average_pairing": [_vp, _vp, _vp, _vp, _vp],
    "skt_tet4_robin_virtual": [_vp, _i32, _vp, _d, _d, _d, _d, _vp, _vp, _vp, _vp],
    "skt_tet4_heat_exchange": [_vp, _vp, _i32, _vp, _d, _d, _d, _d, _vp, _vp, _vp, _vp, _vp],
    "skt_tet4_robin_grad": [_vp, _i32, _vp, _vp, _i32, _vp, _d, _d, _d, _d, _vp, _vp, _vp],
    "skt_tet4_set_extra": [_vp, _vp, _vp],
    "skt_tet4_extra_quadform": [_vp, _i32, _vp, _vp, _vp],
    "skt_design_scatter": [_i64, _vp, _vp, _vp, _vp],
    "skt_oc_candidate": [_i64, _vp, _vp, _vp, _d, _d, _d, _d, _d, _d, _d, _vp, _vp, _vp],
    "skt_oc_bisect": [_vp, _vp, _vp, _vp, _vp, _d, _d, _d, _d, _d, _d, _d, _d, _d, _d,
                      _d, _d, _d, _d, _i32, _vp, _vp, _vp],
    "skt_logmoc_update": [_i64, _vp, _vp, _d, _d, _d, _d, _d, _vp],
    "skt_oc_normalize": [_i64, _vp, _vp, _d, _d, _i32, _vp, _vp, _vp],
    "skt_oc_kkt": [_i64, _vp, _vp, _vp, _vp, _vp, _d, _d, _vp, _vp],
}

_lib = None


def load():
    """Load the shared library once; raise ImportError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python scikit-topt_amd/build.py` "
            "(there is no CPU fallback)")
    lib = C.CDLL(str(LIB_PATH))
    lib.skt_last_error.restype = C.c_char_p
    lib.skt_last_error.argtypes = []
    lib.skt_abi_version.restype = C.c_int
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    _lib = lib
    return lib


def last_error() -> str:
    msg = load().skt_last_error()
    return msg.decode() if msg else ""


def check(status: int, what: str = ""):
    if status == SKT_OK:
        return
    msg = f"{what}: {last_error()}" if what else last_error()
    if status == SKT_EINVAL:
        raise ValueError(msg)
    if status == SKT_ENOCONV:
        raise ConvergenceError(msg)
    if status == SKT_ENOTIMPL:
        raise NotImplementedError(msg)
    raise RuntimeError(msg)


def call(name: str, *args):
    """Call an entry point and raise on a non-zero status."""
    check(getattr(load(), name)(*args), name)


def exported_symbols() -> list[str]:
    return ["skt_last_error", "skt_abi_version", *SIGNATURES.keys()]
