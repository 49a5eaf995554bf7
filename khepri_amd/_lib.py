"""ctypes binding of libkhx.so (the C-ABI declared in include/khx.h).

The library is required: importing an op when libkhx.so is missing raises
immediately -- there is no CPU fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("KHX_LIBRARY", os.path.join(_HERE, "libkhx.so"))

_vp, _dp, _ip, _i, _sz = C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_size_t

# name -> (restype, argtypes); every pointer is passed as c_void_p
SIGNATURES = {
    "khx_last_error": (C.c_char_p, []),
    "khx_abi_version": (_i, []),
    "khx_kvectors": (_i, [_dp, _i, _dp, _dp, _i, _dp, _dp, _vp]),
    "khx_convmat_ws_bytes": (_sz, [_i, _i, _i]),
    "khx_convmat": (_i, [_dp, _i, _i, _i, _i, _i, _i, _dp, _vp, _vp]),
    "khx_zgesv": (_i, [_i, _i, _dp, _dp, _i, _ip, _vp]),
    "khx_zgemm": (_i, [_i, _i, _i, _dp, _dp, _i, C.c_longlong, _dp, _i, C.c_longlong, _dp, _dp, _i, C.c_longlong,
                       _i, _vp]),
    "khx_eig_ws_bytes": (_sz, [_i, _i]),
    "khx_eig": (_i, [_i, _dp, _i, _dp, _dp, _ip, _vp, _vp]),
    "khx_eig_timed": (_i, [_i, _dp, _i, _dp, _dp, _ip, _vp, _vp, _vp]),
    "khx_eig_noalloc": (_i, [_i, _dp, _i, _dp, _dp, _ip, _vp]),
    "khx_eig_noalloc_bytes": (_sz, []),
    "khx_patterned_layer_ws_bytes": (_sz, [_i, _i]),
    "khx_patterned_layer": (_i, [_dp, _dp, _i, _dp, _dp, _i, _i, _dp, _dp, _dp, _dp, _dp, _ip, _vp, _vp]),
    "khx_patterned_layer_cached": (_i, [_dp, _dp, _i, _dp, _dp, _i, _i, _dp, _dp, _dp, _dp, _dp, _ip, _dp, _vp, _vp]),
    "khx_uniform_layer": (_i, [_dp, _dp, _i, _i, _dp, _dp, _dp, _dp, _dp, _dp, _vp]),
    "khx_halfspace": (_i, [_i, _dp, _dp, _i, _i, _dp, _dp, _dp, _dp, _dp, _vp]),
    "khx_redheffer_ws_bytes": (_sz, [_i, _i]),
    "khx_redheffer": (_i, [_dp, _i, _dp, _i, _i, _i, _dp, _ip, _vp, _vp]),
    "khx_poynting_ws_bytes": (_sz, [_i, _i]),
    "khx_poynting": (_i, [_dp, _i, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _i, _dp, _dp, _dp, _dp, _vp, _vp]),
    "khx_incident": (_i, [_i, _dp, _dp, _dp, _dp, _dp, _i, _dp, _vp]),
    "khx_poynting_vec": (_i, [_dp, _dp, _i, _dp, _dp, _dp, _dp, _dp, _i, _dp, _dp, _dp, _dp, _vp]),
    "khx_lu_factor": (_i, [_i, _dp, _i, _ip, _ip, _vp]),
    "khx_lu_solve": (_i, [_i, _i, _dp, _ip, _dp, _i, _vp]),
    "khx_lu_factor_solve": (_i, [_i, _i, _dp, _dp, _i, _ip, _ip, _vp]),
    "khx_diag_scale": (_i, [_dp, _i, C.c_longlong, _dp, _dp, _i, _i, _i, _dp, _i, C.c_longlong, _vp]),
    "khx_identity_minus": (_i, [_dp, _i, _i, _dp, _vp]),
    "khx_zaxpby": (_i, [_dp, _i, C.c_longlong, _dp, _dp, _i, C.c_longlong, _dp, _dp, _i, C.c_longlong, _i, _i, _i,
                        _vp]),
    "khx_propagation": (_i, [_dp, _dp, _i, _i, _dp, _dp, _vp]),
    "khx_eigenbasis": (_i, [_dp, _dp, _i, _i, _dp, _vp]),
    "khx_mode_propagate": (_i, [_dp, _dp, _dp, _i, _i, _dp, _vp]),
    "khx_longitudinal": (_i, [_dp, _dp, _dp, _i, _i, _dp, _dp, _vp]),
    "khx_idft_phase": (_i, [_dp, _dp, C.c_double, _dp, _dp, _i, _i, _dp, _vp]),
    "khx_mode_propagate_batch": (_i, [_dp, _dp, _dp, _i, _i, _i, _dp, _vp]),
    "khx_longitudinal_batch": (_i, [_dp, _dp, _dp, _i, _i, _i, _dp, _dp, _vp]),
    "khx_idft_phase_batch": (_i, [_dp, _dp, _dp, _dp, _dp, _i, _i, _i, _dp, _vp]),
    "khx_joint_subspace": (_i, [_dp, _i, _i, _i, _i, _dp, _vp]),
    "khx_joint_subspace_diag": (_i, [_dp, _i, _i, _i, _dp, _vp]),
    "khx_zgemm_op": (_i, [_i, _i, _i, _i, _i, _dp, _dp, _i, C.c_longlong, _dp, _i, C.c_longlong, _dp, _dp, _i,
                          C.c_longlong, _i, _vp]),
    "khx_conj_transpose": (_i, [_dp, _i, C.c_longlong, _i, _i, _dp, _i, C.c_longlong, _i, _vp]),
    "khx_reduce_dot": (_i, [_dp, _i, C.c_longlong, _dp, _i, C.c_longlong, _i, _i, _i, _i, _dp, C.c_longlong, _i,
                            _vp]),
    "khx_eig_adjoint_mix": (_i, [_dp, _dp, _dp, _i, _i, C.c_double, _vp]),
    "khx_lambda_adjoint": (_i, [_dp, _dp, _dp, _dp, _i, _i, _dp, _vp]),
    "khx_build_pq": (_i, [_dp, _dp, _i, _dp, _dp, _i, _i, _dp, _dp, _vp]),
    "khx_pq_adjoint": (_i, [_dp, _dp, _dp, _dp, _i, _i, _dp, _dp, _vp]),
    "khx_convmat_adjoint_ws_bytes": (_sz, [_i, _i, _i]),
    "khx_convmat_adjoint": (_i, [_dp, _i, _i, _i, _i, _i, _i, _dp, _vp, _vp]),
    "khx_poynting_vec_adjoint": (_i, [_dp, _dp, _i, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _i, _dp, _dp, _vp]),
    "khx_patterned_layer_bwd_ws_bytes": (_sz, [_i, _i]),
    "khx_patterned_layer_bwd": (_i, [_dp, _dp, _i, _dp, _dp, _i, _i, _dp, _dp, _dp, _dp, _dp, C.c_double, _dp, _ip,
                                     _vp, _vp]),
    "khx_patterned_layer_bwd_cached": (_i, [_dp, _dp, _i, _dp, _dp, _i, _i, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double,
                                            _dp, _ip, _vp, _vp]),
    "khx_redheffer_bwd_ws_bytes": (_sz, [_i, _i]),
    "khx_redheffer_bwd": (_i, [_dp, _dp, _i, _i, _dp, _dp, _dp, _ip, _vp, _vp]),
    "khx_beam_fields": (_i, [_i, _vp, _vp, _vp, C.c_longlong, C.c_double, C.c_double, C.c_double, _i, _i, _vp, _vp,
                             _vp, _vp]),
    "khx_bloch_amplitudes": (_i, [_vp, _vp, _vp, _i, _i, _i, C.c_double, C.c_double, _vp, _vp, _i, _vp, _vp]),
    "khx_points_dft": (_i, [_vp, _vp, _vp, C.c_longlong, _vp, _vp, _i, _i, C.c_double, _vp, _vp]),
    "khx_scattering_splitlr": (_i, [_i, _vp, _i, _vp, _vp, _vp]),
    "khx_scattering_eig_ws_bytes": (_sz, [_i, _i]),
    "khx_scattering_eig": (_i, [_i, _vp, _i, C.c_double, C.c_double, _vp, _vp, _vp, _vp, _vp]),
    "khx_scattering_det_ws_bytes": (_sz, [_i, _i]),
    "khx_scattering_det": (_i, [_i, _vp, _i, _vp, _vp, _vp, _vp]),
    "khx_convmat_islands": (_i, [_vp, _vp, _vp, _i, _vp, _vp, C.c_double, _i, _i, _vp, _vp, _vp]),
    "khx_eig_slots": (_i, [_i]),
    "khx_gemm_record": (_i, [_i, _vp, _i]),
    "khx_gemm_force": (_i, [_i]),
    "khx_gemm_tune_set": (_i, [_i, _i, _i, _i, _i, _i, _i]),
    "khx_prof_enable": (_i, [_i]),
    "khx_prof_collect": (_i, [C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_longlong)]),
    "khx_copy": (_i, [_vp, _vp, _sz, _vp]),
}

_lib = None


class KhxError(RuntimeError):
    pass


def load():
    """Load libkhx.so once; raise if it is absent (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m khepri_amd.build` "
            "(the RCWA path has no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc, what):
    if rc != 0:
        msg = load().khx_last_error().decode(errors="replace")
        if rc == -1:
            raise ValueError(f"{what}: {msg}")
        raise KhxError(f"{what} failed ({rc}): {msg}")
