"""ctypes binding of libcfmi.so (the C ABI declared in include/cfmi.h).

The library is loaded from this package directory.  If it is missing the
import raises: there is no CPU fallback for the product path.
"""

from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcfmi.so")

CFMI_OK = 0
CFMI_F64 = 0
CFMI_F32 = 1
CFMI_BF16 = 2
CFMI_BF16X3 = 3
CFMI_GEMM_ACC64 = 128

_c = ctypes
vp = _c.c_void_p
i64 = _c.c_int64
ci = _c.c_int
cd = _c.c_double

# name -> (restype, argtypes); kept in the order of include/cfmi.h.
SIGNATURES = {
    "cfmi_version": (ci, []),
    "cfmi_create": (ci, [ci, vp, _c.POINTER(vp)]),
    "cfmi_destroy": (ci, [vp]),
    "cfmi_last_error": (_c.c_char_p, [vp]),
    "cfmi_synchronize": (ci, [vp]),
    "cfmi_pointings": (ci, [vp, i64, ci, vp, i64, vp]),
    "cfmi_pointings_host": (ci, [i64, ci, vp, i64, vp]),
    "cfmi_signal_00": (ci, [vp, vp, ci, vp, i64, vp, i64, ci, vp, i64, vp, i64]),
    "cfmi_signal_22": (ci, [vp, vp, vp, vp, ci, vp, i64, vp, i64]),
    "cfmi_signal_02": (ci, [vp, vp, vp, ci, vp, i64, vp, i64, vp, i64, vp, i64]),
    "cfmi_deriv_state_doubles": (i64, [i64, i64, ci]),
    "cfmi_deriv_stack_00": (ci, [vp, ci, vp, ci, vp, vp, ci, vp, i64, vp, i64, ci, vp, i64, i64, vp, i64, i64, vp,
                                 ci]),
    "cfmi_deriv_stack_22": (ci, [vp, ci, vp, vp, vp, ci, vp, vp, ci, vp, i64, vp, vp, vp, i64, i64, vp, ci]),
    "cfmi_deriv_stack_02": (ci, [vp, ci, vp, vp, ci, vp, vp, ci, vp, i64, vp, i64, vp, vp, i64, i64, vp, vp,
                                 i64, i64, vp, ci]),
    "cfmi_gemm": (ci, [vp, ci, ci, ci, i64, i64, i64, cd, vp, i64, vp, i64, cd, vp, i64, ci]),
    "cfmi_gemm_packed": (ci, [vp, ci, ci, ci, i64, i64, i64, cd, vp, i64, vp, i64, cd, vp, i64, i64, i64, cd,
                              cd, ci]),
    "cfmi_gemm_batched": (ci, [vp, ci, ci, ci, i64, i64, i64, cd, vp, i64, i64, vp, i64, i64, cd, vp, i64, i64,
                               i64, ci]),
    "cfmi_gemm_packed_batched": (ci, [vp, ci, ci, ci, i64, i64, i64, cd, vp, i64, i64, vp, i64, i64, cd, vp, i64,
                                      i64, i64, i64, cd, cd, i64, ci]),
    "cfmi_gemm_ex": (ci, [vp, ci, ci, ci, ci, i64, i64, i64, cd, vp, i64, i64, i64, vp, i64, i64, i64, cd, vp,
                          i64, i64, i64, ci, ci, i64, i64, i64, cd, cd]),
    "cfmi_gemm_b3": (ci, [vp, ci, i64, i64, i64, cd, vp, i64, i64, i64, i64, vp, i64, i64, i64, i64, cd, vp, i64,
                          i64, i64, i64, ci, i64, i64, i64, cd]),
    "cfmi_split3": (ci, [vp, ci, i64, i64, vp, i64, vp, i64, i64]),
    "cfmi_dot_batched": (ci, [vp, i64, i64, vp, i64, i64, vp, i64, vp, i64]),
    "cfmi_offdiag_nnz": (ci, [vp, i64, vp, i64, _c.POINTER(i64)]),
    "cfmi_join3": (ci, [vp, i64, vp, i64, vp]),
    "cfmi_slice8": (ci, [vp, ci, ci, i64, i64, vp, i64, vp, i64, i64, vp]),
    "cfmi_unslice8": (ci, [vp, ci, i64, i64, vp, i64, i64, vp, vp, i64]),
    "cfmi_transpose": (ci, [vp, i64, i64, vp, i64, vp, i64]),
    "cfmi_s8_kmax": (i64, [ci]),
    "cfmi_gemm_s8": (ci, [vp, ci, i64, i64, i64, cd, vp, i64, i64, vp, i64, vp, i64, i64, vp, i64, cd, vp, i64, ci,
                          i64, i64, i64, cd]),
    "cfmi_jacobi_scale": (ci, [vp, i64, vp, i64, vp]),
    "cfmi_potrf": (ci, [vp, ci, i64, vp, i64, vp, i64, _c.POINTER(ci)]),
    "cfmi_potrf_s8": (ci, [vp, ci, i64, vp, i64, vp, i64, _c.POINTER(ci)]),
    "cfmi_s8_stats_enable": (ci, [vp, ci]),
    "cfmi_s8_stats_read": (ci, [vp, _c.POINTER(i64), _c.POINTER(cd), _c.POINTER(cd), _c.POINTER(cd)]),
    "cfmi_trtri": (ci, [vp, ci, i64, vp, i64, vp, i64]),
    "cfmi_trtri_blocks": (ci, [vp, ci, i64, i64, vp, i64, vp, i64]),
    "cfmi_logdet_chol": (ci, [vp, i64, vp, i64, vp, _c.POINTER(cd)]),
    "cfmi_scale": (ci, [vp, ci, i64, i64, vp, i64, vp]),
    "cfmi_axpby": (ci, [vp, i64, i64, cd, vp, i64, cd, vp, i64]),
    "cfmi_convert": (ci, [vp, ci, i64, i64, vp, i64, vp, i64]),
    "cfmi_copy": (ci, [vp, ci, i64, i64, vp, i64, vp, i64]),
    "cfmi_fill": (ci, [vp, ci, i64, i64, vp, i64, cd]),
    "cfmi_gram": (ci, [vp, ci, i64, i64, i64, vp, i64, vp, i64, vp, i64]),
    "cfmi_harmonic_modes": (ci, [vp, ci, i64, vp, ci, ci, vp, i64, i64, i64, i64, i64, i64]),
    "cfmi_harmonic_fisher": (ci, [vp, vp, vp, i64, ci, ci, vp, vp, ci, vp, vp, vp, ci, vp]),
    "cfmi_harmonic_pairsum": (ci, [vp, i64, vp, i64, i64, vp, i64, i64, i64, ci, ci, vp, vp, ci, vp, vp, vp, ci,
                                   cd, vp, i64]),
    "cfmi_harmonic_coldots": (ci, [vp, i64, vp, i64, vp, i64, vp, ci, i64, ci, ci, vp, vp, ci, vp, vp, vp, ci, cd,
                                   vp]),
    "cfmi_batch_create": (ci, [vp, ci, ci, vp, ci, vp, vp, vp, vp, ci, vp, ci, ci, ci, _c.POINTER(vp)]),
    "cfmi_batch_run": (ci, [vp, vp, vp, vp, vp]),
    "cfmi_batch_capture": (ci, [vp, vp, vp, vp, vp]),
    "cfmi_batch_replay": (ci, [vp]),
    "cfmi_batch_destroy": (ci, [vp]),
    "cfmi_mlp_param_count": (i64, []),
    "cfmi_mlp_workspace_bytes": (i64, [ci, ci]),
    "cfmi_mlp_forward": (ci, [vp, ci, ci, vp, vp, vp, vp, vp, vp]),
    "cfmi_mlp_loss": (ci, [vp, ci, ci, vp, vp, vp, vp, vp]),
    "cfmi_mlp_backward": (ci, [vp, ci, ci, vp, vp, vp, vp, vp, vp, vp, ci]),
    "cfmi_mlp_backward_noalloc": (ci, [vp, ci, ci, vp, vp, vp, vp, vp, vp, ci]),
    "cfmi_mlp_adam": (ci, [vp, i64, vp, vp, vp, vp, cd, cd, cd, cd, ci, vp]),
    "cfmi_mlp_to_bf16": (ci, [vp, i64, vp, vp]),
    "cfmi_mlp_ensemble_stats": (ci, [vp, ci, ci, vp, vp, vp]),
}

_lib = None


class CfmiError(RuntimeError):
    """A libcfmi call returned an API / HIP error."""


def load():
    """Load libcfmi.so once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m cosmoforge_amd.build` "
            "(the HIP path has no CPU fallback)"
        )
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    return list(SIGNATURES)
