"""ctypes binding of libmpskit_hip.so (the C ABI declared in include/mpskit_hip.h).

The library is loaded from this package's ``lib/`` directory (built in-tree by
``mpskit.jl_amd/build.py``).  There is no fallback: if the library or a GPU is
missing, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libmpskit_hip.so"

MK_OK = 0
STATUS_NAMES = {
    0: "MK_OK", 1: "MK_BAD_SHAPE", 2: "MK_BAD_ARG", 3: "MK_OOM", 4: "MK_HIP_ERROR",
    5: "MK_NOT_CONVERGED", 6: "MK_NOT_SUPPORTED", 7: "MK_NONFINITE",
}

FP64, FP32, BF16X3 = 0, 1, 2
PREC_NAMES = {"fp64": FP64, "fp32": FP32, "bf16x3": BF16X3}


class MKError(RuntimeError):
    """Raised for a non-OK status; ``.status`` holds the code name."""

    def __init__(self, status: int, msg: str):
        self.status = STATUS_NAMES.get(status, str(status))
        super().__init__(f"{self.status}: {msg}")


class SpaceMismatch(MKError, ValueError):
    """Shape errors, the analogue of TensorKit's SpaceMismatch."""


class NotConvergedError(MKError):
    pass


_I64 = C.c_int64
_P = C.c_void_p
_D = C.c_double
_INT = C.c_int


class JordanAC2Terms(C.Structure):
    """``mk_jordan_ac2_terms``: device pointers of the constructed two-site Jordan
    terms (NULL = missing) and the AA operator's channel counts."""
    _fields_ = [(n, _P) for n in ("II", "IC", "ID", "CB", "CA", "AB", "BE", "DE", "EE",
                                  "AA_GL", "AA_W1", "AA_W2", "AA_GR")] + \
               [("aa_wl", _I64), ("aa_wm", _I64), ("aa_wr", _I64)]


class EnvLevels(C.Structure):
    """``mk_env_levels``: the levels of a block-stored environment."""
    _fields_ = [("n", C.c_int), ("ptr", C.POINTER(_P)), ("width", C.POINTER(_I64))]


class JordanMPO(C.Structure):
    """``mk_jordan_mpo``: a Jordan MPO tensor as blocks (device d x d) or scalar identities."""
    _fields_ = [("wl", _I64), ("d", _I64), ("wr", _I64), ("nblocks", C.c_int), ("row", C.POINTER(C.c_int)),
                ("col", C.POINTER(C.c_int)), ("block", C.POINTER(_P)), ("scalar", C.POINTER(C.c_double))]


class JordanACTerms(C.Structure):
    """``mk_jordan_ac_terms``: the constructed one-site Jordan terms."""
    _fields_ = [(n, _P) for n in ("D", "I", "E", "C", "B", "A_GL", "A_W", "A_GR")] + \
               [("a_wl", _I64), ("a_wr", _I64)]

_SIGS = {
    "mk_version": ([], C.c_int),
    "mk_last_error": ([], C.c_char_p),
    "mk_ctx_create": ([_INT, C.POINTER(_P)], _INT),
    "mk_ctx_destroy": ([_P], _INT),
    "mk_ctx_set_stream": ([_P, _P], _INT),
    "mk_ctx_synchronize": ([_P], _INT),
    "mk_prepare_c": ([_P, _P, _P] + [_I64] * 5 + [_INT, C.POINTER(_P)], _INT),
    "mk_prepare_ac_generic": ([_P, _P, _P, _P] + [_I64] * 7 + [_INT, C.POINTER(_P)], _INT),
    "mk_prepare_ac2_generic": ([_P, _P, _P, _P, _P] + [_I64] * 9 + [_INT, C.POINTER(_P)], _INT),
    "mk_prepare_ac_jordan": ([_P, _P, _P, _P] + [_I64] * 5 + [_INT, C.POINTER(_P)], _INT),
    "mk_prepare_ac2_jordan": ([_P, _P, _P, _P, _P] + [_I64] * 7 + [_INT, C.POINTER(_P)], _INT),
    "mk_prepare_ac2_jordan_terms": ([_P, C.POINTER(JordanAC2Terms), _I64, _I64, _I64, _I64, _INT, C.POINTER(_P)],
                                    _INT),
    "mk_prepare_ac_jordan_terms": ([_P, C.POINTER(JordanACTerms), _I64, _I64, _I64, _INT, C.POINTER(_P)], _INT),
    "mk_prepare_ac2_jordan_levels": ([_P, C.POINTER(EnvLevels), C.POINTER(JordanMPO), C.POINTER(JordanMPO),
                                      C.POINTER(EnvLevels), _I64, _I64, _INT, C.POINTER(_P)], _INT),
    "mk_prepare_ac_jordan_levels": ([_P, C.POINTER(EnvLevels), C.POINTER(JordanMPO), C.POINTER(EnvLevels), _I64,
                                     _I64, _INT, C.POINTER(_P)], _INT),
    "mk_prepare_c_levels": ([_P, C.POINTER(EnvLevels), C.POINTER(EnvLevels)] + [_I64] * 4 + [_INT, C.POINTER(_P)],
                            _INT),
    "mk_transfer_left_mpo_levels": ([_P, C.POINTER(EnvLevels), C.POINTER(JordanMPO), _P, _P] + [_I64] * 4 +
                                    [C.POINTER(EnvLevels)], _INT),
    "mk_transfer_right_mpo_levels": ([_P, C.POINTER(EnvLevels), C.POINTER(JordanMPO), _P, _P] + [_I64] * 4 +
                                     [C.POINTER(EnvLevels)], _INT),
    "mk_op_set_charge_blocks": ([_P, _INT, C.POINTER(_I64), C.POINTER(_INT), _INT, C.POINTER(_I64), C.POINTER(_INT),
                                 C.POINTER(_INT)], _INT),
    "mk_op_compact_charge_blocks": ([_P], _INT),
    "mk_op_charge_blocks": ([_P, C.POINTER(_I64), _INT, C.POINTER(_INT)], _INT),
    "mk_krylov_orthogonalize": ([_P, _P, _I64, _INT, _P, _I64, C.POINTER(_D)], _INT),
    "mk_potrf": ([_P, _P, _I64, _I64, C.POINTER(_INT)], _INT),
    "mk_apply": ([_P, _P, _P, _D, _D, _INT], _INT),
    "mk_apply_noalloc": ([_P, _P, _P, _D, _D, _INT], _INT),
    "mk_op_info": ([_P, C.POINTER(_I64), C.POINTER(_I64), C.POINTER(_D)], _INT),
    "mk_op_free": ([_P], _INT),
    "mk_transfer_left_mps": ([_P, _P, _P, _P] + [_I64] * 5 + [_P], _INT),
    "mk_transfer_right_mps": ([_P, _P, _P, _P] + [_I64] * 5 + [_P], _INT),
    "mk_prepare_transfer_left": ([_P, _P, _P, _P] + [_I64] * 7 + [_INT, C.POINTER(_P)], _INT),
    "mk_prepare_transfer_right": ([_P, _P, _P, _P] + [_I64] * 7 + [_INT, C.POINTER(_P)], _INT),
    "mk_transfer_left_mpo": ([_P, _P, _P, _P, _P] + [_I64] * 7 + [_P], _INT),
    "mk_transfer_right_mpo": ([_P, _P, _P, _P, _P] + [_I64] * 7 + [_P], _INT),
    "mk_transfer_left_excited": ([_P, _P, _P, _P, _P] + [_I64] * 8 + [_P], _INT),
    "mk_transfer_right_excited": ([_P, _P, _P, _P, _P] + [_I64] * 8 + [_P], _INT),
    "mk_transfer_left_leg": ([_P, _P, _P, _P, _P] + [_I64] * 8 + [_P], _INT),
    "mk_transfer_right_leg": ([_P, _P, _P, _P, _P] + [_I64] * 8 + [_P], _INT),
    "mk_mpo_extend": ([_P, _P] + [_I64] * 5 + [_P], _INT),
    "mk_regularize": ([_P, _P, _P, _P, _I64, _I64, _I64], _INT),
    "mk_eigsolve": ([_P, _P, _INT, _INT, _D, _INT, C.POINTER(_D), C.POINTER(_D), C.POINTER(_INT),
                     C.POINTER(_INT)], _INT),
    "mk_gemm": ([_P, _INT, _INT, _I64, _I64, _I64, _D, _P, _I64, _P, _I64, _D, _P, _I64, _INT], _INT),
    "mk_batch_create": ([_P, _INT, C.POINTER(_P), C.POINTER(_P)], _INT),
    "mk_batch_layout": ([_P, C.POINTER(_I64), C.POINTER(_I64), C.POINTER(_INT)], _INT),
    "mk_batch_apply": ([_P, _P, _P, _INT], _INT),
    "mk_batch_eigsolve": ([_P, _P, _INT, _INT, _D, _INT, _INT, C.POINTER(_D), C.POINTER(_D), C.POINTER(_INT),
                           C.POINTER(_INT)], _INT),
    "mk_batch_free": ([_P], _INT),
    "mk_autotune_stats": ([C.POINTER(_INT), C.POINTER(_INT)], _INT),
    "mk_mlp_create": ([_P, C.c_uint64, _I64, C.POINTER(_P)], _INT),
    "mk_mlp_free": ([_P], _INT),
    "mk_mlp_param_count": ([C.POINTER(_I64)], _INT),
    "mk_mlp_get_params": ([_P, _P], _INT),
    "mk_mlp_set_params": ([_P, _P], _INT),
    "mk_mlp_forward": ([_P, _P, _I64, _P], _INT),
    "mk_mlp_gradient": ([_P, _P, _P, _I64, _P, C.POINTER(C.c_float)], _INT),
    "mk_mlp_dataset_size": ([_P, C.POINTER(_I64)], _INT),
    "mk_mlp_train": ([_P, _INT, _INT, C.c_float, C.c_uint64, _I64, _I64, C.POINTER(C.c_float),
                      C.POINTER(C.c_float)], _INT),
    "mk_mlp_eval": ([_P, _I64, _I64, C.POINTER(C.c_float)], _INT),
    "mk_davidson": ([_P, _P, _INT, _INT, _D, _P, _INT, _INT, C.POINTER(_D), C.POINTER(_D), C.POINTER(_INT),
                     C.POINTER(_INT)], _INT),
    "mk_prof_enable": ([_INT], _INT),
    "mk_prof_reset": ([], _INT),
    "mk_prof_query": ([C.c_char_p, C.POINTER(_D), C.POINTER(_I64)], _INT),
}

EXPORTED = tuple(_SIGS)

_lib = None


def load() -> C.CDLL:
    """Load (once) and return the library; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} not found: build it with `python mpskit.jl_amd/build.py` "
            "(there is no CPU fallback)")
    lib = C.CDLL(str(LIB_PATH), mode=os.RTLD_LOCAL)
    for name, (args, res) in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def check(status: int) -> None:
    if status != MK_OK:
        msg = load().mk_last_error().decode(errors="replace")
        if status == 1:
            raise SpaceMismatch(status, msg)
        if status == 5:
            raise NotConvergedError(status, msg)
        raise MKError(status, msg)
