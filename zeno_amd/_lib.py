"""Loader for zeno_amd/libzeno_hip.so (the HIP backend and its host side).

The library is built in-tree for gfx950 (zeno_amd/csrc/Makefile).  There is
no fallback: if it is missing, importing the backend raises.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libzeno_hip.so")
CSRC = os.path.join(_HERE, "csrc")

ZB_OK, ZB_ERR_ARG, ZB_ERR_UNSUPPORTED, ZB_ERR_MEMORY, ZB_ERR_DEVICE = 0, -1, -2, -3, -4
KERNEL_NAMES = {0: "k_zero", 1: "k_1q", 2: "k_cx", 3: "k_diag1", 4: "k_diagk", 5: "k_window",
                **{10 + k: f"k_kq<{k}>" for k in range(2, 7)},
                **{20 + k: f"k_kq_mfma<{k}>" for k in range(4, 7)}}

# Every exported symbol of include/zeno_hip.h with its ctypes signature.
P, U64, I64, I32, U32, D = C.c_void_p, C.c_uint64, C.c_int64, C.c_int, C.c_uint32, C.c_double
SIGNATURES = {
    "zb_create": (I32, [U32, I32, C.POINTER(C.c_void_p)]),
    "zb_destroy": (None, [P]),
    "zb_name": (C.c_char_p, [P]),
    "zb_apply_unitary": (I32, [P, P, U32, P]),
    "zb_apply_diagonal": (I32, [P, P, U32, P]),
    "zb_apply_cx": (I32, [P, U32, U32]),
    "zb_prob_one": (D, [P, U32]),
    "zb_measure": (I32, [P, U32, D]),
    "zb_reset_all": (I32, [P]),
    "zb_sample": (I32, [P, U64, U64, P]),
    "zb_sample_noalloc": (I32, [P, U64, U64, P]),
    "zb_statevector": (I32, [P, P]),
    "zb_norm_sqr": (D, [P]),
    "zb_finish": (I32, [P]),
    "zb_last_error": (C.c_char_p, []),
    "zb_state_raw": (I32, [P, P, P]),
    "zb_state_set_raw": (I32, [P, P, P]),
    "zb_state_device_ptrs": (I32, [P, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    "zb_n_qubits": (U32, [P]),
    "zb_precision": (I32, [P]),
    "zb_chunk_sums": (I32, [P, P]),
    "zb_set_mode": (I32, [P, I32]),
    "zb_stream": (P, [P]),
    "zb_max_abs_diff": (I32, [P, P, P]),
    "zb_profile_begin": (I32, [P]),
    "zb_tune": (I32, [P, C.c_char_p, I64]),
    "zb_tune_default": (I32, [U32, I32, C.c_char_p, I64]),
    "zb_profile_end": (I64, [P, P, P, P, I64]),
    "zb_profile_detail": (I64, [P, P, I64]),
    "zr_compile": (P, [P, I64, U32, P, U32, I32, I32]),
    "zr_compiled_free": (None, [P]),
    "zr_compiled_nops": (I64, [P]),
    "zr_compiled_op": (I32, [P, I64, P, P, P, P, P, P]),
    "zr_compiled_stats": (None, [P, P]),
    "zr_compiled_final_measures": (I64, [P, P]),
    "zr_apply_ops": (I32, [P, P, I64, I64]),
    "zr_run": (I64, [P, I32, U64, U64, P, I32, P, P, I64, P]),
    "zr_budget_bytes": (U64, [U64, D]),
    "zr_plan": (I32, [U32, I32, U64, P]),
    "zr_resolve_precision": (I32, [U32, I32, U64, P, P, C.c_char_p, C.c_size_t]),
    "zb_device_count": (I32, []),
    "zb_mem_info": (I32, [P, P]),
    "zb_calibration_copy": (I32, [U64, I32, P]),
    "zb_copy_probe": (I32, [U64, I32, I32, P]),
    "zb_layout": (I32, [P, P]),
    "zb_create_bound": (I32, [U32, I32, P, P, P]),
    "zb_rebind": (I32, [P, U32, I32, P, P]),
    "zs_create": (I32, [U64, I32, I32, P]),
    "zs_submit": (I64, [P, P, I32, U64, U64, P]),
    "zs_wait": (I64, [P, I64, P, P, I64, P]),
    "zs_status": (None, [P, P]),
    "zs_set_batch": (I32, [P, I32]),
    "zs_counters": (None, [P, P]),
    "zs_destroy": (None, [P]),
    "zr_parse_qasm": (I64, [C.c_char_p, P, I64, P, P, P, U32]),
    "zr_trajectories_supported": (I32, [P]),
    "zr_batch_supported": (I32, [P, I32, U64]),
    "zr_run_batch": (I32, [P, P, P, P, I64, P, P]),
    "zr_run_trajectories": (I32, [P, I32, U64, U64, P, P, P]),
    "zr_gen_random": (I64, [U32, U32, U64, P, I64]),
    "zr_gen_qft": (I64, [U32, P, I64]),
    "zr_gen_u3_qft": (I64, [U32, U64, P, I64]),
    "zr_gen_matching": (I64, [U32, U32, U64, U32, P, I64]),
    "zr_gate_build": (I32, [C.c_char_p, P, I32, P, P]),
}


def build() -> str:
    """Compile every HIP/C++ source for gfx950 into zeno_amd/libzeno_hip.so."""
    subprocess.run(["make", "-s", "-B", "-C", CSRC, "-j8"], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is not built; run `make -C {CSRC}` (or __graft_entry__.build()) first"
            )
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    return (lib().zb_last_error() or b"").decode()


def ptr(a):
    return a.ctypes.data_as(C.c_void_p)
