"""ctypes binding of librwig_hip.so (declarations in include/rwig_hip.h).

The library is the product: there is no CPU fallback. Loading fails loudly
when the shared object has not been built.
"""
from __future__ import annotations

import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librwig_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "rwig_hip.h")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_d = ctypes.c_double
_i = ctypes.c_int
_vp = ctypes.c_void_p

# argument types of every exported entry point
SIGNATURES = {
    "rwig_available": ([], _i),
    "rwig_last_error": ([], ctypes.c_char_p),
    "rwig_sinkhorn_vanilla": ([_dp, _dp, _dp, _dp, _dp, _ip, _dp, _dp, _dp, _dp, _i, _i, _d,
                               _i, _i, _d, _ip], _i),
    "rwig_sinkhorn_log": ([_dp, _dp, _dp, _dp, _dp, _ip, _dp, _dp, _dp, _dp, _i, _i, _d, _i,
                           _i, _d, _ip], _i),
    "rwig_barycenter_parallel": ([_dp, _dp, _dp, _dp, _dp, _dp, _ip, _dp, _dp, _dp, _dp, _dp,
                                  _i, _i, _i, _d, _i, _i, _d, _ip], _i),
    "rwig_barycenter_log": ([_dp, _dp, _dp, _dp, _dp, _dp, _ip, _dp, _dp, _dp, _dp, _dp, _i,
                             _i, _i, _d, _i, _i, _d, _ip], _i),
    "rwig_barycenter_log_ck": ([_dp, _dp, _dp, _dp, _dp, _dp, _ip, _dp, _dp, _dp, _dp, _dp,
                                _i, _i, _i, _d, _i, _i, _d, _i, _i, _i, _ip], _i),
    "rwig_barycenter_log_ck_dev": ([_vp, _vp, _vp, _vp, _vp, _vp, _ip, _dp, _vp, _vp, _vp, _vp,
                                    _i, _i, _i, _d, _i, _i, _d, _i, _i, _i, _ip, _vp], _i),
    "rwig_barycenter_log_batched": ([_dp, _ip, _dp, _dp, _dp, _dp, _i, _i, _i, _i, _d, _i,
                                     _d], _i),
    "rwig_sinkhorn_batched_dev": ([_vp, ctypes.c_long, _i, _vp, _i, _vp, _vp, _vp, _i, _d, _i,
                                   _d, _i, _i, _vp, ctypes.c_long, _i, _vp, _vp, _vp, _vp, _vp,
                                   _vp], _i),
    "rwig_sinkhorn_log_dev": ([_vp, _vp, _vp, _i, _i, _d, _i, _i, _d, _vp, _vp, _vp, _vp, _vp,
                               _ip, _dp, _vp], _i),
    "rwig_gemm_workspace_bytes": ([_i, _i, _i, _i, ctypes.POINTER(ctypes.c_long)], _i),
    "rwig_gemm_tn_bf16_dev": ([_i, _i, _i, _vp, ctypes.c_long, _vp, ctypes.c_long, _vp,
                               ctypes.c_long, _i, _vp, _vp], _i),
    "rwig_gemm_tn_f64_dev": ([_i, _i, _i, _vp, ctypes.c_long, _vp, ctypes.c_long, _vp,
                              ctypes.c_long, _i, _vp, _vp], _i),
    "rwig_dense_kernel_create": ([_vp, ctypes.c_long, _i, _i, _d, _vp,
                                  ctypes.POINTER(ctypes.c_void_p)], _i),
    "rwig_dense_kernel_destroy": ([_vp], _i),
    "rwig_sinkhorn_batched_mfma_dev": ([_vp, _vp, _i, _vp, _vp, _vp, _i, _vp, _d, _i, _d, _i,
                                        _i, _d, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    "rwig_wdl": ([_dp, _dp, _dp, _dp, _dp, _i, _i, _i, _d, _i, _d, _i, _i, _i, _d, _d, _d, _d,
                  _d, _i, _i], _i),
    "rwig_wdl_init": ([_dp, _dp, _dp, _dp, _dp, _i, _i, _i, _d, _i, _d, _i, _i, _i, _d, _d, _d,
                       _d, _d, _dp, _dp, _i, _i], _i),
    "rwig_wdl_batch_grad_dev": ([_vp, _vp, ctypes.c_long, _i, _i, _i, _i, _vp, _vp, _vp, _vp,
                                 _vp, _vp, _i, _i, _vp], _i),
    "rwig_wdl_step_dev": ([_vp, _vp, ctypes.c_long, _vp, _i, _i, _i, _i, _i, _i, _vp, _vp, _vp,
                           _vp, _vp, _vp, _i, _i, _d, _d, _d, _d, _d, _vp, _i, _i, _vp], _i),
    "rwig_gibbs_lowp_dev": ([_vp, ctypes.c_long, _i, _i, _d, _i, _vp, ctypes.c_long, _vp], _i),
    "rwig_euclidean": ([_dp, _i, _i, _dp, _i], _i),
    "rwig_euclidean_dev": ([_vp, _i, _i, ctypes.c_long, _vp, ctypes.c_long, _i, _vp], _i),
    "rwig_doc2dist": ([ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_long), _i,
                       ctypes.POINTER(ctypes.c_char_p), _i, _dp], _i),
    "rwig_tsvd": ([_dp, _ip, _dp, _dp, _i, _i, _i, _i], _i),
    "rwig_wig_scores": ([_dp, _dp, _dp, _dp, _i, _i, _i, _i], _i),
    "rwig_wig_index": ([_ip, _dp, _ip, _dp, _ip, _i, _i, _i], _i),
    "rwig_softmin_workspace_bytes": ([_i, _i, _i, _i, _i, ctypes.POINTER(ctypes.c_long)], _i),
    "rwig_softmin_rows_dev": ([_vp, _i, _i, ctypes.c_long, _vp, ctypes.c_long, _vp,
                               ctypes.c_long, _i, _d, _vp, _i, _vp, _vp], _i),
    "rwig_softmin_rows_noalloc_dev": ([_vp, _i, _i, ctypes.c_long, _vp, ctypes.c_long, _vp,
                                       ctypes.c_long, _i, _d, _vp, _i, _vp], _i),
    "rwig_softmin_cols_dev": ([_vp, _i, _i, ctypes.c_long, _vp, ctypes.c_long, _vp,
                               ctypes.c_long, _i, _d, _vp, _i, _vp, _vp], _i),
    "rwig_serve_create": ([_d, _i, _i, _d, _i, ctypes.POINTER(ctypes.c_void_p)], _i),
    "rwig_serve_destroy": ([_vp], _i),
    "rwig_serve_footprint": ([_i, _i, _i, _dp, _dp], _i),
    "rwig_serve_register_cost": ([_vp, _dp, _i, _i, _ip], _i),
    "rwig_serve_submit_sinkhorn": ([_vp, _i, _d, _dp, _i, _dp, _i, _i, _d,
                                    ctypes.POINTER(ctypes.c_long)], _i),
    "rwig_serve_submit_barycenter": ([_vp, _i, _d, _dp, _i, _i, _dp, _i, _d,
                                      ctypes.POINTER(ctypes.c_long)], _i),
    "rwig_serve_wait": ([_vp, ctypes.c_long, _d, _ip], _i),
    "rwig_serve_result": ([_vp, ctypes.c_long, _dp, _dp, _dp, _ip, _dp, _dp, _ip], _i),
    "rwig_serve_stats": ([_vp, _dp, _i], _i),
    "rwig_serve_reset_stats": ([_vp], _i),
}

_lib = None


class RwigError(RuntimeError):
    """A non-zero status from librwig_hip.so."""

    def __init__(self, status, message):
        super().__init__(f"librwig_hip status {status}: {message}")
        self.status = status


def header_symbols():
    """Names of the functions include/rwig_hip.h declares."""
    with open(HEADER) as fh:
        text = fh.read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*|void)\s*\*?\s*(rwig_\w+)\s*\(",
                                 text, flags=re.M)))


def lib():
    """Load librwig_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m rwig_amd.build`")
        # The PyTorch ROCm wheel ships its own HIP and HSA runtimes. The HSA
        # runtime has the same soname as the one this library links, so the
        # process keeps whichever is loaded first. Torch's HIP does not run on
        # the system's newer HSA runtime (torch then finds no GPU), while this
        # library runs on torch's. Where torch is installed it is therefore
        # imported before the library is loaded, whatever the caller's import
        # order; without torch nothing changes.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        handle = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.argtypes = args
            fn.restype = res
        _lib = handle
    return _lib


def check(status):
    if status != 0:
        msg = lib().rwig_last_error()
        raise RwigError(status, msg.decode() if msg else "")
