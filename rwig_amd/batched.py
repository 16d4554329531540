"""Batched one-source-many-targets Sinkhorn on the device
(rwig_sinkhorn_batched_dev in include/rwig_hip.h).

torch provides the device buffers and the stream; the solve itself runs in
librwig_hip.so. The reference solves such a batch as one sinkhorn_log_cpp call
per target (src/sinkhorn.cpp:142); each target here gives the same result as
that call on its reduced problem (the support of b), solved side by side.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from ._lib import check, lib

PRECISIONS = {"fp64": 0, "fp32": 1, "bf16": 2}
ERR_MODES = {"reference": 0, "l1": 1}


def _ptr(t):
    return None if t is None else t.data_ptr()


def _stream_ptr(stream):
    if stream is None:
        stream = torch.cuda.current_stream()
    return stream.cuda_stream


class TargetBatch:
    """Ragged target supports in the layout the kernels read (device tensors).

    supports: list of index arrays into the columns of C; weights: list of
    positive mass arrays of the same lengths."""

    def __init__(self, supports, weights, device="cuda"):
        if len(supports) != len(weights):
            raise ValueError("one weight array per support")
        sizes = [len(s) for s in supports]
        if min(sizes) < 1:
            raise ValueError("every target needs a non-empty support")
        ptr = np.zeros(len(sizes) + 1, dtype=np.int32)
        ptr[1:] = np.cumsum(sizes)
        idx = np.concatenate([np.asarray(s, dtype=np.int32) for s in supports])
        val = np.concatenate([np.asarray(w, dtype=np.float64) for w in weights])
        if np.any(val <= 0):
            raise ValueError("target masses on the support must be positive")
        self.B = len(sizes)
        self.max_nt = int(max(sizes))
        self.nnz = int(ptr[-1])
        self.col_ptr = torch.as_tensor(ptr, device=device)
        self.col_idx = torch.as_tensor(idx, device=device)
        self.tgt_of = torch.as_tensor(np.repeat(np.arange(self.B, dtype=np.int32), sizes),
                                      device=device)
        self.b = torch.as_tensor(val, device=device)
        self.offsets = ptr


def gibbs_lowp(C, reg, precision, stream=None):
    """Low-precision Gibbs kernel exp(-C/reg) for precision 'fp32' or 'bf16'.

    Returns (tensor, ldk); rows are padded to a multiple of 8."""
    M, N = C.shape[0], C.shape[1]
    ldk = (M + 7) // 8 * 8
    dt = torch.float32 if precision == "fp32" else torch.int16
    K = torch.empty((N, ldk), dtype=dt, device=C.device)  # column-major M x N
    check(lib().rwig_gibbs_lowp_dev(_ptr(C), C.stride(1), M, N, float(reg),
                                    PRECISIONS[precision], _ptr(K), ldk, _stream_ptr(stream)))
    return K, ldk


def sinkhorn_batched(C, a, targets: TargetBatch, reg, precision="fp64", max_iter=1000,
                     zero_tol=1e-6, err="reference", absorb_every=20, Klo=None,
                     with_loss=True, stream=None, out=None):
    """Solve the batch; C is a column-major (Fortran-ordered) float64 device
    tensor of shape (M, N) i.e. C.stride() == (1, M); a is (M,)."""
    if C.dtype != torch.float64 or C.stride(0) != 1:
        raise ValueError("C must be a column-major float64 device tensor")
    M = C.shape[0]
    prec = PRECISIONS[precision]
    if prec != 0 and Klo is None:
        Klo = gibbs_lowp(C, reg, precision, stream)
    kt, ldk = (None, 0) if prec == 0 else Klo
    dev = C.device
    if out is None:
        out = {
            "f": torch.empty((targets.B, M), dtype=torch.float64, device=dev),
            "g": torch.empty(targets.nnz, dtype=torch.float64, device=dev),
            "iter": torch.empty(targets.B, dtype=torch.int32, device=dev),
            "err": torch.empty(targets.B, dtype=torch.float64, device=dev),
            "loss": torch.empty((targets.B, 2), dtype=torch.float64, device=dev)
            if with_loss else None,
        }
    check(lib().rwig_sinkhorn_batched_dev(
        _ptr(C), C.stride(1), M, _ptr(a), targets.B, _ptr(targets.col_ptr),
        _ptr(targets.col_idx), _ptr(targets.b), targets.max_nt, float(reg), int(max_iter),
        float(zero_tol), ERR_MODES[err], prec, _ptr(kt), ldk, int(absorb_every),
        _ptr(out["f"]), _ptr(out["g"]), _ptr(out["iter"]), _ptr(out["err"]),
        _ptr(out["loss"]), _stream_ptr(stream)))
    return out


def return_status(out, max_iter, zero_tol):
    """Per-target return_status of a finished batched solve, as the reference's
    shim computes it (src/sinkhorn.cpp:88-95): 0 where err <= zero_tol, 1 where
    the target ran max_iter iterations, else 2 (a target whose error turned
    non-finite stops and lands here). Call after the solve's stream has
    finished."""
    err, it = out["err"], out["iter"]
    st = torch.full_like(it, 2)
    st = torch.where(it >= int(max_iter), torch.ones_like(st), st)
    return torch.where(err <= float(zero_tol), torch.zeros_like(st), st)


class DenseKernel:
    """exp(-C/reg) in BF16 and FP64, both orientations, and exp(-C/reg) .* C in
    FP64, for the MFMA path (rwig_dense_kernel_create). Built once per cost
    matrix and reg."""

    def __init__(self, C, reg, stream=None):
        if C.dtype != torch.float64 or C.stride(0) != 1:
            raise ValueError("C must be a column-major float64 device tensor")
        self.M, self.Nc = C.shape
        self.reg = float(reg)
        h = ctypes.c_void_p()
        check(lib().rwig_dense_kernel_create(_ptr(C), C.stride(1), self.M, self.Nc, self.reg,
                                             _stream_ptr(stream), ctypes.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            check(lib().rwig_dense_kernel_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sinkhorn_batched_mfma(C, a, targets: TargetBatch, reg, kernel: DenseKernel = None,
                          max_iter=1000, zero_tol=1e-6, err="l1", absorb_every=20,
                          switch_thr=1e-2, with_loss=True, stream=None, out=None):
    """Batched solve with every Gibbs product on BF16 MFMA (all targets as GEMM
    columns), FP64 absorption every absorb_every iterations."""
    if kernel is None:
        kernel = DenseKernel(C, reg, stream)
    M = C.shape[0]
    dev = C.device
    if out is None:
        out = {
            "f": torch.empty((targets.B, M), dtype=torch.float64, device=dev),
            "g": torch.empty(targets.nnz, dtype=torch.float64, device=dev),
            "iter": torch.empty(targets.B, dtype=torch.int32, device=dev),
            "err": torch.empty(targets.B, dtype=torch.float64, device=dev),
            "loss": torch.empty((targets.B, 2), dtype=torch.float64, device=dev)
            if with_loss else None,
        }
    check(lib().rwig_sinkhorn_batched_mfma_dev(
        kernel.handle, _ptr(a), targets.B, _ptr(targets.col_ptr),
        _ptr(targets.col_idx), _ptr(targets.tgt_of), targets.nnz, _ptr(targets.b), float(reg),
        int(max_iter), float(zero_tol), ERR_MODES[err], int(absorb_every), float(switch_thr),
        _ptr(out["f"]), _ptr(out["g"]), _ptr(out["iter"]), _ptr(out["err"]), _ptr(out["loss"]),
        _stream_ptr(stream)))
    return out
