"""GPU tests of the MFMA GEMM (rwig_gemm_tn_*_dev) against torch.

BF16 inputs: the reference is the FP32 product of the same BF16-rounded
inputs, tolerance 1e-5 relative to the row scale (only the accumulation order
differs). FP64: tolerance 1e-13 relative to the row scale.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from rwig_amd._lib import check, lib  # noqa: E402


def run_bf16(A, Bt, M, N, K):
    # A: (M, lda) bf16 row-major; Bt: (N, ldb) bf16 (column n of B is row n)
    D = torch.empty((N, M), dtype=torch.float32, device="cuda")
    check(lib().rwig_gemm_tn_bf16_dev(M, N, K, A.data_ptr(), A.shape[1], Bt.data_ptr(),
                                      Bt.shape[1], D.data_ptr(), M,
                                      torch.cuda.current_stream().cuda_stream))
    return D


@pytest.mark.parametrize("M,N,K", [(16, 16, 32), (128, 128, 64), (300, 70, 200), (1000, 257, 999)])
def test_bf16_gemm_matches_torch(M, N, K):
    g = torch.Generator(device="cpu").manual_seed(M + N + K)
    lda = (K + 7) // 8 * 8
    A = torch.zeros((M, lda), dtype=torch.bfloat16)
    A[:, :K] = torch.randn((M, K), generator=g).to(torch.bfloat16)
    Bt = torch.zeros((N, lda), dtype=torch.bfloat16)
    Bt[:, :K] = torch.randn((N, K), generator=g).to(torch.bfloat16)
    # poison the padding: the kernel must ignore k >= K
    A[:, K:] = 7.0
    Bt[:, K:] = 7.0
    Ad, Bd = A.cuda(), Bt.cuda()
    D = run_bf16(Ad, Bd, M, N, K)
    torch.cuda.synchronize()
    ref = (Bt[:, :K].double() @ A[:, :K].double().T)  # (N, M)
    scale = (Bt[:, :K].double().abs() @ A[:, :K].double().abs().T).clamp_min(1e-30)
    rel = ((D.cpu().double() - ref).abs() / scale).max().item()
    assert rel < 1e-5, rel


@pytest.mark.parametrize("M,N,K", [(16, 16, 4), (64, 64, 16), (300, 70, 201), (1000, 257, 999)])
def test_f64_gemm_matches_torch(M, N, K):
    g = torch.Generator(device="cpu").manual_seed(M * 3 + N + K)
    lda = (K + 1) // 2 * 2
    A = torch.zeros((M, lda), dtype=torch.float64)
    A[:, :K] = torch.randn((M, K), generator=g, dtype=torch.float64)
    Bt = torch.zeros((N, lda), dtype=torch.float64)
    Bt[:, :K] = torch.randn((N, K), generator=g, dtype=torch.float64)
    A[:, K:] = 7.0
    Bt[:, K:] = 7.0
    Ad, Bd = A.cuda(), Bt.cuda()
    D = torch.empty((N, M), dtype=torch.float64, device="cuda")
    check(lib().rwig_gemm_tn_f64_dev(M, N, K, Ad.data_ptr(), lda, Bd.data_ptr(), lda,
                                     D.data_ptr(), M, torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    ref = Bt[:, :K] @ A[:, :K].T
    scale = (Bt[:, :K].abs() @ A[:, :K].abs().T).clamp_min(1e-300)
    rel = ((D.cpu() - ref).abs() / scale).max().item()
    assert rel < 1e-13, rel


def test_bf16_gemm_throughput_report():
    M, N, K = 4096, 1024, 4096
    A = torch.randn((M, K), device="cuda").to(torch.bfloat16)
    Bt = torch.randn((N, K), device="cuda").to(torch.bfloat16)
    D = run_bf16(A, Bt, M, N, K)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(20):
        run_bf16(A, Bt, M, N, K)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / 20
    print(f"bf16 GEMM {M}x{N}x{K}: {dt*1e6:.1f} us, {2*M*N*K/dt/1e12:.1f} TFLOP/s", flush=True)
    ref = (Bt[:64].float() @ A.float().T)
    assert torch.allclose(D[:64], ref, rtol=1e-2, atol=1e-1)
