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


def workspace(dtype, M, N, ks):
    import ctypes
    nb = ctypes.c_long()
    check(lib().rwig_gemm_workspace_bytes(dtype, M, N, ks, ctypes.byref(nb)))
    return torch.empty(nb.value, dtype=torch.uint8, device="cuda")


def run_bf16(A, Bt, M, N, K, ks=1, deferred=False):
    # A: (M, lda) bf16 row-major; Bt: (N, ldb) bf16 (column n of B is row n)
    D = torch.empty((N, M), dtype=torch.float32, device="cuda")
    w = workspace(0, M, N, ks) if deferred else None
    check(lib().rwig_gemm_tn_bf16_dev(M, N, K, A.data_ptr(), A.shape[1], Bt.data_ptr(),
                                      Bt.shape[1], D.data_ptr(), M, ks,
                                      None if w is None else w.data_ptr(),
                                      torch.cuda.current_stream().cuda_stream))
    return D


# (ks, deferred): fused epilogue; deferred epilogue; k split in 2 and 5.
# Rows 16-byte aligned take the direct-to-LDS path; when K is not a multiple
# of the k-step, its last, partial step goes through registers.
SPLITS = [(1, False), (1, True), (2, True), (5, True)]


@pytest.mark.parametrize("ks,deferred", SPLITS)
@pytest.mark.parametrize("M,N,K", [(16, 16, 32), (128, 128, 64), (300, 70, 200), (1000, 257, 999),
                                   (300, 70, 256), (129, 600, 64), (1000, 257, 1024)])
def test_bf16_gemm_matches_torch(M, N, K, ks, deferred):
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
    D = run_bf16(Ad, Bd, M, N, K, ks, deferred)
    torch.cuda.synchronize()
    ref = (Bt[:, :K].double() @ A[:, :K].double().T)  # (N, M)
    scale = (Bt[:, :K].double().abs() @ A[:, :K].double().abs().T).clamp_min(1e-30)
    rel = ((D.cpu().double() - ref).abs() / scale).max().item()
    assert rel < 1e-5, rel


@pytest.mark.parametrize("ks,deferred", SPLITS)
@pytest.mark.parametrize("M,N,K", [(16, 16, 4), (64, 64, 16), (300, 70, 201), (1000, 257, 999)])
def test_f64_gemm_matches_torch(M, N, K, ks, deferred):
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
    w = workspace(1, M, N, ks) if deferred else None
    check(lib().rwig_gemm_tn_f64_dev(M, N, K, Ad.data_ptr(), lda, Bd.data_ptr(), lda,
                                     D.data_ptr(), M, ks, None if w is None else w.data_ptr(),
                                     torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    ref = Bt[:, :K] @ A[:, :K].T
    scale = (Bt[:, :K].abs() @ A[:, :K].abs().T).clamp_min(1e-300)
    rel = ((D.cpu() - ref).abs() / scale).max().item()
    assert rel < 1e-13, rel


@pytest.mark.parametrize("ks,deferred", [(1, False), (2, True)])
def test_bf16_gemm_throughput_report(ks, deferred):
    M, N, K = 4096, 1024, 4096
    A = torch.randn((M, K), device="cuda").to(torch.bfloat16)
    Bt = torch.randn((N, K), device="cuda").to(torch.bfloat16)
    D = run_bf16(A, Bt, M, N, K, ks, deferred)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(20):
        run_bf16(A, Bt, M, N, K, ks, deferred)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / 20
    print(f"bf16 GEMM {M}x{N}x{K} ks={ks} deferred={deferred}: {dt*1e6:.1f} us, "
          f"{2*M*N*K/dt/1e12:.1f} TFLOP/s", flush=True)
    ref = (Bt[:64].float() @ A.float().T)
    assert torch.allclose(D[:64], ref, rtol=1e-2, atol=1e-1)


@pytest.mark.parametrize("ks,deferred", [(1, False), (2, True)])
@pytest.mark.parametrize("M,N,K", [(300, 70, 256), (1000, 257, 1024), (300, 70, 200),
                                   (1000, 257, 1000), (129, 65, 40)])
def test_bf16_direct_to_lds_equals_register_path(M, N, K, ks, deferred):
    # the same operands once 16-byte aligned (direct-to-LDS path, with a
    # partial last k-step when K % 64 != 0, only that step when K < 64) and
    # once shifted by 4 elements (rows no longer 16-byte aligned: register path)
    g = torch.Generator(device="cpu").manual_seed(7 * M + K)
    lda = K + 8
    A = torch.randn((M, lda), generator=g).to(torch.bfloat16).cuda()
    Bt = torch.randn((N, lda), generator=g).to(torch.bfloat16).cuda()
    Da = run_bf16(A, Bt, M, N, K, ks, deferred)
    As = torch.empty((M, lda + 8), dtype=torch.bfloat16, device="cuda")
    Bs = torch.empty((N, lda + 8), dtype=torch.bfloat16, device="cuda")
    As[:, 4:4 + lda] = A
    Bs[:, 4:4 + lda] = Bt
    D = torch.empty((N, M), dtype=torch.float32, device="cuda")
    w = workspace(0, M, N, ks) if deferred else None
    check(lib().rwig_gemm_tn_bf16_dev(M, N, K, As[:, 4:].data_ptr(), lda + 8,
                                      Bs[:, 4:].data_ptr(), lda + 8, D.data_ptr(), M, ks,
                                      None if w is None else w.data_ptr(),
                                      torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    assert torch.equal(Da, D)


@pytest.mark.parametrize("ks,deferred", [(1, False), (3, True)])
@pytest.mark.parametrize("M,N,K", [(300, 900, 256), (300, 900, 201), (257, 70, 10)])
def test_f64_direct_to_lds_equals_register_path(M, N, K, ks, deferred):
    # as above for FP64 (k-steps of 16): rows shifted by one element are no
    # longer 16-byte aligned and take the register path
    g = torch.Generator(device="cpu").manual_seed(5 * M + K)
    lda = K + 2 + K % 2
    A = torch.randn((M, lda), generator=g, dtype=torch.float64).cuda()
    Bt = torch.randn((N, lda), generator=g, dtype=torch.float64).cuda()
    Da = run_f64(A, Bt, M, N, K, ks, deferred)
    As = torch.empty((M, lda + 2), dtype=torch.float64, device="cuda")
    Bs = torch.empty((N, lda + 2), dtype=torch.float64, device="cuda")
    As[:, 1:1 + lda] = A
    Bs[:, 1:1 + lda] = Bt
    D = torch.empty((N, M), dtype=torch.float64, device="cuda")
    w = workspace(1, M, N, ks) if deferred else None
    check(lib().rwig_gemm_tn_f64_dev(M, N, K, As[:, 1:].data_ptr(), lda + 2,
                                     Bs[:, 1:].data_ptr(), lda + 2, D.data_ptr(), M, ks,
                                     None if w is None else w.data_ptr(),
                                     torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    assert torch.equal(Da, D)


def run_f64(A, Bt, M, N, K, ks=1, deferred=False):
    D = torch.empty((N, M), dtype=torch.float64, device="cuda")
    w = workspace(1, M, N, ks) if deferred else None
    check(lib().rwig_gemm_tn_f64_dev(M, N, K, A.data_ptr(), A.shape[1], Bt.data_ptr(),
                                     Bt.shape[1], D.data_ptr(), M, ks,
                                     None if w is None else w.data_ptr(),
                                     torch.cuda.current_stream().cuda_stream))
    return D


@pytest.mark.parametrize("ks,deferred", [(1, False), (1, True), (3, True)])
@pytest.mark.parametrize("dtype,K", [("bf16", 256), ("bf16", 200), ("f64", 256), ("f64", 201)])
def test_every_tile_shape_gives_the_same_columns(dtype, K, ks, deferred):
    """The tile shape follows the column count (smaller tiles for fewer
    columns). A product over the first n columns must equal those columns of
    the product over all of them bit for bit, for n in every shape's range.
    K = 256 takes the direct-to-LDS tiles in whole k-steps, K = 200 / 201 ends
    with a partial k-step."""
    M = 300
    g = torch.Generator(device="cpu").manual_seed(K + ks)
    if dtype == "bf16":
        Nall, subs, run = 700, (100, 200, 400), run_bf16
        lda = (K + 7) // 8 * 8
        A = torch.randn((M, lda), generator=g).to(torch.bfloat16).cuda()
        Bt = torch.randn((Nall, lda), generator=g).to(torch.bfloat16).cuda()
    else:
        Nall, subs, run = 900, (200, 400, 700), run_f64
        lda = (K + 1) // 2 * 2
        A = torch.randn((M, lda), generator=g, dtype=torch.float64).cuda()
        Bt = torch.randn((Nall, lda), generator=g, dtype=torch.float64).cuda()
    full = run(A, Bt, M, Nall, K, ks, deferred)
    for n in subs:
        part = run(A, Bt[:n].contiguous(), M, n, K, ks, deferred)
        torch.cuda.synchronize()
        assert torch.equal(part, full[:n]), (dtype, K, n)


def test_bf16_product_only_leaves_the_tiles_in_the_workspace():
    """D null with a workspace: only the product runs, and the workspace holds
    the FP32 tiles (column-major within each 128 x 128 tile at > 512 columns)
    that the deferred epilogue would have written to D. Without a workspace
    a null D is refused."""
    M, N, K = 256, 600, 256
    g = torch.Generator(device="cpu").manual_seed(7)
    A = torch.randn((M, K), generator=g).to(torch.bfloat16).cuda()
    Bt = torch.randn((N, K), generator=g).to(torch.bfloat16).cuda()
    D = run_bf16(A, Bt, M, N, K, 1, True)
    w = workspace(0, M, N, 1)
    st = torch.cuda.current_stream().cuda_stream
    check(lib().rwig_gemm_tn_bf16_dev(M, N, K, A.data_ptr(), K, Bt.data_ptr(), K, None, M, 1,
                                      w.data_ptr(), st))
    torch.cuda.synchronize()
    wf = w.view(torch.float32).cpu()
    gx, T = (M + 127) // 128, 128
    m = torch.arange(M).view(1, M)
    n = torch.arange(N).view(N, 1)
    idx = ((n // T) * gx + m // T) * T * T + (n % T) * T + m % T
    assert torch.equal(wf[idx], D.cpu())
    rc = lib().rwig_gemm_tn_bf16_dev(M, N, K, A.data_ptr(), K, Bt.data_ptr(), K, None, M, 1,
                                     None, st)
    assert rc != 0
