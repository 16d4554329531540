"""GPU tests of the fused soft-min passes (rwig_softmin_rows_dev /
rwig_softmin_cols_dev) against an FP64 torch reference,
softmin_x r = -reg * logsumexp(-r / reg).

Bars: FP64 1e-12 relative to the result's scale; FP32 (hardware exp)
2e-5 relative to the scale of R (the results are differences of O(|R|)
quantities; the exp's error enters through reg * log)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from rwig_amd._lib import check, lib  # noqa: E402


def workspace(which, M, N, S, prec):
    nb = ctypes.c_long()
    check(lib().rwig_softmin_workspace_bytes(which, M, N, S, prec, ctypes.byref(nb)))
    return torch.empty(nb.value, dtype=torch.uint8, device="cuda")


def problem(seed, M, N, S):
    g = torch.Generator().manual_seed(seed)
    C = torch.rand((M, N), generator=g, dtype=torch.float64)
    F = 0.1 * torch.randn((M, S), generator=g, dtype=torch.float64)
    G = 0.1 * torch.randn((N, S), generator=g, dtype=torch.float64)
    return C, F, G


def reference(C, F, G, reg):
    R = C[:, :, None] - F[:, None, :] - G[None, :, :]  # (M, N, S)
    rows = -reg * torch.logsumexp(-R / reg, dim=1)      # (M, S)
    cols = -reg * torch.logsumexp(-R / reg, dim=0)      # (N, S)
    return rows, cols, R.abs().max().item()


@pytest.mark.parametrize("prec,tol", [(0, 1e-12), (1, 2e-5)])
@pytest.mark.parametrize("M,N,S,reg", [(100, 70, 1, 0.05), (300, 257, 3, 0.01),
                                       (1000, 600, 8, 0.004), (64, 2000, 5, 0.1),
                                       (200, 150, 16, 0.02), (130, 90, 32, 0.05)])
def test_softmin_passes_match_reference(M, N, S, reg, prec, tol):
    C, F, G = problem(M + N + S, M, N, S)
    rows_ref, cols_ref, scale = reference(C, F, G, reg)
    dt = torch.float64 if prec == 0 else torch.float32
    # column-major device copies: a (rows x cols) column-major matrix is the
    # row-major (cols x rows) tensor
    Cd = C.t().contiguous().to(dt).cuda()     # C column-major, ldc = M
    CTd = C.contiguous().to(dt).cuda()        # C^T column-major (N x M), ldct = N
    Fd = F.t().contiguous().to(dt).cuda()     # M x S column-major
    Gd = G.t().contiguous().to(dt).cuda()     # N x S column-major
    out_r = torch.empty((S, M), dtype=dt, device="cuda")
    out_c = torch.empty((S, N), dtype=dt, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    wr, wc = workspace(0, M, N, S, prec), workspace(1, M, N, S, prec)
    check(lib().rwig_softmin_rows_dev(Cd.data_ptr(), M, N, M, Fd.data_ptr(), M, Gd.data_ptr(), N,
                                      S, reg, out_r.data_ptr(), prec, wr.data_ptr(), st))
    check(lib().rwig_softmin_cols_dev(CTd.data_ptr(), M, N, N, Fd.data_ptr(), M, Gd.data_ptr(),
                                      N, S, reg, out_c.data_ptr(), prec, wc.data_ptr(), st))
    torch.cuda.synchronize()
    er = (out_r.double().cpu().t() - rows_ref).abs().max().item()
    ec = (out_c.double().cpu().t() - cols_ref).abs().max().item()
    assert er <= tol * scale, (er, scale)
    assert ec <= tol * scale, (ec, scale)


def test_invalid_arguments():
    with pytest.raises(Exception):
        check(lib().rwig_softmin_rows_dev(None, 10, 10, 10, None, 10, None, 10, 40, 0.1, None, 0,
                                          None, None))
