"""GPU test of rwig_softmin_rows_noalloc_dev, the soft-min row pass that
allocates nothing (its chunk partials live in a static device array): its
output must equal rwig_softmin_rows_dev's bit for bit (same kernels, same
chunking), on two calls in a row, the second smaller, in FP64 and FP32; and a
pass whose partials exceed the array is refused with a clear error."""
import ctypes

import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from rwig_amd._lib import RwigError, check, lib  # noqa: E402


def _problem(seed, M, N, S, dt):
    g = torch.Generator().manual_seed(seed)
    C = torch.rand((N, M), generator=g, dtype=torch.float64)  # C column-major (M x N)
    F = 0.1 * torch.randn((S, M), generator=g, dtype=torch.float64)
    G = 0.1 * torch.randn((S, N), generator=g, dtype=torch.float64)
    return C.to(dt).cuda(), F.to(dt).cuda(), G.to(dt).cuda()


def _with_workspace(C, F, G, M, N, S, reg, prec, st):
    nb = ctypes.c_long()
    check(lib().rwig_softmin_workspace_bytes(0, M, N, S, prec, ctypes.byref(nb)))
    work = torch.empty(nb.value, dtype=torch.uint8, device="cuda")
    out = torch.full((S, M), float("nan"), dtype=C.dtype, device="cuda")
    check(lib().rwig_softmin_rows_dev(C.data_ptr(), M, N, M, F.data_ptr(), M, G.data_ptr(), N,
                                      S, reg, out.data_ptr(), prec, work.data_ptr(), st))
    return out


def _no_alloc(C, F, G, M, N, S, reg, prec, st):
    out = torch.full((S, M), float("nan"), dtype=C.dtype, device="cuda")
    check(lib().rwig_softmin_rows_noalloc_dev(C.data_ptr(), M, N, M, F.data_ptr(), M,
                                              G.data_ptr(), N, S, reg, out.data_ptr(), prec, st))
    return out


@pytest.mark.parametrize("prec", [0, 1])
def test_no_alloc_pass_equals_the_workspace_pass_on_two_calls(prec):
    dt = torch.float64 if prec == 0 else torch.float32
    st = torch.cuda.current_stream().cuda_stream
    # the first call at the C3 shape (the scratch array's sizing case), the
    # second smaller, so stale partials of the first would show
    for (M, N, S, reg) in [(4096, 4096, 8, 0.004), (300, 257, 3, 0.01)]:
        print(f"no-alloc soft-min rows: M={M} N={N} S={S} precision={prec}", flush=True)
        C, F, G = _problem(M + N + S, M, N, S, dt)
        ref = _with_workspace(C, F, G, M, N, S, reg, prec, st)
        got = _no_alloc(C, F, G, M, N, S, reg, prec, st)
        torch.cuda.synchronize()
        assert torch.isfinite(ref).all()
        assert torch.equal(got, ref)


def test_no_alloc_pass_refuses_partials_beyond_the_array():
    # FP64, S = 32 at M = N = 4096 needs 64 MiB of partials: refused before
    # any launch (no buffers are touched, so null pointers of the right kind
    # are not needed; the size check comes after the argument check)
    M = N = 4096
    S = 32
    C = torch.zeros(1, dtype=torch.float64, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    with pytest.raises(RwigError, match="static scratch array"):
        check(lib().rwig_softmin_rows_noalloc_dev(C.data_ptr(), M, N, M, C.data_ptr(), M,
                                                  C.data_ptr(), N, S, 0.01, C.data_ptr(), 0, st))
