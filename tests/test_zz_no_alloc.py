"""rskx_rskelf_sv_noalloc (work rows in the library's static device array)
against rskx_rskelf_sv, through the C ABI: two calls in a row on one GPU,
the second on a smaller problem, must each give rskx_rskelf_sv's result
bit for bit (same kernels, same data); a sweep beyond the array's size is
refused before anything runs."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _lib():
    from flam_amd import _lib as L
    return L


def _factor(N, prec):
    import flam_amd
    from flam_amd import geometry
    g = geometry.circle(N)
    return flam_amd.rskelf(g.kernel, g.x, 64, 1e-8, g.proxy, dict(symm="p", precision=prec))


def _both(F, B, dtype):
    """(rskx_rskelf_sv result, rskx_rskelf_sv_noalloc result) of B (N x nrhs)."""
    L = _lib()
    lib = L.lib()
    lib.rskx_set_stream(torch.cuda.current_stream().cuda_stream)
    code = L.RSKX_F64 if B.dtype == torch.float64 else L.RSKX_F32
    X1 = B.t().contiguous()                      # column-major N x nrhs
    X2 = B.t().contiguous()
    N, nrhs = B.shape
    assert lib.rskx_rskelf_sv(F.handle, X1.data_ptr(), N, nrhs, b"n", code) == 0, L.last_error()
    rc = lib.rskx_rskelf_sv_noalloc(F.handle, X2.data_ptr(), N, nrhs, b"n", code)
    assert rc == 0, L.last_error()
    torch.cuda.synchronize()
    return X1.t(), X2.t()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_noalloc_matches_rskelf_sv(prec):
    print(f"noalloc solve: {prec} circle N=16384 nrhs 8, then N=3000 nrhs 3", flush=True)
    dt = torch.float64 if prec == "fp64" else torch.float32
    rng = np.random.default_rng(0)
    F1 = _factor(16384, prec)
    B1 = torch.from_numpy(rng.standard_normal((16384, 8))).to(dt).cuda()
    a, b = _both(F1, B1, dt)
    assert torch.equal(a, b)
    F2 = _factor(3000, prec)                     # second call: smaller input
    B2 = torch.from_numpy(rng.standard_normal((3000, 3))).to(dt).cuda()
    a, b = _both(F2, B2, dt)
    assert torch.equal(a, b)
    # the first factor again after the smaller call left its rows behind
    a, b = _both(F1, B1[:, :5].contiguous(), dt)
    assert torch.equal(a, b)


def test_noalloc_refuses_beyond_its_array():
    print("noalloc solve: a sweep beyond the static array is refused", flush=True)
    L = _lib()
    lib = L.lib()
    N = 1 << 18
    F = _factor(N, "fp64")
    nrhs = 40                                    # 2^18 x 40 doubles = 80 MiB > 64 MiB
    X = torch.zeros(nrhs, N, dtype=torch.float64, device="cuda")
    X[0, 0] = 7.0
    rc = lib.rskx_rskelf_sv_noalloc(F.handle, X.data_ptr(), N, nrhs, b"n", L.RSKX_F64)
    assert rc == L.RSKX_ERR_MEM and L.last_error().startswith("FLAM:rskx:scratch")
    torch.cuda.synchronize()
    assert X[0, 0].item() == 7.0 and int((X != 0).sum()) == 1
