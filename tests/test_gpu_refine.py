"""The refinement's fused orthogonalization (rskx_cgs2) against a plain
FP64 torch restatement of classical Gram-Schmidt applied twice."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nrhs,nv,N", [(8, 1, 1 << 16), (3, 5, 100003), (1, 17, 4096), (2, 32, 777)])
def test_cgs2_matches_torch(nrhs, nv, N):
    from flam_amd.refine import _cgs2
    print(f"cgs2 nrhs={nrhs} nv={nv} N={N}", flush=True)
    g = torch.Generator(device="cpu").manual_seed(nv)
    nb = nv + 3                                   # basis storage longer than the used part
    V = torch.randn(nrhs, nb, N, dtype=torch.float64, generator=g)
    for r in range(nrhs):                         # orthonormal columns, as Arnoldi keeps them
        q, _ = torch.linalg.qr(V[r, :nv].t())
        V[r, :nv] = q.t()
    w = torch.randn(nrhs, N, dtype=torch.float64, generator=g)
    Vd, wd = V.cuda(), w.clone().cuda()
    hout = _cgs2(Vd, nv, wd).cpu()
    Vj = V[:, :nv]
    h1 = torch.einsum("rin,rn->ri", Vj, w)
    w1 = w - torch.einsum("rin,ri->rn", Vj, h1)
    h2 = torch.einsum("rin,rn->ri", Vj, w1)
    w2 = w1 - torch.einsum("rin,ri->rn", Vj, h2)
    scale = torch.linalg.vector_norm(w, dim=1).max().item()
    assert torch.allclose(hout[:, :nv], h1 + h2, rtol=0, atol=1e-12 * scale)
    assert torch.allclose(hout[:, nv], torch.linalg.vector_norm(w2, dim=1), rtol=1e-12)
    assert torch.allclose(wd.cpu(), w2, rtol=0, atol=1e-12 * scale)
    # the result is orthogonal to the basis
    assert torch.einsum("rin,rn->ri", Vj, wd.cpu()).abs().max().item() <= 1e-10 * scale
