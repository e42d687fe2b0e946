"""The refinement's fused orthogonalization (rskx_cgs2) against a plain
FP64 torch restatement of classical Gram-Schmidt applied twice."""
import math

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


def test_closed_form_inner_operator():
    """The unit circle's closed form A = a I + b 1 1' agrees with the FFT
    product of the entry-wise matrix to ~1e-11 at N = 2^20, refuses a
    non-circle curve, and as FGMRES's inner operator still yields a true
    relative residual (FFT product) <= 1e-10."""
    import flam_amd
    from flam_amd import geometry
    from flam_amd.refine import fgmres
    N = 1 << 20
    print(f"closed-form inner operator: circle N={N}", flush=True)
    g = geometry.circle(N)
    X = torch.from_numpy(np.random.default_rng(8).standard_normal((N, 4))).cuda()
    Y1, Y2 = g.closed_form_matvec(X), g.circulant_matvec(X)
    rel = (torch.linalg.vector_norm(Y1 - Y2, dim=0) / torch.linalg.vector_norm(Y2, dim=0)).max()
    assert rel.item() < 1e-10
    e = geometry.ellipse(4096, 1.0, 0.5)
    with pytest.raises(ValueError):
        e.closed_form_matvec(X[:4096])
    F = flam_amd.rskelf(g.kernel, g.x, 256, 1e-6, g.proxy, dict(symm="p", precision="fp32"))
    Xs, info = fgmres(g.circulant_matvec, X, lambda V: flam_amd.rskelf_sv(F, V), tol=1e-10,
                      inner_matvec=g.closed_form_matvec)
    R = X - g.circulant_matvec(Xs)
    rr = (torch.linalg.vector_norm(R, dim=0) / torch.linalg.vector_norm(X, dim=0)).max().item()
    print(f"fgmres iters {info['iters']} outer {info['outer']} relres {rr:.2e}", flush=True)
    assert rr <= 1e-10
    # the factor itself as the preconditioner: in-place solves on the
    # stored directions (a strided slice, leading dimension restart * N)
    # give the same iterates as the copying path
    Xf, info_f = fgmres(g.circulant_matvec, X, F, tol=1e-10, inner_matvec=g.closed_form_matvec)
    assert info_f["iters"] == info["iters"]
    assert (Xf - Xs).abs().max().item() <= 1e-13 * Xs.abs().max().item()


def test_inplace_solve_on_strided_rows():
    """rskelf_sv_ on rows of a larger array (leading dimension > N) equals
    rskelf_sv on a contiguous copy and leaves the other rows untouched."""
    import flam_amd
    from flam_amd import geometry
    from flam_amd.rskelf import rskelf_sv_
    N = 5000
    print(f"in-place solve: ellipse N={N}", flush=True)
    g = geometry.ellipse(N, 1.0, 0.5)
    F = flam_amd.rskelf(g.kernel, g.x, 64, 1e-9, g.proxy, dict(symm="n"))
    Z = torch.from_numpy(np.random.default_rng(9).standard_normal((3, 4, N))).cuda()
    Z0 = Z.clone()
    rskelf_sv_(F, Z[:, 1, :])
    ref = flam_amd.rskelf_sv(F, Z0[:, 1, :].t())            # (N, 3)
    assert torch.equal(Z[:, 1, :], ref.t())
    assert torch.equal(Z[:, 0, :], Z0[:, 0, :]) and torch.equal(Z[:, 2:, :], Z0[:, 2:, :])


def test_factor_operator_refines_against_exact_product():
    """A tol-1e-12 FP64 factor's forward sweep as the refinement operator
    (no fast exact product for ellipses): it agrees with the exact
    on-the-fly product to ~1e-12, and FGMRES on it with an FP32
    preconditioner reaches a true residual (exact product) <= 1e-10 in
    every problem of a batch."""
    import flam_amd
    from flam_amd import geometry
    from flam_amd.refine import factor_operator, fgmres
    probs = [geometry.ellipse(n, 1.0, b) for n, b in [(3000, 0.3), (9000, 0.8), (20000, 0.5)]]
    print("factor operator: 3 ellipses N=3000+9000+20000", flush=True)
    batch = flam_amd.stack_problems(probs)
    F_op = flam_amd.rskelf_batch(batch, 128, 1e-12, dict(symm="n"))
    op = factor_operator(F_op)
    exact = lambda V: flam_amd.kernel_matvec_batch(batch, V)
    X = torch.from_numpy(np.random.default_rng(10).standard_normal((batch.N, 3))).cuda()
    Y1, Y2 = op(X), exact(X)
    rel = (torch.linalg.vector_norm(Y1 - Y2, dim=0) / torch.linalg.vector_norm(Y2, dim=0)).max()
    print(f"operator vs exact product: {rel.item():.2e}", flush=True)
    assert rel.item() <= 1e-11
    F = flam_amd.rskelf_batch(batch, 128, 1e-6, dict(symm="n", precision="fp32"))
    segs = list(zip(batch.offsets[:-1], batch.offsets[1:]))
    Xs, info = fgmres(op, X, F, tol=1e-10, segments=segs)
    R = X - exact(Xs)
    worst = max((torch.linalg.vector_norm(R[a:b], dim=0) /
                 torch.linalg.vector_norm(X[a:b], dim=0)).max().item() for a, b in segs)
    print(f"fgmres iters {info['iters']}, true residual {worst:.2e}", flush=True)
    assert worst <= 1e-10


def test_fgmres_zero_column():
    """A zero right-hand side among random ones comes back exactly zero with
    a finite residual (MATLAB's gmres returns x = 0 for b = 0) while the
    other columns converge as usual."""
    import flam_amd
    from flam_amd import geometry
    from flam_amd.refine import fgmres
    N = 8192
    print(f"fgmres with a zero column, circle N={N}", flush=True)
    g = geometry.circle(N)
    F = flam_amd.rskelf(g.kernel, g.x, 128, 1e-6, g.proxy, dict(symm="p", precision="fp32"))
    B = torch.from_numpy(np.random.default_rng(4).standard_normal((N, 3))).cuda()
    B[:, 1] = 0
    X, info = fgmres(g.circulant_matvec, B, F, tol=1e-10)
    assert bool(torch.isfinite(X).all())
    assert X[:, 1].abs().max().item() == 0.0
    assert all(math.isfinite(r) for r in info["relres"])
    R = B - g.circulant_matvec(X)
    for c in (0, 2):
        assert (torch.linalg.vector_norm(R[:, c]) / torch.linalg.vector_norm(B[:, c])).item() <= 1e-10


def test_graph_captured_solve():
    """rskelf_sv captured in a HIP graph: every replay equals the eager
    solve bit for bit, with eager work and synchronizations between the
    replays, and FGMRES preconditioned by the graph takes the same iterates
    as with the factor itself."""
    import flam_amd
    from flam_amd import geometry
    from flam_amd.graph import SolveGraph
    from flam_amd.refine import fgmres
    N = 1 << 16
    print(f"graph-captured solve: circle N={N} fp32", flush=True)
    g = geometry.circle(N)
    F = flam_amd.rskelf(g.kernel, g.x, 256, 1e-6, g.proxy, dict(symm="p", precision="fp32"))
    G = SolveGraph(F, 8)
    rng = np.random.default_rng(9)
    for rep in range(4):
        B = torch.from_numpy(rng.standard_normal((8, N))).cuda()
        eager = flam_amd.rskelf_sv(F, B.t()).t()
        flam_amd.rskelf_mv(F, B.t())                  # other eager work in between
        torch.cuda.synchronize()
        Xt = B.clone()
        G.solve_(Xt)
        assert torch.equal(Xt, eager), rep
    B = torch.from_numpy(rng.standard_normal((N, 8))).cuda()
    X1, i1 = fgmres(g.circulant_matvec, B, F, tol=1e-10, inner_matvec=g.closed_form_matvec)
    X2, i2 = fgmres(g.circulant_matvec, B, G, tol=1e-10, inner_matvec=g.closed_form_matvec)
    assert i1["iters"] == i2["iters"] and torch.equal(X1, X2)


def test_tuned_shapes_give_the_same_factor():
    """Launch-shape parameters change which kernel shape a box runs on, not
    the result: the level table is unchanged; concurrent classes off gives
    the same factor bit for bit, the 8-wave register shape narrowed sends
    boxes to the 16-wave shape (same arithmetic up to FP32 rounding), and a
    small QR shared-memory budget forces the QR ID's Householder vectors to
    global memory (FP64 QRCP route)."""
    import flam_amd
    from flam_amd import geometry, tune
    N = 1 << 17
    print(f"tuning parameters: circle N={N}", flush=True)
    g = geometry.circle(N)
    B = torch.from_numpy(np.random.default_rng(3).standard_normal((N, 2))).cuda()
    o32 = dict(symm="p", precision="fp32")
    F0 = flam_amd.rskelf(g.kernel, g.x, 256, 1e-6, g.proxy, o32)
    X0 = flam_amd.rskelf_sv(F0, B)
    try:
        tune.set_tuning("concurrent", 0)
        F1 = flam_amd.rskelf(g.kernel, g.x, 256, 1e-6, g.proxy, o32)
        assert F1.level_table() == F0.level_table()
        assert torch.equal(flam_amd.rskelf_sv(F1, B), X0)
        tune.set_tuning("concurrent", 1)
        tune.set_tuning("elim8_nmax", 128)
        F2 = flam_amd.rskelf(g.kernel, g.x, 256, 1e-6, g.proxy, o32)
        assert F2.level_table() == F0.level_table()
        X2 = flam_amd.rskelf_sv(F2, B)
        assert (X2 - X0).abs().max().item() <= 1e-5 * X0.abs().max().item()
        tune.set_tuning("elim8_nmax", 1 << 30)
        o64 = dict(symm="n", id_route="qr")
        Fq = flam_amd.rskelf(g.kernel, g.x, 128, 1e-10, g.proxy, o64)
        tune.set_tuning("id_reg", 0)               # every QR box to the global kernel
        tune.set_tuning("id_lds", 4096)            # ... with its vectors in global memory
        Fg = flam_amd.rskelf(g.kernel, g.x, 128, 1e-10, g.proxy, o64)
        assert Fg.level_table() == Fq.level_table()
        Xq, Xg = flam_amd.rskelf_sv(Fq, B), flam_amd.rskelf_sv(Fg, B)
        assert (Xg - Xq).abs().max().item() <= 1e-12 * Xq.abs().max().item()
    finally:
        for k, v in (("concurrent", 1), ("elim8_nmax", 1 << 30), ("id_lds", 160 * 1024),
                     ("id_reg", 1)):
            tune.set_tuning(k, v)
