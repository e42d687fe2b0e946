"""3D Laplace single layer on the triangulated unit sphere (the octree path).

Points are triangle centroids of a subdivided icosahedron, A_ij = area_j /
(4 pi |x_i - x_j|), the self-panel integral on the diagonal, 512 proxy
points on a sphere of one box size (rskelf/test/ie_sphere.m:36-40,
174-187).  The oracle runs on exactly the arrays the device factors.

Tolerances: the ID's rank decisions may break ties differently, so the
per-level skeleton totals must agree within 2%; solves must agree with the
dense FP64 LU to the accuracy the factor tolerance allows: the oracle's own
error times 10, and at most 1000 tol (the first-kind matrix's condition
number, which grows with N, amplifies the factor's tolerance; the oracle's
own error is the binding bound).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import flam as ofl  # noqa: E402
from oracle import problems as opr  # noqa: E402


def _oracle_problem(g, symm="n"):
    return opr.SphereSL.from_arrays(g.x.cpu().numpy(), g.w.cpu().numpy(), g.diag.cpu().numpy(),
                                    g.proxy.pts.cpu().numpy(), symm=symm)


def test_sphere_matvec_matches_dense():
    import flam_amd
    from flam_amd import geometry
    g = geometry.sphere(1280)
    print("sphere matvec nface=1280", flush=True)
    pr = _oracle_problem(g)
    X = np.random.default_rng(0).standard_normal((g.N, 3))
    Y = flam_amd.kernel_matvec(g.kernel, g.x, torch.from_numpy(X).cuda()).cpu().numpy()
    Ye = pr.dense() @ X
    assert np.abs(Y - Ye).max() <= 1e-13 * np.abs(Ye).max()


@pytest.mark.parametrize("nface,occ,tol", [(1280, 32, 1e-4), (5120, 128, 1e-6)])
def test_sphere_fp64_vs_oracle_and_dense(nface, occ, tol):
    import flam_amd
    from flam_amd import geometry
    g = geometry.sphere(nface)
    print(f"sphere fp64 nface={nface} occ={occ} tol={tol} symm=n", flush=True)
    F = flam_amd.rskelf(g.kernel, g.x, occ, tol, g.proxy, dict(symm="n"))
    pr = _oracle_problem(g)
    Fo = ofl.rskelf(pr.afun, pr.x, occ, tol, pr.pxyfun, dict(symm="n", order="level"))
    got, ref = F.level_table(), ofl.rank_table(Fo)
    assert [r[:3] for r in got][:1] == [r[:3] for r in ref][:1]
    assert len(got) == len(ref)
    for a, b in zip(got, ref):
        assert (a[0], a[1]) == (b[0], b[1]), (a, b)
        assert abs(a[3] - b[3]) <= max(2, 0.02 * b[3]), (a, b)
    A = pr.dense()
    b = np.random.default_rng(1).standard_normal(g.N)
    xe = np.linalg.solve(A, b)
    x = flam_amd.rskelf_sv(F, torch.from_numpy(b).cuda()).cpu().numpy()
    xo = ofl.rskelf_sv(Fo, b)
    err = np.linalg.norm(x - xe) / np.linalg.norm(xe)
    err_o = np.linalg.norm(xo - xe) / np.linalg.norm(xe)
    print(f"solve error: device {err:.2e}, oracle {err_o:.2e}", flush=True)
    assert err <= max(10 * err_o, 1e-12) and err <= 1000 * tol
    y = flam_amd.rskelf_mv(F, torch.from_numpy(xe).cuda()).cpu().numpy()
    assert np.linalg.norm(y - b) / np.linalg.norm(b) <= max(10 * err_o, 1e-12)
    # adjoint sweeps against the transposed dense solve
    xt = flam_amd.rskelf_sv(F, torch.from_numpy(b).cuda(), "t").cpu().numpy()
    xte = np.linalg.solve(A.T, b)
    assert np.linalg.norm(xt - xte) / np.linalg.norm(xte) <= max(100 * err_o, 1e-10)


def test_sphere_fp32_refined_to_1e9():
    """Mixed precision on the sphere: FP32 factor at tol 1e-5 (occ 256),
    FP64 on-the-fly matvec, flexible GMRES to relative residual 1e-9."""
    import flam_amd
    from flam_amd import geometry
    from flam_amd.refine import fgmres
    g = geometry.sphere(20480)
    print("sphere fp32 nface=20480 occ=256 tol=1e-5 + fgmres 1e-9", flush=True)
    F = flam_amd.rskelf(g.kernel, g.x, 256, 1e-5, g.proxy, dict(symm="n", precision="fp32"))
    print("levels:", F.level_table(), flush=True)
    B = torch.from_numpy(np.random.default_rng(2).standard_normal((g.N, 4))).cuda()
    mv = lambda V: flam_amd.kernel_matvec(g.kernel, g.x, V)
    X, info = fgmres(mv, B, lambda V: flam_amd.rskelf_sv(F, V), tol=1e-9)
    R = B - mv(X)
    rel = (torch.linalg.vector_norm(R, dim=0) / torch.linalg.vector_norm(B, dim=0)).max().item()
    print(f"fgmres iters {info['iters']} relres {rel:.2e}", flush=True)
    assert rel <= 1e-9 and info["iters"] <= 16


def test_sphere_batch_equals_standalone():
    """Two spheres of different sizes in one batched factorization."""
    import flam_amd
    from flam_amd import geometry
    probs = [geometry.sphere(1280), geometry.sphere(5120)]
    print("sphere batch nface=1280+5120", flush=True)
    # one proxy set for the batch: the problems' incoming-row weights differ
    probs[1].proxy.pts = probs[0].proxy.pts
    batch = flam_amd.stack_problems(probs)
    Fb = flam_amd.rskelf_batch(batch, 128, 1e-6, dict(symm="n"))
    B = torch.from_numpy(np.random.default_rng(3).standard_normal((batch.N, 2))).cuda()
    Yb = batch.split(flam_amd.rskelf_sv(Fb, B))
    for p, Bj, Yj in zip(probs, batch.split(B), Yb):
        F = flam_amd.rskelf(p.kernel, p.x, 128, 1e-6, p.proxy, dict(symm="n"))
        Ys = flam_amd.rskelf_sv(F, Bj)
        assert (Yj - Ys).abs().max().item() <= 1e-13 * Ys.abs().max().item()
