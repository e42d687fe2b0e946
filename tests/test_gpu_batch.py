"""Many independent problems in one factorization (rskx_rskelf_batch).

The batch factors the block-diagonal matrix diag(A_1, ..., A_P) with every
problem keeping its own tree; the boxes of one tree level of all problems
go through the same launches.  A box's computation depends only on its own
tree, so each problem's block of the batch factor must be the factor that
rskx_rskelf computes for that problem alone -- already checked against the
CPU oracle in test_gpu_parity.py.  Here: the batch's solves and products
equal the per-problem ones to rounding (1e-13 relative; identical kernels
on identical data), its level table is the sum of the per-problem tables,
and the FP32 batch factor preconditions FGMRES to 1e-10 per problem.

Problems: ellipses x = (cos t, b sin t) (rskelf/test/ie_ellipse.m:28-37,
103-132, double-layer kernel, symm 'n'), b ~ U[0.3, 1] and N log-uniform,
seeded, plus a 1-point and a 37-point problem as edge cases.
"""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _problems(nprob, nmin, nmax, seed, edge=True):
    from flam_amd import geometry
    rng = np.random.default_rng(seed)
    bs = rng.uniform(0.3, 1.0, nprob)
    Ns = np.round(np.exp(rng.uniform(math.log(nmin), math.log(nmax), nprob))).astype(int)
    probs = [geometry.ellipse(int(n), 1.0, float(b)) for n, b in zip(Ns, bs)]
    if edge:
        probs.insert(1, geometry.ellipse(1, 1.0, 0.5))
        probs.insert(3, geometry.ellipse(37, 1.0, 0.7))
    return probs


def _expected_table(tables, Ns, nlvl):
    """Level table of the batch from the per-problem tables: trees aligned
    at their roots, so reference level lvl of the batch is level lvl of
    every tree; a tree with fewer levels contributes its leaves and its
    untouched points to the deeper rows."""
    rows = []
    for lvl in range(nlvl, 0, -1):
        nblk = start = end = 0
        for tab, N in zip(tables, Ns):
            by = {r[0]: r for r in tab}
            if lvl in by:
                _, b, s, e = by[lvl]
            else:                         # deeper than this tree: nothing factored yet
                b, s, e = by[max(by)][1], N, N
            nblk += b
            start += s
            end += e
        rows.append((lvl, nblk, start, end))
    return rows


@pytest.mark.parametrize("symm", ["n", "p"])
def test_batch_equals_standalone_factors(symm):
    import flam_amd
    if symm == "p":
        from flam_amd import geometry
        # SPD problems: benchmark-sign circles of uneven N, whose matrix
        # (1/2 - 1/N) I + 1/(2N) 11' is positive definite for N >= 3
        probs = [geometry.circle(n) for n in (700, 3, 3000, 37, 1500)]
    else:
        probs = _problems(5, 300, 6000, seed=1)
    occ, tol = 64, 1e-10
    Ns = [p.N for p in probs]
    print(f"batch symm={symm} N={Ns} occ={occ} tol={tol}", flush=True)
    batch = flam_amd.stack_problems(probs)
    Fb = flam_amd.rskelf_batch(batch, occ, tol, dict(symm=symm))
    Fs = [flam_amd.rskelf(p.kernel, p.x, occ, tol, p.proxy, dict(symm=symm)) for p in probs]
    tables = [F.level_table() for F in Fs]
    nlvl = max(len(t) for t in tables)
    assert Fb.level_table() == _expected_table(tables, Ns, nlvl)
    assert Fb.info()["nfactor"] == sum(F.info()["nfactor"] for F in Fs)
    rng = np.random.default_rng(0)
    B = torch.from_numpy(rng.standard_normal((batch.N, 3))).cuda()
    for fn, trans in ((flam_amd.rskelf_sv, "n"), (flam_amd.rskelf_sv, "c"),
                      (flam_amd.rskelf_mv, "n"), (flam_amd.rskelf_mv, "t")):
        Yb = batch.split(fn(Fb, B, trans))
        for F, Bj, Yj in zip(Fs, batch.split(B), Yb):
            Ys = fn(F, Bj, trans)
            err = (Yj - Ys).abs().max().item()
            assert err <= 1e-13 * max(Ys.abs().max().item(), 1.0), (fn.__name__, trans, err)


def test_batch_matvec_is_block_diagonal():
    import flam_amd
    probs = _problems(6, 200, 3000, seed=2)
    batch = flam_amd.stack_problems(probs)
    print(f"batch matvec N={[p.N for p in probs]}", flush=True)
    X = torch.from_numpy(np.random.default_rng(1).standard_normal((batch.N, 2))).cuda()
    Yb = batch.split(flam_amd.kernel_matvec_batch(batch, X))
    for p, Xj, Yj in zip(probs, batch.split(X), Yb):
        Ys = flam_amd.kernel_matvec(p.kernel, p.x, Xj)
        assert torch.allclose(Yj, Ys, rtol=1e-14, atol=1e-14 * Ys.abs().max().item())


def test_batch_single_problem_equals_rskelf():
    import flam_amd
    from flam_amd import geometry
    g = geometry.ellipse(5000, 1.0, 0.4)
    print("batch of one: ellipse N=5000 b=0.4", flush=True)
    Fb = flam_amd.rskelf_batch(flam_amd.stack_problems([g]), 64, 1e-10, dict(symm="n"))
    F = flam_amd.rskelf(g.kernel, g.x, 64, 1e-10, g.proxy, dict(symm="n"))
    assert Fb.level_table() == F.level_table()
    B = torch.from_numpy(np.random.default_rng(3).standard_normal(g.N)).cuda()
    assert torch.equal(flam_amd.rskelf_sv(Fb, B), flam_amd.rskelf_sv(F, B))


def test_batch_fp32_refined_per_problem():
    """Mixed precision on the batch: FP32 factor (tol 1e-6), FP64 block-
    diagonal matvec, FGMRES until every problem's relative residual of
    every column is <= 1e-10."""
    import flam_amd
    from flam_amd.refine import fgmres
    probs = _problems(10, 500, 20000, seed=3)
    batch = flam_amd.stack_problems(probs)
    print(f"batch fp32 N={[p.N for p in probs]}", flush=True)
    F = flam_amd.rskelf_batch(batch, 128, 1e-6, dict(symm="n", precision="fp32"))
    B = torch.from_numpy(np.random.default_rng(4).standard_normal((batch.N, 4))).cuda()
    segs = list(zip(batch.offsets[:-1], batch.offsets[1:]))
    X, info = fgmres(lambda V: flam_amd.kernel_matvec_batch(batch, V), B,
                     lambda V: flam_amd.rskelf_sv(F, V), tol=1e-10, segments=segs)
    R = B - flam_amd.kernel_matvec_batch(batch, X)
    for (a, b) in segs:
        rel = (torch.linalg.vector_norm(R[a:b], dim=0) / torch.linalg.vector_norm(B[a:b], dim=0)).max()
        assert rel.item() <= 1e-10, (a, b, rel.item(), info)
    assert info["iters"] <= 12


def test_batch_rejects_mixed_proxy_rings():
    import flam_amd
    from flam_amd import geometry
    from flam_amd.rskelf import FlamError
    a, b = geometry.circle(300), geometry.ellipse(300, 1.0, 0.5)
    b.scale = a.scale
    with pytest.raises(FlamError) as e:
        flam_amd.rskelf_batch(flam_amd.stack_problems([a, b]), 64, 1e-8, dict(symm="n"))
    assert e.value.ident == "FLAM:rskx:batch"
