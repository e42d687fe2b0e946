"""The learned rank model on the GPU: its backward kernel against the CPU
restatement (oracle/mlp.py, itself pinned by finite differences), training
on ranks harvested from the library's own factorizations, and the
factorization with the model: buffers sized by the predictions, the same
results as without it, and the fallback to full buffers when predictions
are too low (forced with rank_scale < 1)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import mlp  # noqa: E402


def _dev():
    import flam_amd
    from flam_amd import geometry, rank
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return flam_amd, geometry, rank


def test_device_gradient_matches_cpu_restatement():
    flam_amd, geometry, rank = _dev()
    print("rank model: device forward / gradient vs the CPU restatement", flush=True)
    rng = np.random.default_rng(1)
    M = rank.RankModel(seed=3)
    p = 0.4 * rng.standard_normal(M.nparams())
    M.set_params(p)
    assert np.array_equal(M.params(), p)
    for n in (1, 31, 32, 1000):                  # partial and full gradient blocks
        X = rng.standard_normal((n, rank.NFEAT))
        t = rng.standard_normal(n)
        loss_d, g_d = M.grad(X, t)
        loss_c, g_c = mlp.loss_grad(p, X, t)
        assert abs(loss_d - loss_c) <= 1e-12 * loss_c
        assert np.abs(g_d - g_c).max() <= 1e-12 * np.abs(g_c).max()
        y_c = mlp.forward(p, X)[0]
        assert np.abs(M.predict(X) - y_c).max() <= 1e-13 * max(1.0, np.abs(y_c).max())


def _harvest_set(flam_amd, geometry):
    """Exact ranks of FP32 factorizations at the Gram route's tolerances."""
    Fs = []
    for N in (1 << 15, 1 << 16, 1 << 17):
        g = geometry.circle(N)
        for occ, tol in ((256, 1e-6), (128, 1e-7)):
            Fs.append(flam_amd.rskelf(g.kernel, g.x, occ, tol, g.proxy,
                                      dict(symm="p", precision="fp32", record_boxes=True)))
    for b in (0.3, 0.6):
        g = geometry.ellipse(1 << 16, 1.0, b)
        Fs.append(flam_amd.rskelf(g.kernel, g.x, 256, 1e-6, g.proxy,
                                  dict(symm="n", precision="fp32", record_boxes=True)))
    return Fs


def _trained(flam_amd, geometry, rank):
    X, t = rank.harvest(_harvest_set(flam_amd, geometry))
    M = rank.RankModel(seed=0)
    loss0 = M.grad(X, t)[0]
    loss = M.train(X, t, epochs=300, batch=256, lr=3e-3)
    return M, X, t, loss0, loss


def test_training_on_harvested_ranks():
    """Adam with the device's own kernels fits the harvested ranks: the mean
    squared error of log2(1 + k) / 10 falls by two orders of magnitude and
    most boxes' predicted ranks are within 25% + 4 of the exact ones."""
    flam_amd, geometry, rank = _dev()
    print("rank model: harvest and train", flush=True)
    M, X, t, loss0, loss = _trained(flam_amd, geometry, rank)
    k = rank.rank_of(t)
    kh = rank.rank_of(M.predict(X))
    inside = np.mean(k <= 1.25 * kh + 4)
    print(f"{len(t)} boxes, loss {loss0:.3e} -> {loss:.3e}, "
          f"within cap {inside:.4f}, median |k^ - k| {np.median(np.abs(kh - k)):.2f}", flush=True)
    assert loss < 1e-2 * loss0 and inside > 0.97


def test_factorization_with_the_model_and_its_fallback():
    """The benchmark's route (FP32, Gram ID, occ 256, tol 1e-6) at a size
    the model never saw: with the model the ID's work buffers shrink, the
    level table and the solve are bit for bit those without it, and forcing
    predictions low (rank_scale 0.25) redoes the under-sized boxes with full
    buffers and still gives the same results."""
    flam_amd, geometry, rank = _dev()
    M = _trained(flam_amd, geometry, rank)[0]
    N = 1 << 18
    print(f"rank model ablation: circle N={N} occ 256 tol 1e-6 fp32", flush=True)
    g = geometry.circle(N)
    base = dict(symm="p", precision="fp32")
    F0 = flam_amd.rskelf(g.kernel, g.x, 256, 1e-6, g.proxy, base)
    F1 = flam_amd.rskelf(g.kernel, g.x, 256, 1e-6, g.proxy, dict(base, rank_model=M))
    F2 = flam_amd.rskelf(g.kernel, g.x, 256, 1e-6, g.proxy,
                         dict(base, rank_model=M, rank_scale=0.25))
    i0, i1, i2 = F0.info(), F1.info(), F2.info()
    print(f"id scratch: {i0['id_scratch']/2**20:.1f} MiB without, {i1['id_scratch']/2**20:.1f} MiB "
          f"with the model ({i1['rank_fallbacks']} fallbacks); forced low: "
          f"{i2['rank_fallbacks']} fallbacks", flush=True)
    assert F1.level_table() == F0.level_table() == F2.level_table()
    assert i1["id_scratch"] * 3 < i0["id_scratch"]
    assert i1["rank_fallbacks"] <= 0.01 * i1["nfactor"]
    assert i2["rank_fallbacks"] > 0
    B = torch.from_numpy(np.random.default_rng(0).standard_normal((N, 2))).cuda()
    X0 = flam_amd.rskelf_sv(F0, B)
    assert torch.equal(flam_amd.rskelf_sv(F1, B), X0)
    assert torch.equal(flam_amd.rskelf_sv(F2, B), X0)
