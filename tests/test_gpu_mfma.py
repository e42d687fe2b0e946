"""Matrix-core kernels against the vector kernels they replace.

The 'p' elimination on 16 x 16 MFMA tiles (every tile shape: boxes up to
n = 256 at occ 256 reach the largest) and the Gram ID with G accumulated on
FP64 MFMA change only the order of the arithmetic: the rank tables must be
the same as with the vector kernels, and solves must agree to the factor
precision's rounding.  Both are also pinned against the oracle through the
other GPU tests, which run the default kernels.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _solve_pair(flam_amd, tune, knob, g, B, occ, tol, opts):
    tune.set_tuning(knob, 1)
    Fa = flam_amd.rskelf(g.kernel, g.x, occ, tol, g.proxy, opts)
    tune.set_tuning(knob, 0)
    Fb = flam_amd.rskelf(g.kernel, g.x, occ, tol, g.proxy, opts)
    return Fa, Fb, flam_amd.rskelf_sv(Fa, B), flam_amd.rskelf_sv(Fb, B)


@pytest.mark.parametrize("prec,rtol", [("fp32", 2e-5), ("fp64", 1e-12)])
def test_mfma_elimination_matches_register_elimination(prec, rtol):
    import flam_amd
    from flam_amd import geometry, tune
    N = 1 << 17
    print(f"mfma elimination vs register elimination: circle N={N} {prec}", flush=True)
    g = geometry.circle(N)
    B = torch.from_numpy(np.random.default_rng(5).standard_normal((N, 3))).cuda()
    try:
        Fa, Fb, Xa, Xb = _solve_pair(flam_amd, tune, "elim_mfma", g, B, 256, 1e-6,
                                     dict(symm="p", precision=prec, id_route="gram"))
    finally:
        tune.set_tuning("elim_mfma", 1)
    assert Fa.level_table() == Fb.level_table()
    err = ((Xa - Xb).abs().max() / Xb.abs().max()).item()
    print(f"max relative difference of the solves {err:.2e}", flush=True)
    assert err <= rtol
    # the exported factors agree too (L, E of one leaf factor)
    va, vb = Fa.export(0), Fb.export(0)
    for key in ("L", "E", "T"):
        a, b = np.asarray(va[key]), np.asarray(vb[key])
        assert a.shape == b.shape
        assert np.abs(a - b).max() <= 50 * rtol * max(np.abs(b).max(), 1.0), key


def test_mfma_gram_id_matches_vector_gram_id():
    import flam_amd
    from flam_amd import geometry, tune
    N = 1 << 17
    print(f"mfma Gram ID vs vector Gram ID: circle N={N} fp32", flush=True)
    g = geometry.circle(N)
    B = torch.from_numpy(np.random.default_rng(6).standard_normal((N, 2))).cuda()
    try:
        Fa, Fb, Xa, Xb = _solve_pair(flam_amd, tune, "gram_mfma", g, B, 256, 1e-6,
                                     dict(symm="p", precision="fp32", id_route="gram"))
    finally:
        tune.set_tuning("gram_mfma", 0)
    assert Fa.level_table() == Fb.level_table()
    err = ((Xa - Xb).abs().max() / Xb.abs().max()).item()
    print(f"max relative difference of the solves {err:.2e}", flush=True)
    assert err <= 1e-4
