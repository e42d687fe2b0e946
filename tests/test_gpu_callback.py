"""Matrices given by host callbacks (RSKX_KERNEL_CALLBACK).

The reference takes the matrix as a function handle A(I,J) and the proxy
as pxyfun(x,slf,nbr,l,ctr) (rskelf.m:205,210-211,219).  Here the same
Python callables that drive the oracle's rskelf (oracle/problems.py) drive
the library's callback path: the library calls them on the host for each
box, uploads K and A(slf,slf), and compresses and eliminates on the device.
So the device factor and the oracle's see bit-identical matrix entries.

Checks: FP64 per-level rank tables identical to the oracle's (level
order) on the circle ('p') and the ellipse ('n'); solves against the
closed form / dense LU to 1e-10; the callback factor of the ellipse against
the built-in double-layer kernel's factor of the same matrix; the 3D
sphere (d = 3, two-sided proxy rows) against dense LU; an FP32 callback
factor's one-shot error; the proxy buffer growing on request; and an
exception raised inside A(I,J) surfacing from rskelf.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import flam as ofl  # noqa: E402
from oracle import problems as opr  # noqa: E402


def _dev():
    import flam_amd
    from flam_amd import geometry
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return flam_amd, geometry


def _np(t):
    return t.detach().cpu().numpy()


def _sv(flam_amd, F, B, trans="n"):
    return _np(flam_amd.rskelf_sv(F, torch.from_numpy(np.ascontiguousarray(B)).cuda(), trans))


@pytest.mark.parametrize("kind,symm", [("circle", "p"), ("ellipse", "n"), ("circle", "h")])
def test_callback_fp64_matches_oracle(kind, symm):
    flam_amd, _ = _dev()
    N, occ, tol = 4096, 64, 1e-10
    print(f"callback fp64: {kind} N={N} occ={occ} tol={tol} symm {symm}", flush=True)
    pr = opr.CurveDL(N, kind, a=2.0 if kind == "ellipse" else 1.0)
    F = flam_amd.rskelf(pr.afun, pr.x, occ, tol, pr.pxyfun, dict(symm=symm))
    Fo = ofl.rskelf(pr.afun, pr.x, occ, tol, pr.pxyfun, dict(symm=symm, order="level"))
    got, ref = F.level_table(), ofl.rank_table(Fo)
    print("device", got, "\noracle", ref, flush=True)
    assert got == ref
    B = np.random.default_rng(0).standard_normal((N, 3))
    X = _sv(flam_amd, F, B)
    if kind == "circle":
        Xe = pr.closed_form_inverse_apply(B)
    else:
        Xe = np.linalg.solve(pr.dense(), B)
    err = np.linalg.norm(X - Xe) / np.linalg.norm(Xe)
    print(f"solve error {err:.3e}", flush=True)
    assert err < 1e-9
    Y = _np(flam_amd.rskelf_mv(F, torch.from_numpy(X).cuda()))
    assert np.linalg.norm(Y - B) < 1e-9 * np.linalg.norm(B)


def test_callback_equals_builtin_kernel():
    """The ellipse through callbacks and through the built-in double-layer
    kernel (the same matrix, entries evaluated by NumPy on the host or by
    the device): rank tables within 2 per level, solutions agree to the
    tolerance, for trans 'n' and 'c'."""
    flam_amd, geometry = _dev()
    N, occ, tol = 8192, 64, 1e-9
    print(f"callback vs built-in: ellipse N={N} symm n", flush=True)
    g = geometry.ellipse(N, 2.0, 1.0)
    pr = opr.CurveDL(N, "ellipse", a=2.0)
    assert np.abs(_np(g.x) - pr.x).max() < 1e-14
    Fc = flam_amd.rskelf(pr.afun, pr.x, occ, tol, pr.pxyfun, dict(symm="n"))
    Fb = flam_amd.rskelf(g.kernel, g.x, occ, tol, g.proxy, dict(symm="n"))
    for a, b in zip(Fc.level_table(), Fb.level_table()):
        assert a[:3] == b[:3] and abs(a[3] - b[3]) <= 2, (a, b)
    B = np.random.default_rng(1).standard_normal((N, 2))
    for trans in ("n", "c"):
        Xc, Xb = _sv(flam_amd, Fc, B, trans), _sv(flam_amd, Fb, B, trans)
        assert np.linalg.norm(Xc - Xb) < 100 * tol * np.linalg.norm(Xb), trans


def test_callback_sphere_two_sided():
    """d = 3: the sphere's single layer with incoming proxy rows (symm
    'n', ie_sphere.m:174-187 through the oracle's pxyfun), solve against
    dense LU as accurate as the oracle's own factor."""
    flam_amd, _ = _dev()
    print("callback sphere: 1280 triangles occ 64 tol 1e-6 symm n", flush=True)
    pr = opr.SphereSL(1280, symm="n")
    tol = 1e-6
    F = flam_amd.rskelf(pr.afun, pr.x, 64, tol, pr.pxyfun, dict(symm="n"))
    Fo = ofl.rskelf(pr.afun, pr.x, 64, tol, pr.pxyfun, dict(symm="n", order="level"))
    got, ref = F.level_table(), ofl.rank_table(Fo)
    print("device", got, "\noracle", ref, flush=True)
    for a, b in zip(got, ref):
        assert a[:3] == b[:3] and abs(a[3] - b[3]) <= max(2, 0.02 * b[3]), (a, b)
    A = pr.dense()
    B = np.random.default_rng(2).standard_normal((pr.x.shape[1], 2))
    Xe = np.linalg.solve(A, B)
    err = np.linalg.norm(_sv(flam_amd, F, B) - Xe) / np.linalg.norm(Xe)
    err_o = np.linalg.norm(ofl.rskelf_sv(Fo, B) - Xe) / np.linalg.norm(Xe)
    print(f"solve error {err:.3e} (oracle {err_o:.3e})", flush=True)
    assert err < 10 * max(err_o, tol)


def test_callback_fp32_and_proxy_buffer_growth():
    """FP32 storage from FP64 callbacks: one-shot solve error against the
    closed form below 1e-4; a proxy buffer of one row (grown to 64 on each
    box's request) gives the same factor as the default buffer."""
    flam_amd, _ = _dev()
    N = 1 << 14
    print(f"callback fp32: circle N={N} occ 128 tol 1e-6 symm p", flush=True)
    pr = opr.CurveDL(N, "circle")
    F = flam_amd.rskelf(pr.afun, pr.x, 128, 1e-6, pr.pxyfun, dict(symm="p", precision="fp32"))
    F1 = flam_amd.rskelf(pr.afun, pr.x, 128, 1e-6, pr.pxyfun,
                         dict(symm="p", precision="fp32", proxy_rows=1))
    assert F.level_table() == F1.level_table()
    B = np.random.default_rng(3).standard_normal((N, 2))
    X, X1 = _sv(flam_amd, F, B), _sv(flam_amd, F1, B)
    assert np.array_equal(X, X1)
    Xe = pr.closed_form_inverse_apply(B)
    err = np.linalg.norm(X - Xe) / np.linalg.norm(Xe)
    print(f"fp32 one-shot error {err:.3e}", flush=True)
    assert err < 1e-4


def test_callback_exception_surfaces():
    flam_amd, _ = _dev()
    pr = opr.CurveDL(2048, "circle")
    calls = []

    def afun(I, J):
        calls.append(len(I))
        if len(calls) == 5:
            raise ArithmeticError("entries failed")
        return pr.afun(I, J)

    with pytest.raises(ArithmeticError, match="entries failed"):
        flam_amd.rskelf(afun, pr.x, 64, 1e-8, pr.pxyfun, dict(symm="p"))
    # the library is usable afterwards
    F = flam_amd.rskelf(pr.afun, pr.x, 64, 1e-8, pr.pxyfun, dict(symm="p"))
    B = np.random.default_rng(4).standard_normal((2048, 1))
    Xe = pr.closed_form_inverse_apply(B)
    assert np.linalg.norm(_sv(flam_amd, F, B) - Xe) < 1e-7 * np.linalg.norm(Xe)
