"""The C-ABI library: it loads, exports every entry point of include/rskx.h,
and rejects bad arguments with the reference's error identifiers before it
touches a GPU.  CPU only (no compute calls)."""
import ctypes
import math
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _lib():
    from flam_amd import _lib as L
    if not os.path.exists(L.LIB_PATH):
        from flam_amd import build
        build.build()
    return L


def _header_functions():
    src = open(os.path.join(ROOT, "include", "rskx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(rskx_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_expected_surface():
    names = _header_functions()
    for want in ("rskx_rskelf", "rskx_rskelf_sv", "rskx_rskelf_mv", "rskx_free",
                 "rskx_last_error", "rskx_export_factor", "rskx_level_table"):
        assert want in names


def test_library_exports_every_header_symbol():
    L = _lib()
    lib = ctypes.CDLL(L.LIB_PATH)
    for name in _header_functions():
        assert hasattr(lib, name), name
    assert set(_header_functions()) == set(L.SIGNATURES)


def test_opts_defaults_match_reference():
    L = _lib()
    o = L.Opts()
    assert L.lib().rskx_opts_default(ctypes.byref(o)) == 0
    # rskelf.m:112-120
    assert o.lvlmax == 0 and not o.ext and o.Tmax == 2.0 and math.isinf(o.rrqr_iter)
    assert o.symm == b"n" and o.verb == 0 and o.precision == L.RSKX_PREC_FP64
    assert o.stop < 0                     # opts.stop = Inf


def _call(L, *, symm=b"n", occ=64, tol=1e-6, Tmax=2.0, rrqr=math.inf, lvlmax=0, d=2,
          kind=None, pkind=None):
    o = L.Opts()
    L.lib().rskx_opts_default(ctypes.byref(o))
    o.symm, o.Tmax, o.rrqr_iter, o.lvlmax = symm, Tmax, rrqr, lvlmax
    dummy = 16  # never dereferenced: validation fails first
    k = L.Kernel(kind or L.RSKX_KERNEL_LAPLACE2D_DL, dummy, dummy, dummy, 1.0)
    p = L.Proxy(pkind if pkind is not None else L.RSKX_PROXY_RING_SL, 64, 1.5, 1.0)
    h = ctypes.c_void_p()
    rc = L.lib().rskx_rskelf(dummy, d, 100, occ, tol, ctypes.byref(k), ctypes.byref(p),
                             ctypes.byref(o), ctypes.byref(h))
    return rc, L.last_error()


@pytest.mark.parametrize("kw,ident", [
    (dict(symm=b"q"), "FLAM:chksymm:invalidSymm"),
    (dict(occ=-1), "FLAM:hypoct:invalidOcc"),
    (dict(lvlmax=-3), "FLAM:hypoct:invalidLvlmax"),
    (dict(tol=-1.0), "FLAM:id:invalidRankOrTol"),
    (dict(Tmax=0.5), "FLAM:id:invalidTmax"),
    (dict(rrqr=-1.0), "FLAM:id:invalidRRQRIter"),
])
def test_argument_errors_carry_reference_ids(kw, ident):
    L = _lib()
    rc, msg = _call(L, **kw)
    assert rc == L.RSKX_ERR_ARG
    assert msg.startswith(ident)


@pytest.mark.parametrize("kw", [dict(d=3), dict(pkind=0), dict(tol=5.0)])
def test_unbuilt_options_are_refused(kw):
    L = _lib()
    rc, msg = _call(L, **kw)
    assert rc == L.RSKX_ERR_UNSUP
    assert msg.startswith("FLAM:rskx:")


def test_bad_trans_is_refused():
    L = _lib()
    rc = L.lib().rskx_rskelf_sv(None, None, 1, 1, b"x", L.RSKX_F64)
    assert rc == L.RSKX_ERR_ARG       # NULL factor checked first
    assert L.last_error().startswith("FLAM:rskx:null")


def test_python_mirror_rejects_without_gpu():
    """The Python layer raises FlamError (not a silent CPU path)."""
    import flam_amd
    from flam_amd.rskelf import FlamError
    assert issubclass(FlamError, ValueError)
    assert hasattr(flam_amd, "rskelf") and hasattr(flam_amd, "rskelf_sv")


def test_inplace_solve_checks_its_view():
    """rskelf_sv_ refuses a view it cannot hand to the C ABI as a
    column-major block (host memory, wrong length, non-unit stride)."""
    import torch
    from flam_amd.rskelf import Factor, FlamError, rskelf_sv_
    F = Factor(None, 10, None)
    for bad in (torch.zeros(2, 10), torch.zeros(2, 9), torch.zeros(10, 2).t()):
        with pytest.raises(FlamError):
            rskelf_sv_(F, bad)


def _batch_call(L, offsets, proxies, ext=False):
    o = L.Opts()
    L.lib().rskx_opts_default(ctypes.byref(o))
    buf = (ctypes.c_double * 4)(0, 0, 1, 1)
    if ext:
        o.ext = buf
    dummy = 16  # never dereferenced: validation fails first
    k = L.Kernel(L.RSKX_KERNEL_LAPLACE2D_DL, dummy, dummy, dummy, 1.0)
    offs = (ctypes.c_int64 * len(offsets))(*offsets)
    pxs = (L.Proxy * len(proxies))(*proxies)
    h = ctypes.c_void_p()
    rc = L.lib().rskx_rskelf_batch(len(offsets) - 1, offs, dummy, 2, 64, 1e-6, ctypes.byref(k),
                                   pxs, ctypes.byref(o), ctypes.byref(h))
    return rc, L.last_error()


def test_batch_arguments_checked_before_device_use():
    L = _lib()
    ring = L.Proxy(L.RSKX_PROXY_RING_DL, 64, 1.5, 0.0)
    rc, msg = _batch_call(L, [1, 10, 20], [ring, ring])
    assert rc == L.RSKX_ERR_ARG and msg.startswith("FLAM:rskx:batch")
    rc, msg = _batch_call(L, [0, 10, 10], [ring, ring])           # empty problem
    assert rc == L.RSKX_ERR_ARG and msg.startswith("FLAM:rskx:batch")
    rc, msg = _batch_call(L, [0, 10, 20], [ring, L.Proxy(L.RSKX_PROXY_RING_SL, 64, 1.5, 1.0)])
    assert rc == L.RSKX_ERR_UNSUP and msg.startswith("FLAM:rskx:batch")
    rc, msg = _batch_call(L, [0, 10, 20], [ring, ring], ext=True)
    assert rc == L.RSKX_ERR_UNSUP and msg.startswith("FLAM:rskx:batch")


def _cfie_call(L, *, symm=b"n", pkind=None, k=20.0, stop=-1):
    o = L.Opts()
    L.lib().rskx_opts_default(ctypes.byref(o))
    o.symm = symm
    o.stop = stop
    dummy = 16  # never dereferenced: validation fails first
    kern = L.Kernel(L.RSKX_KERNEL_HELMHOLTZ2D_CFIE, dummy, dummy, dummy, 1.0, k, k)
    p = L.Proxy(pkind if pkind is not None else L.RSKX_PROXY_RING_CFIE, 64, 1.5, 1.0)
    h = ctypes.c_void_p()
    rc = L.lib().rskx_rskelf(dummy, 2, 100, 64, 1e-6, ctypes.byref(kern), ctypes.byref(p),
                             ctypes.byref(o), ctypes.byref(h))
    return rc, L.last_error()


@pytest.mark.parametrize("kw,code,ident", [
    (dict(symm=b"p"), "UNSUP", "FLAM:rskx:symm"),
    (dict(symm=b"h"), "UNSUP", "FLAM:rskx:symm"),
    (dict(stop=2), "UNSUP", "FLAM:rskx:stop"),
    (dict(pkind=1), "UNSUP", "FLAM:rskx:proxy"),
    (dict(k=0.0), "ARG", "FLAM:rskx:kernel"),
])
def test_complex_kernel_arguments(kw, code, ident):
    """The combined field takes symm 'n' / 's', no partial factorization,
    its own proxy ring and a positive wavenumber; checked before any
    device work."""
    L = _lib()
    rc, msg = _cfie_call(L, **kw)
    assert rc == getattr(L, "RSKX_ERR_" + code) and msg.startswith(ident), msg


def test_complex_rhs_needs_complex_factor():
    import torch
    from flam_amd.rskelf import Factor, FlamError, _rhs_dtype
    F = Factor(None, 10, None, cplx=False)
    with pytest.raises(FlamError):
        _rhs_dtype(F, torch.zeros(10, dtype=torch.complex128))
    Fz = Factor(None, 10, None, cplx=True)
    assert _rhs_dtype(Fz, torch.zeros(10)) == torch.complex128
    assert _rhs_dtype(Fz, torch.zeros(10, dtype=torch.complex64)) == torch.complex64


def test_tuning_parameters_checked():
    L = _lib()
    lib = L.lib()
    assert lib.rskx_set_tuning(b"no_such_knob", 1.0) == L.RSKX_ERR_ARG
    assert L.last_error().startswith("FLAM:rskx:tuning")
    assert lib.rskx_set_tuning(b"elim8_lds", 200 * 1024) == L.RSKX_ERR_ARG
    v = ctypes.c_double()
    assert lib.rskx_get_tuning(b"concurrent", ctypes.byref(v)) == 0 and v.value == 1


def _cb_call(L, *, kind=None, pkind=None, entries=True, fn=True, d=2, p=8):
    o = L.Opts()
    L.lib().rskx_opts_default(ctypes.byref(o))
    dummy = 16  # never dereferenced: validation fails first
    ent = L.ENTRIES_FN(lambda *a: 0) if entries else L.ENTRIES_FN()
    pfn = L.PROXY_FN(lambda *a: 0) if fn else L.PROXY_FN()
    kern = L.Kernel(kind if kind is not None else L.RSKX_KERNEL_CALLBACK, None, None, None, 1.0,
                    0.0, 0.0, ent, None)
    px = L.Proxy(pkind if pkind is not None else L.RSKX_PROXY_CALLBACK, p, 1.5, 0.0, None, pfn, None)
    h = ctypes.c_void_p()
    rc = L.lib().rskx_rskelf(dummy, d, 100, 64, 1e-6, ctypes.byref(kern), ctypes.byref(px),
                             ctypes.byref(o), ctypes.byref(h))
    return rc, L.last_error()


@pytest.mark.parametrize("kw,code,ident", [
    (dict(entries=False), "ARG", "FLAM:rskx:null"),
    (dict(fn=False), "ARG", "FLAM:rskx:null"),
    (dict(p=-1), "ARG", "FLAM:rskx:proxy"),
    (dict(d=1), "UNSUP", "FLAM:rskx:kernel"),
    (dict(pkind=1), "UNSUP", "FLAM:rskx:proxy"),           # callback matrix, device ring
])
def test_callback_kernel_arguments(kw, code, ident):
    """A host-callback matrix needs its entries function, a callback proxy
    with its function, and d = 2 or 3; checked before any device work."""
    L = _lib()
    rc, msg = _cb_call(L, **kw)
    assert rc == getattr(L, "RSKX_ERR_" + code) and msg.startswith(ident), msg


def test_callback_product_and_batch_refused():
    L = _lib()
    dummy = 16
    kern = L.Kernel(L.RSKX_KERNEL_CALLBACK, None, None, None, 1.0, 0.0, 0.0,
                    L.ENTRIES_FN(lambda *a: 0), None)
    rc = L.lib().rskx_kernel_matvec(dummy, 2, 10, ctypes.byref(kern), dummy, 10, dummy, 10, 1)
    assert rc == L.RSKX_ERR_UNSUP and L.last_error().startswith("FLAM:rskx:kernel")


def test_callback_trampolines_on_the_host():
    """The Python trampolines write A(I, J) column-major and the proxy
    block with its leading dimension, compact the kept neighbours in place,
    ask for a larger buffer when Kpxy has more rows, and hold a raised
    exception for rskelf to raise again."""
    import numpy as np
    from flam_amd.rskelf import _Callbacks
    L = _lib()
    M = np.arange(100.0).reshape(10, 10)
    xh = np.vstack([np.arange(10.0), np.zeros(10)])

    def pxy(x, slf, nbr, l, ctr):
        assert x is xh and l.shape == (2,) and ctr.shape == (2,)
        return M[:3, slf] + l[0], nbr[nbr % 2 == 0]

    cb = _Callbacks(lambda I, J: M[np.ix_(I, J)], pxy, xh, 2)
    I = (ctypes.c_int64 * 3)(1, 4, 7)
    J = (ctypes.c_int64 * 2)(0, 9)
    out = (ctypes.c_double * 6)()
    assert cb.entries(None, I, 3, J, 2, out) == 0
    assert list(out) == [M[1, 0], M[4, 0], M[7, 0], M[1, 9], M[4, 9], M[7, 9]]
    slf = (ctypes.c_int64 * 2)(2, 5)
    nbr = (ctypes.c_int64 * 4)(1, 2, 3, 4)
    nn, rows = ctypes.c_int64(4), ctypes.c_int64(0)
    lv, cv = (ctypes.c_double * 2)(0.5, 0.5), (ctypes.c_double * 2)(0, 0)
    K = (ctypes.c_double * 4)()
    assert cb.proxy_fn(None, slf, 2, nbr, ctypes.byref(nn), lv, cv, K, 2, ctypes.byref(rows)) == 0
    assert rows.value == 3 and nn.value == 4                    # too many rows: nothing written
    K = (ctypes.c_double * 6)()
    assert cb.proxy_fn(None, slf, 2, nbr, ctypes.byref(nn), lv, cv, K, 3, ctypes.byref(rows)) == 0
    assert rows.value == 3 and nn.value == 2 and list(nbr)[:2] == [2, 4]
    assert list(K) == [M[0, 2] + .5, M[1, 2] + .5, M[2, 2] + .5, M[0, 5] + .5, M[1, 5] + .5, M[2, 5] + .5]

    def boom(I, J):
        raise KeyError("from A")
    cb2 = _Callbacks(boom, pxy, xh, 2)
    assert cb2.entries(None, I, 3, J, 2, out) == 1 and isinstance(cb2.exc, KeyError)
