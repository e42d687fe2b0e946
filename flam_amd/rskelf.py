"""Python mirror of FLAM's rskelf interface over the HIP library.

    F = rskelf(A, x, occ, rank_or_tol, pxyfun=None, opts=None)   # rskelf/rskelf.m:109
    Y = rskelf_sv(F, X, trans='n')                                # rskelf/rskelf_sv.m:13
    Y = rskelf_mv(F, X, trans='n')                                # rskelf/rskelf_mv.m:13

A is a kernel description (LaplaceDL2D, ...) and pxyfun a proxy
description (ProxyRing, ...) for the matrices the device evaluates itself;
any other real matrix is given as the reference gives it, by Python
callables A(I, J) and pxyfun(x, slf, nbr, l, ctr) (0-based indices), which
the library calls on the host for every box (a slow path: the compression
and elimination still run on the device).  Device arrays are torch
tensors on the GPU.  Argument errors
raise FlamError carrying the reference's identifiers (FLAM:chksymm:...,
FLAM:id:..., FLAM:hypoct:...).  Like MATLAB, rskelf_sv / rskelf_mv return
a new array and leave X untouched.
"""
import ctypes
import math
from dataclasses import dataclass

import torch

from . import _lib as L


class FlamError(ValueError):
    """Error raised by the library; `ident` is the FLAM:* identifier."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code
        self.ident = msg.split(": ")[0] if ": " in msg else msg


def _check(rc):
    if rc != L.RSKX_OK:
        raise FlamError(rc, L.last_error())


def _on_torch_stream():
    """Run the library on torch's current stream, so its work is ordered
    after the torch work that made its inputs and before what reads its
    outputs (the sweeps and products return once their work is queued)."""
    _check(L.lib().rskx_set_stream(torch.cuda.current_stream().cuda_stream))


def _dev(t, dtype=torch.float64):
    if not isinstance(t, torch.Tensor):
        t = torch.as_tensor(t, dtype=dtype)
    if not t.is_cuda:
        t = t.to("cuda")
    return t.to(dtype).contiguous()


_RDT = {torch.float64: L.RSKX_F64, torch.float32: L.RSKX_F32,
        torch.complex128: L.RSKX_Z64, torch.complex64: L.RSKX_Z32}


def _rhs_dtype(F, X):
    """The element type a sweep of F takes X in: a complex factor takes
    complex right-hand sides (real ones are promoted), a real factor real
    ones; FP32 / complex-FP32 inputs keep their width."""
    if getattr(F, "cplx", False) or X.is_complex():
        if not getattr(F, "cplx", False):
            raise FlamError(L.RSKX_ERR_ARG, "FLAM:rskx:dtype: complex right-hand sides need a complex factor")
        return torch.complex64 if X.dtype == torch.complex64 else torch.complex128
    return X.dtype if X.dtype in (torch.float64, torch.float32) else torch.float64


@dataclass
class LaplaceDL2D:
    """A_ij = scale w_j ((x_i - x_j).nu_j) / (2 pi |x_i - x_j|^2), A_ii = diag_i.

    nu: (2, N) unit normals, w: (N,) weights, diag: (N,) diagonal.  This is
    the Afun of rskelf/test/ie_circle.m:85-103 (scale 1) and
    rskelf/test/ie_ellipse.m:103-123, written as data."""
    nu: torch.Tensor
    w: torch.Tensor
    diag: torch.Tensor
    scale: float = 1.0

    def _struct(self):
        self.nu = _dev(self.nu).reshape(2, -1)
        self.w = _dev(self.w).reshape(-1)
        self.diag = _dev(self.diag).reshape(-1)
        # column-major d x N = row-major N x d
        self._nu_cm = self.nu.t().contiguous()
        return L.Kernel(L.RSKX_KERNEL_LAPLACE2D_DL, self._nu_cm.data_ptr(), self.w.data_ptr(),
                        self.diag.data_ptr(), float(self.scale))


@dataclass
class LaplaceSL3D:
    """A_ij = scale w_j / (4 pi |x_i - x_j|), A_ii = diag_i, points in 3D.

    w: (N,) quadrature weights (triangle areas), diag: (N,) self-panel
    integrals.  The single-layer kernel of rskelf/test/ie_sphere.m:149-160
    with one-point quadrature off the diagonal."""
    w: torch.Tensor
    diag: torch.Tensor
    scale: float = 1.0
    nu = None

    def _struct(self):
        self.w = _dev(self.w).reshape(-1)
        self.diag = _dev(self.diag).reshape(-1)
        return L.Kernel(L.RSKX_KERNEL_LAPLACE3D_SL, None, self.w.data_ptr(),
                        self.diag.data_ptr(), float(self.scale))


@dataclass
class HelmholtzCFIE2D:
    """Exterior Dirichlet Helmholtz combined field, complex:
    A_ij = w_j (dG/dn_j(x_i, x_j) - i eta G(x_i, x_j)), G = (i/4) H0^(1)(k r)
    (the kernel of rskelf/test/ie_square3.m:44,107-110), A_ii = diag_i
    (complex).  nu: (2, N) unit normals, w: (N,) weights, diag: (N,) complex
    diagonal; k the wavenumber, eta the single layer's coupling (k by
    default).  Pairs with ProxyRing(kind='cfie')."""
    nu: torch.Tensor
    w: torch.Tensor
    diag: torch.Tensor
    k: float
    eta: float = None

    def _struct(self):
        self.nu = _dev(self.nu).reshape(2, -1)
        self.w = _dev(self.w).reshape(-1)
        self.diag = _dev(self.diag, torch.complex128).reshape(-1)
        self._nu_cm = self.nu.t().contiguous()
        eta = float(self.k if self.eta is None else self.eta)
        return L.Kernel(L.RSKX_KERNEL_HELMHOLTZ2D_CFIE, self._nu_cm.data_ptr(), self.w.data_ptr(),
                        self.diag.data_ptr(), 1.0, float(self.k), eta)


@dataclass
class ProxyPoints:
    """pxyfun of rskelf/test/ie_sphere.m:174-187: proxy points pts (3, p) in
    box-size units about the box centre (pxy = pts l + ctr), single-layer
    rows weighted by the column's w (outgoing) and, for symm 'n', by
    `weight` (incoming, 4 pi / N there); neighbours kept inside `radius`
    box sizes (1.5)."""
    pts: torch.Tensor
    radius: float = 1.5
    weight: float = 0.0

    @property
    def p(self):
        return int(self.pts.shape[1])

    def _struct(self):
        self.pts = _dev(self.pts).reshape(3, -1)
        self._pts_cm = self.pts.t().contiguous()      # column-major 3 x p
        return L.Proxy(L.RSKX_PROXY_POINTS_SL, self.p, float(self.radius), float(self.weight),
                       self._pts_cm.data_ptr())


@dataclass
class ProxyRing:
    """pxyfun of the demos: p points on a ring of radius `radius` box sizes
    around the box centre; neighbours inside the ring are kept explicitly.
    kind 'sl': single layer times `weight` (ie_circle.m:106-115);
    kind 'dl': double layer of the box points (ie_ellipse.m:126-132);
    kind 'cfie': the combined-field ring of HelmholtzCFIE2D, at least p
    points and more as k times the box size grows, incoming rows weighted
    by `weight` (the mean quadrature weight)."""
    p: int = 64
    radius: float = 1.5
    kind: str = "sl"
    weight: float = 0.0

    def _struct(self):
        kind = {"sl": L.RSKX_PROXY_RING_SL, "dl": L.RSKX_PROXY_RING_DL,
                "cfie": L.RSKX_PROXY_RING_CFIE}[self.kind]
        return L.Proxy(kind, int(self.p), float(self.radius), float(self.weight), None)


_PREC = {"fp64": L.RSKX_PREC_FP64, "fp32": L.RSKX_PREC_FP32}


class Factor:
    """Device-resident factorization (opaque handle of the C ABI)."""

    def __init__(self, handle, N, keep, cplx=False):
        self._h = handle
        self.N = N
        self._keep = keep
        self.cplx = cplx

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                L.lib().rskx_free(h)
            except Exception:
                pass
            self._h = None

    @property
    def handle(self):
        return self._h

    def info(self):
        inf = L.Info()
        _check(L.lib().rskx_factor_info(self._h, ctypes.byref(inf)))
        return dict(N=inf.N, d=inf.d, nlvl=inf.nlvl, nlvl_done=inf.nlvl_done,
                    nfactor=inf.nfactor, bytes=inf.bytes,
                    precision="fp64" if inf.precision == L.RSKX_PREC_FP64 else "fp32",
                    symm=inf.symm.decode(), id_scratch=inf.id_scratch,
                    rank_fallbacks=inf.rank_fallbacks)

    def box_features(self):
        """(X, rank): the rank model's features (nbox x 8, numpy) and the
        exact rank of every box the factorization compressed (recorded with
        opts record_boxes or a rank model)."""
        import numpy as np
        n = ctypes.c_int64()
        _check(L.lib().rskx_box_features(self._h, ctypes.byref(n), None, None))
        X = np.zeros((n.value, 8))
        k = np.zeros(n.value, np.int64)
        _check(L.lib().rskx_box_features(self._h, ctypes.byref(n), X.ctypes.data, k.ctypes.data))
        return X, k

    def level_table(self):
        """(lvl, nblk, start npts, end npts) per level, as rskelf.m:264-272 prints."""
        n = L.lib().rskx_level_table(self._h, None, 0)
        rows = (ctypes.c_int64 * (4 * max(n, 1)))()
        L.lib().rskx_level_table(self._h, rows, n)
        return [tuple(rows[4 * i: 4 * i + 4]) for i in range(n)]

    def export(self, i):
        """Host copy of F.factors(i) (0-based): dict sk, rd, T, L, U, p, E, F
        as numpy arrays (indices 0-based)."""
        import numpy as np
        v = L.FactorView()
        _check(L.lib().rskx_export_factor(self._h, i, ctypes.byref(v)))
        k, r = v.k, v.r
        lu = self.info()["symm"] in ("n", "s")
        dt = np.complex128 if self.cplx else np.float64   # complex: interleaved (re, im)
        arr = dict(sk=np.zeros(k, np.int64), rd=np.zeros(r, np.int64),
                   T=np.zeros((r, k), dt), L=np.zeros((r, r), dt), U=np.zeros((r, r), dt),
                   p=np.zeros(r, np.int64), E=np.zeros((r, k), dt), F=np.zeros((k, r), dt))

        def ptr(a, t):
            return a.ctypes.data_as(ctypes.POINTER(t))
        v.sk, v.rd, v.p = ptr(arr["sk"], ctypes.c_int64), ptr(arr["rd"], ctypes.c_int64), ptr(arr["p"], ctypes.c_int64)
        for key in ("T", "L", "U", "E", "F"):
            setattr(v, key, ptr(arr[key], ctypes.c_double))
        _check(L.lib().rskx_export_factor(self._h, i, ctypes.byref(v)))
        ldl = self.info()["symm"] == "h"            # U = D for 'h' (rskelf.m:231)
        out = dict(sk=arr["sk"], rd=arr["rd"], T=arr["T"].T.copy(), L=arr["L"].T.copy(),
                   E=arr["E"].T.copy(), U=arr["U"].T.copy() if (lu or ldl) else None,
                   p=arr["p"] if lu else None, F=arr["F"].T.copy() if lu else None)
        return out

    def kernel_times(self):
        n = L.lib().rskx_kernel_times(self._h, None, None, None, None, None, 0)
        names = (ctypes.c_char_p * max(n, 1))()
        ms = (ctypes.c_double * max(n, 1))()
        cnt = (ctypes.c_int64 * max(n, 1))()
        byt = (ctypes.c_int64 * max(n, 1))()
        fl = (ctypes.c_double * max(n, 1))()
        L.lib().rskx_kernel_times(self._h, names, ms, cnt, byt, fl, n)
        return {names[i].decode(): dict(ms=ms[i], launches=cnt[i], bytes=byt[i], flops=fl[i])
                for i in range(n)}


def _opts(opts, d):
    """rskx_opts from the options dict (rskelf.m:112-120 names and defaults);
    returns the struct and the host buffer ext points into."""
    opts = dict(opts or {})
    o = L.Opts()
    L.lib().rskx_opts_default(ctypes.byref(o))
    lvlmax = opts.get("lvlmax", math.inf)
    o.lvlmax = 0 if lvlmax == math.inf else int(lvlmax)
    ext = opts.get("ext")
    ext_buf = None
    if ext is not None:
        import numpy as np
        e = np.asarray(ext, dtype=np.float64).reshape(d, 2)
        ext_buf = (ctypes.c_double * (2 * d))(*e.T.ravel())   # lower bounds, then upper
        o.ext = ext_buf
    o.Tmax = float(opts.get("Tmax", 2.0))
    o.rrqr_iter = float(opts.get("rrqr_iter", math.inf))
    o.symm = str(opts.get("symm", "n")).encode()[:1]
    o.verb = int(opts.get("verb", 0))
    prec = opts.get("precision", "fp64")
    if prec not in _PREC:
        raise FlamError(L.RSKX_ERR_ARG, "FLAM:rskx:precision: precision must be 'fp64' or 'fp32'")
    o.precision = _PREC[prec]
    o.profile = int(opts.get("profile", 0))
    stop = opts.get("stop", math.inf)
    if callable(stop):
        raise FlamError(L.RSKX_ERR_UNSUP, "FLAM:rskx:stop: only a numeric opts.stop (levels) is built")
    o.stop = -1 if stop is None or stop == math.inf else int(stop)
    route = opts.get("id_route", "auto")
    if route not in ("auto", "gram", "qr"):
        raise FlamError(L.RSKX_ERR_ARG, "FLAM:rskx:id_route: id_route must be 'auto', 'gram' or 'qr'")
    o.id_gram = {"auto": 0, "gram": 1, "qr": -1}[route]
    model = opts.get("rank_model")
    o.rank_model = model.handle if model is not None else None
    o.rank_scale = float(opts.get("rank_scale", 1.0))
    o.record_boxes = int(bool(opts.get("record_boxes", False)))
    return o, ext_buf


class _Callbacks:
    """ctypes trampolines for a matrix given by Python callables: A(I, J)
    returns the (len(I), len(J)) block, pxyfun(x, slf, nbr, l, ctr) returns
    (Kpxy, kept neighbours) as rskelf.m:205 calls it (x: the d x N host
    points).  An exception inside a callback aborts the factorization and
    is raised again by rskelf."""

    def __init__(self, A, pxyfun, xh, prows):
        import numpy as np
        self.np = np
        self.A, self.pxyfun, self.xh = A, pxyfun, xh
        self.exc = None
        self.pending = None                 # a proxy block too large for the buffer
        self.entries = L.ENTRIES_FN(self._entries)
        self.proxy_fn = L.PROXY_FN(self._proxy)
        self.prows = int(prows)

    def _entries(self, ctx, I, nI, J, nJ, out):
        np = self.np
        try:
            Ii = np.ctypeslib.as_array(I, shape=(nI,)).copy() if nI else np.zeros(0, np.int64)
            Jj = np.ctypeslib.as_array(J, shape=(nJ,)).copy() if nJ else np.zeros(0, np.int64)
            blk = np.asarray(self.A(Ii, Jj), dtype=np.float64).reshape(nI, nJ)
            if nI and nJ:
                np.ctypeslib.as_array(out, shape=(nJ, nI))[...] = blk.T
            return 0
        except BaseException as e:          # noqa: BLE001 - raised again after the call
            self.exc = self.exc or e
            return 1

    def _proxy(self, ctx, slf, nslf, nbr, nnbr, l, ctr, Kpxy, maxrows, rows):
        np = self.np
        try:
            d = self.xh.shape[0]
            s_ = np.ctypeslib.as_array(slf, shape=(nslf,)).copy()
            n0 = nnbr[0]
            nb = np.ctypeslib.as_array(nbr, shape=(n0,)).copy() if n0 else np.zeros(0, np.int64)
            lv = np.ctypeslib.as_array(l, shape=(d,)).copy()
            cv = np.ctypeslib.as_array(ctr, shape=(d,)).copy()
            K, kept = self.pxyfun(self.xh, s_, nb, lv, cv)
            K = np.asarray(K, dtype=np.float64).reshape(-1, nslf)
            kept = np.asarray(kept, dtype=np.int64).reshape(-1)
            rows[0] = K.shape[0]
            if K.shape[0] > maxrows:        # called again with a larger buffer
                return 0
            if len(kept):
                np.ctypeslib.as_array(nbr, shape=(n0,))[:len(kept)] = kept
            nnbr[0] = len(kept)
            if K.shape[0] and nslf:
                np.ctypeslib.as_array(Kpxy, shape=(nslf, maxrows))[:, :K.shape[0]] = K.T
            return 0
        except BaseException as e:          # noqa: BLE001
            self.exc = self.exc or e
            return 1

    def structs(self):
        k = L.Kernel(L.RSKX_KERNEL_CALLBACK, None, None, None, 1.0, 0.0, 0.0, self.entries, None)
        p = L.Proxy(L.RSKX_PROXY_CALLBACK, self.prows, 1.5, 0.0, None, self.proxy_fn, None)
        return k, p


def rskelf(A, x, occ, rank_or_tol, pxyfun=None, opts=None):
    """F = rskelf(A, x, occ, rank_or_tol, pxyfun, opts) on the GPU.

    opts keys (rskelf.m:112-120): lvlmax (inf), ext (None), Tmax (2),
    rrqr_iter (inf), symm ('n'), verb (0), stop (inf: a numeric stop
    factors that many levels, rskelf.m:128,172); plus precision ('fp64' or
    'fp32'), profile (0), id_route ('auto': the Gram-matrix ID for FP32
    factors at tol >= 1e-7 and, in FP64, for boxes beyond the register
    Gram classes (n > 284, the top of 3D trees), column-pivoted QR
    otherwise; 'gram': the Gram ID for every FP64 box too; 'qr': always
    QRCP), rank_model (a trained
    flam_amd.rank.RankModel: the Gram ID's buffers sized by its predicted
    ranks, boxes beyond their cap redone with full buffers), rank_scale (1)
    and record_boxes (False: keep features and exact ranks of every box,
    Factor.box_features()).

    A and pxyfun may instead be Python callables, as the reference takes
    them: A(I, J) -> the (len(I), len(J)) block of the real matrix,
    pxyfun(x, slf, nbr, l, ctr) -> (Kpxy, kept neighbours), with 0-based
    index arrays and x the d x N host points; opts proxy_rows (64) is the
    first size of the Kpxy buffer (grown on request).  The library calls
    them on the host for every box (rskelf.m:205,210-219) and runs the
    compression and elimination on the device."""
    cbs = None
    if not hasattr(A, "_struct"):
        if not callable(A) or not callable(pxyfun):
            raise FlamError(L.RSKX_ERR_ARG, "FLAM:rskx:kernel: A must be a kernel description, or "
                                            "A(I, J) and pxyfun both callables")
        xh = x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else __import__("numpy").asarray(x)
        cbs = _Callbacks(A, pxyfun, xh.astype("float64"), (opts or {}).get("proxy_rows", 64))
    x = _dev(x)
    if x.dim() != 2:
        raise FlamError(L.RSKX_ERR_ARG, "FLAM:rskx:size: x must be d x N")
    d, N = x.shape
    xc = x.t().contiguous()                # column-major d x N
    o, ext_buf = _opts(opts, d)
    if cbs is not None:
        k, p = cbs.structs()
    else:
        k = A._struct()
        p = pxyfun._struct() if pxyfun is not None else L.Proxy(L.RSKX_PROXY_NONE, 0, 1.5, 0.0)
    h = ctypes.c_void_p()
    _on_torch_stream()
    rc = L.lib().rskx_rskelf(xc.data_ptr(), d, N, int(occ), float(rank_or_tol),
                             ctypes.byref(k), ctypes.byref(p), ctypes.byref(o), ctypes.byref(h))
    if cbs is not None and cbs.exc is not None:
        raise cbs.exc
    _check(rc)
    return Factor(h.value, N, keep=(A, xc, opts and opts.get("rank_model")),
                  cplx=isinstance(A, HelmholtzCFIE2D))


@dataclass
class Batch:
    """Independent problems stacked for rskelf_batch: the points, normals,
    weights and diagonals of every problem concatenated (problem j owns
    columns offsets[j]:offsets[j+1]), and one proxy ring per problem."""
    x: torch.Tensor
    kernel: LaplaceDL2D
    offsets: list
    proxies: list

    @property
    def N(self):
        return self.offsets[-1]

    def split(self, X):
        """The rows of each problem of a stacked (N, ...) array."""
        return [X[a:b] for a, b in zip(self.offsets[:-1], self.offsets[1:])]


def stack_problems(problems):
    """Batch of problems given as objects with x (d x N_j), kernel (all
    LaplaceDL2D or all LaplaceSL3D, one common scale) and proxy."""
    scales = {float(p.kernel.scale) for p in problems}
    if len(scales) != 1:
        raise FlamError(L.RSKX_ERR_UNSUP, "FLAM:rskx:batch: the problems' kernels must share one scale")
    offs = [0]
    for p in problems:
        offs.append(offs[-1] + int(p.x.shape[1]))
    kinds = {type(p.kernel) for p in problems}
    if len(kinds) != 1:
        raise FlamError(L.RSKX_ERR_UNSUP, "FLAM:rskx:batch: the problems' kernels must be of one kind")
    x = torch.cat([_dev(p.x) for p in problems], dim=1)
    w = torch.cat([_dev(p.kernel.w).reshape(-1) for p in problems])
    diag = torch.cat([_dev(p.kernel.diag).reshape(-1) for p in problems])
    if kinds.pop() is LaplaceSL3D:
        kern = LaplaceSL3D(w, diag, scales.pop())
    else:
        nu = torch.cat([_dev(p.kernel.nu).reshape(2, -1) for p in problems], dim=1)
        kern = LaplaceDL2D(nu, w, diag, scales.pop())
    return Batch(x, kern, offs, [p.proxy for p in problems])


def rskelf_batch(batch, occ, rank_or_tol, opts=None):
    """One factorization of many independent problems (rskx_rskelf_batch):
    the factor of the block-diagonal matrix diag(A_1, ..., A_P), each
    problem's tree factored level by level together with the others'.
    rskelf_sv / rskelf_mv apply it to the stacked right-hand sides."""
    x = _dev(batch.x)
    d, N = x.shape
    if N != batch.N:
        raise FlamError(L.RSKX_ERR_ARG, "FLAM:rskx:batch: offsets do not match the points")
    xc = x.t().contiguous()
    o, ext_buf = _opts(opts, d)
    k = batch.kernel._struct()
    nprob = len(batch.offsets) - 1
    offs = (ctypes.c_int64 * (nprob + 1))(*batch.offsets)
    structs = [p._struct() for p in batch.proxies]
    if structs[0].kind == L.RSKX_PROXY_POINTS_SL:
        # the library takes one proxy point set for the batch (weights may differ)
        p0 = batch.proxies[0]
        for p, st in zip(batch.proxies[1:], structs[1:]):
            if p.pts.shape != p0.pts.shape or not torch.equal(p.pts, p0.pts):
                raise FlamError(L.RSKX_ERR_UNSUP, "FLAM:rskx:batch: the problems' proxy "
                                                  "surfaces must share their points")
            st.pts = structs[0].pts
    pxs = (L.Proxy * nprob)(*structs)
    h = ctypes.c_void_p()
    _on_torch_stream()
    _check(L.lib().rskx_rskelf_batch(nprob, offs, xc.data_ptr(), d, int(occ), float(rank_or_tol),
                                     ctypes.byref(k), pxs, ctypes.byref(o), ctypes.byref(h)))
    return Factor(h.value, N, keep=(batch.kernel, xc))


def kernel_matvec_batch(batch, X):
    """FP64 Y = diag(A_1, ..., A_P) X, each kernel evaluated on the fly."""
    x = _dev(batch.x)
    d, N = x.shape
    xc = x.t().contiguous()
    X = _dev(X)
    vec = X.dim() == 1
    X2 = X.reshape(-1, 1) if vec else X
    Xc = X2.t().contiguous()
    Yc = torch.empty_like(Xc)
    k = batch.kernel._struct()
    nprob = len(batch.offsets) - 1
    offs = (ctypes.c_int64 * (nprob + 1))(*batch.offsets)
    _on_torch_stream()
    _check(L.lib().rskx_kernel_matvec_batch(nprob, offs, xc.data_ptr(), d, ctypes.byref(k),
                                            Xc.data_ptr(), N, Yc.data_ptr(), N, Xc.shape[0]))
    Y = Yc.t()
    return Y.reshape(-1) if vec else Y


def _sweep(fn, F, X, trans):
    X = X if isinstance(X, torch.Tensor) else torch.as_tensor(X)
    if not X.is_cuda:
        X = X.to("cuda")
    X = X.to(_rhs_dtype(F, X))
    vec = X.dim() == 1
    X2 = X.reshape(-1, 1) if vec else X
    if X2.shape[0] != F.N:
        raise FlamError(L.RSKX_ERR_ARG, "FLAM:rskx:size: X must have N rows")
    # column-major N x nrhs; always a fresh copy (a (1, N) view of a vector
    # is already "contiguous" and would alias the caller's storage)
    Xc = X2.t().clone(memory_format=torch.contiguous_format)
    dt = _RDT[Xc.dtype]
    t = str(trans).encode()[:1] if trans else b"n"
    _on_torch_stream()
    _check(fn(F.handle, Xc.data_ptr(), F.N, Xc.shape[0], t, dt))
    Y = Xc.t()
    return Y.reshape(-1) if vec else Y


def _sweep_mode(fn, F, X, trans, mode):
    """One sweep (mode 1 upward, 2 downward) on a fresh column-major copy."""
    X2 = X if X.dim() == 2 else X.reshape(-1, 1)
    Xc = X2.t().clone(memory_format=torch.contiguous_format)
    dt = _RDT[Xc.dtype]
    _on_torch_stream()
    _check(fn(F.handle, Xc.data_ptr(), F.N, Xc.shape[0], str(trans).encode()[:1], dt, mode))
    return Xc.t()


def rskelf_partial_info(F):
    """[sk, S] = rskelf_partial_info(F) (rskelf/rskelf_partial_info.m):
    the points left by a partial factorization (ascending, 0-based, a CPU
    int64 tensor) and the sparse modification S (CPU, FP64, COO) of their
    system, which is A(sk, sk) + S."""
    nsk, nnz = ctypes.c_int64(), ctypes.c_int64()
    _check(L.lib().rskx_partial_info(F.handle, ctypes.byref(nsk), ctypes.byref(nnz),
                                     None, None, None, None))
    sk = torch.empty(nsk.value, dtype=torch.int64)
    I = torch.empty(nnz.value, dtype=torch.int64)
    J = torch.empty(nnz.value, dtype=torch.int64)
    V = torch.empty(nnz.value, dtype=torch.float64)
    p64 = lambda t: ctypes.cast(t.data_ptr(), ctypes.POINTER(ctypes.c_int64))
    _check(L.lib().rskx_partial_info(F.handle, None, None, p64(sk), p64(I), p64(J),
                                     ctypes.cast(V.data_ptr(), ctypes.POINTER(ctypes.c_double))))
    S = torch.sparse_coo_tensor(torch.stack([I, J]), V, (nsk.value, nsk.value))
    return sk, S


def _partial(fn, F, X, fun, trans):
    X = X if isinstance(X, torch.Tensor) else torch.as_tensor(X)
    if not X.is_cuda:
        X = X.to("cuda")
    if X.dtype not in (torch.float64, torch.float32):
        X = X.to(torch.float64)
    vec = X.dim() == 1
    X2 = X.reshape(-1, 1) if vec else X
    if X2.shape[0] != F.N:
        raise FlamError(L.RSKX_ERR_ARG, "FLAM:rskx:size: X must have N rows")
    t = str(trans).lower()
    if t not in ("n", "t", "c"):
        raise FlamError(L.RSKX_ERR_ARG, "FLAM:chktrans:invalidTrans: Transpose parameter must be one of 'N', 'T', or 'C'.")
    if t == "t":                        # real data: transpose = conjugate transpose
        t = "c"
    sk, _ = rskelf_partial_info(F)
    Y = _sweep_mode(fn, F, X2, t, 1)
    if sk.numel():
        skd = sk.to(Y.device)
        Y = Y.clone()
        Y[skd] = fun(Y[skd], t).to(Y.dtype)
    Y = _sweep_mode(fn, F, Y, t, 2)
    return Y.reshape(-1) if vec else Y


def rskelf_partial_sv(F, X, svfun, trans="n"):
    """rskelf/rskelf_partial_sv.m:35-61: upward sweep, svfun on the remaining
    skeletons (svfun(Y, trans) = (A(sk,sk) + S)^{-1} Y, or its transpose),
    downward sweep."""
    return _partial(L.lib().rskx_sweep_sv, F, X, svfun, trans)


def rskelf_partial_mv(F, X, mvfun, trans="n"):
    """rskelf/rskelf_partial_mv.m:35-61, with mvfun applying A(sk,sk) + S."""
    return _partial(L.lib().rskx_sweep_mv, F, X, mvfun, trans)


def rskelf_sv(F, X, trans="n"):
    """Apply F^{-1} (rskelf/rskelf_sv.m:13-36)."""
    return _sweep(L.lib().rskx_rskelf_sv, F, X, trans)


def rskelf_sv_(F, Xt, trans="n"):
    """In-place F^{-1} on a column-major N x nrhs block given as its
    (nrhs, N) row view Xt: unit stride along N, rows any stride >= N (a
    slice of a larger array; the C ABI's leading dimension).  Returns Xt."""
    if (Xt.dim() != 2 or not Xt.is_cuda or Xt.shape[1] != F.N or Xt.stride(1) != 1
            or Xt.stride(0) < F.N or Xt.dtype not in _RDT
            or Xt.is_complex() != getattr(F, "cplx", False)):
        raise FlamError(L.RSKX_ERR_ARG, "FLAM:rskx:size: Xt must be an (nrhs, N) CUDA view "
                                        "with unit stride along N, complex for a complex factor")
    dt = _RDT[Xt.dtype]
    _on_torch_stream()
    _check(L.lib().rskx_rskelf_sv(F.handle, Xt.data_ptr(), Xt.stride(0), Xt.shape[0],
                                  str(trans).encode()[:1], dt))
    return Xt


def rskelf_mv(F, X, trans="n"):
    """Apply F (rskelf/rskelf_mv.m:13-36)."""
    return _sweep(L.lib().rskx_rskelf_mv, F, X, trans)


def rskelf_logdet(F):
    """log det of the factored matrix, imaginary part in [0, 2 pi)
    (rskelf/rskelf_logdet.m:11)."""
    re, im = ctypes.c_double(), ctypes.c_double()
    _on_torch_stream()
    _check(L.lib().rskx_rskelf_logdet(F.handle, ctypes.byref(re), ctypes.byref(im)))
    return complex(re.value, im.value)


def rskelf_cholmv(F, X, trans="n"):
    """C X (or C' X) for the generalized Cholesky factor of F = C C', symm
    'p' (rskelf/rskelf_cholmv.m:15)."""
    return _sweep(L.lib().rskx_rskelf_cholmv, F, X, trans)


def rskelf_cholsv(F, X, trans="n"):
    """C \\ X (or C' \\ X), symm 'p' (rskelf/rskelf_cholsv.m:15)."""
    return _sweep(L.lib().rskx_rskelf_cholsv, F, X, trans)


def kernel_matvec(A, x, X):
    """FP64 Y = A X with the kernel evaluated on the fly (no compression)."""
    x = _dev(x)
    d, N = x.shape
    xc = x.t().contiguous()
    X = _dev(X, torch.complex128 if isinstance(A, HelmholtzCFIE2D) else torch.float64)
    vec = X.dim() == 1
    X2 = X.reshape(-1, 1) if vec else X
    Xc = X2.t().contiguous()
    Yc = torch.empty_like(Xc)
    k = A._struct()
    _on_torch_stream()
    _check(L.lib().rskx_kernel_matvec(xc.data_ptr(), d, N, ctypes.byref(k), Xc.data_ptr(), N,
                                      Yc.data_ptr(), N, Xc.shape[0]))
    Y = Yc.t()
    return Y.reshape(-1) if vec else Y


def release_cache():
    """Give the library's idle device memory (workspaces and the cache of
    freed factor blocks) back to the device; live factors are untouched."""
    _check(L.lib().rskx_release_cache())


def rskelf_import(F, N=None, precision="fp64"):
    """A device factor from one in the reference's layout (F.factors,
    F.lvp, F.symm of rskelf.m:156-158,249-257,263) -- e.g. the oracle's,
    or a factor exported with Factor.export: F is a dict with 'factors'
    (dicts sk, rd, T, L, U, p, E, F; 0-based indices), 'lvp' and 'symm'
    (and 'N').  The inverse of Factor.export (rskx_import)."""
    import numpy as np
    facs = F["factors"]
    lvp = np.ascontiguousarray(F["lvp"], dtype=np.int64)
    symm = str(F["symm"]).lower()
    N = int(F.get("N", N) if N is None else N)
    keep = []
    views = (L.FactorView * max(len(facs), 1))()

    def arr(a, dt):
        if a is None:
            return None
        a = np.asfortranarray(np.asarray(a, dtype=dt))
        keep.append(a)
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64 if dt == np.int64 else ctypes.c_double))
    for i, f in enumerate(facs):
        v = views[i]
        v.k, v.r = len(f["sk"]), len(f["rd"])
        v.sk, v.rd = arr(f["sk"], np.int64), arr(f["rd"], np.int64)
        v.p = arr(f.get("p"), np.int64)
        for key in ("T", "L", "U", "E", "F"):
            setattr(v, key, arr(f.get(key), np.float64))
    h = ctypes.c_void_p()
    _check(L.lib().rskx_import(N, symm.encode()[:1], _PREC[precision], len(lvp) - 1,
                               lvp.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                               ctypes.cast(views, ctypes.c_void_p), ctypes.byref(h)))
    return Factor(h.value, N, keep=None)


def export_all(F):
    """The whole factor in the reference's layout: dict(N, symm, lvp,
    factors) with every F.factors(i) from Factor.export and F.lvp
    (cumulative factor counts per level, leaves first)."""
    info = F.info()
    n = L.lib().rskx_factor_lvp(F.handle, None, 0)
    lvp = (ctypes.c_int64 * max(n, 1))()
    L.lib().rskx_factor_lvp(F.handle, lvp, n)
    return dict(N=info["N"], symm=info["symm"], lvp=list(lvp[:n]),
                factors=[F.export(i) for i in range(info["nfactor"])])
