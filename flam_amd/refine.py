"""Krylov refinement with an rskelf preconditioner.

The reference solves with rskelf as a preconditioner inside MATLAB's gmres
(rskelf/test/ie_square3.m:97: gmres(mv,B,32,1e-12,32,@(x)rskelf_sv(F,x))).
Here the preconditioner is usually a low-precision (FP32) factorization and
the operator an exact FP64 product, so this is flexible GMRES (the
preconditioned directions are kept, so a preconditioner that is linear only
up to FP32 rounding does not limit the attainable FP64 residual), run on
all right-hand sides at once and stopped on the true relative residual
||b - A x|| / ||b|| of every column.

Vectors are stored one right-hand side per row, (nrhs, N), which is the
column-major N x nrhs layout the sweeps take; the Krylov basis of every
column is one (nrhs, restart + 1, N) array, so each orthogonalization pass
is one batched product.  Classical Gram-Schmidt is applied twice per step,
by the library's fused kernels (rskx_cgs2): each pass reads the basis and
the new vector once, with no temporaries.
"""
import numpy as np
import torch

from . import _lib as L
from .rskelf import Factor, rskelf_mv, rskelf_sv_


def factor_operator(F):
    """A X through a factor's forward sweep (rskelf_mv, rskelf/rskelf_mv.m:13).

    An FP64 factor built at a tight tolerance reproduces A to about that
    tolerance at the cost of one sweep, so it serves as the accurate
    operator of the refinement where A has no fast exact product (the
    reference's demos use an ifmm for that, rskelf/test/ie_sphere.m:99-125;
    this is the other accurate operator the library can offer).  The
    attained residual then holds for this operator; the caller confirms it
    against the exact product."""
    return lambda V: rskelf_mv(F, V)


def _cgs2(V, nv, w):
    """Orthogonalize w (nrhs, N) against V[:, :nv, :] twice, in place; returns
    a (nrhs, nv + 1) CPU tensor of the projections and the new norm."""
    nrhs, nb, N = V.shape
    hout = torch.empty(nrhs, nv + 1, dtype=torch.float64, device=V.device)
    lib = L.lib()
    lib.rskx_set_stream(torch.cuda.current_stream().cuda_stream)
    rc = lib.rskx_cgs2(V.data_ptr(), N, nb * N, nv, w.data_ptr(), N, N, nrhs, hout.data_ptr())
    if rc != L.RSKX_OK:
        raise RuntimeError(L.last_error())
    return hout


def _cgs2_torch(V, nv, w):
    """The same orthogonalization for complex vectors, with torch's batched
    products (the fused kernel is real): h = V' w twice, w -= V h, and the
    new norm; returns (nrhs, nv + 1) with the norm last."""
    Vj = V[:, :nv, :].contiguous()
    Vc = Vj.conj().resolve_conj()                               # (nrhs, nv, N)
    h1 = torch.bmm(Vc, w.unsqueeze(2)).squeeze(2)               # V' w
    w -= torch.bmm(h1.unsqueeze(1), Vj).squeeze(1)
    h2 = torch.bmm(Vc, w.unsqueeze(2)).squeeze(2)
    w -= torch.bmm(h2.unsqueeze(1), Vj).squeeze(1)
    nrm = _norms(w)
    return torch.cat([h1 + h2, nrm.to(w.dtype).unsqueeze(1)], dim=1)


def _rowsum(P):
    """Sum over the last dimension in two stages (1024 partial sums per
    row first), so a few very long rows still spread over the whole GPU."""
    n = P.shape[-1]
    if n % 1024 == 0 and n >= 1 << 16:
        return P.reshape(*P.shape[:-1], 1024, n // 1024).sum(-1).sum(-1)
    return P.sum(-1)


def _norms(W):
    if W.is_complex():
        return torch.sqrt(_rowsum(W.real * W.real + W.imag * W.imag))
    return torch.sqrt(_rowsum(W * W))


def _seg_norms(W, segments):
    """(nrhs, len(segments)) norms of the row blocks W[:, a:b]."""
    return torch.stack([_norms(W[:, a:b]) for a, b in segments], dim=1)


def fgmres(matvec, B, precond, tol=1e-10, restart=16, maxit=8, segments=None, inner_matvec=None):
    """Solve A X = B.  matvec, precond: callables on (N, nrhs) FP64 CUDA
    tensors; precond may also be a Factor, whose solve then runs in place
    on the stored preconditioned direction, or a graph.SolveGraph (the
    same solve replayed from a HIP graph).  The small least-squares
    problems are kept in QR form by Givens rotations, one new rotation per
    step, on the host.  segments: optional row ranges [(a, b), ...] of independent
    problems stacked in one system (a block-diagonal A); then every
    problem's own relative residual must reach tol.  inner_matvec: an
    optional cheaper operator close to A for the Arnoldi steps (inexact
    Krylov); convergence is still decided on true residuals B - A X taken
    with matvec after each cycle, so the result meets tol for matvec's A.
    Returns (X, info) with info = dict(iters, outer, relres)."""
    vec = B.dim() == 1
    cplx = B.is_complex() or getattr(precond, "cplx", False)
    dt = torch.complex128 if cplx else torch.float64
    B2 = B.reshape(B.shape[0], -1).to(dt)
    N, nrhs = B2.shape
    Bt = B2.t().contiguous()                       # (nrhs, N)
    bnorm = _norms(Bt)
    bnorm = torch.where(bnorm > 0, bnorm, torch.ones_like(bnorm))
    bnorm_h = bnorm.cpu().numpy()
    if segments is not None:
        bseg = _seg_norms(Bt, segments)
        bseg = torch.where(bseg > 0, bseg, torch.ones_like(bseg))
        # the Arnoldi estimate bounds the stacked residual, which bounds
        # every block's: stop it at the smallest block's share
        inner_ref = bseg.min(dim=1).values.cpu().numpy()
    else:
        inner_ref = bnorm_h
    Xt = torch.zeros_like(Bt)
    V = torch.empty(nrhs, restart + 1, N, dtype=dt, device=B2.device)
    Z = torch.empty(nrhs, restart, N, dtype=dt, device=B2.device)
    hdt = np.complex128 if cplx else np.float64

    def op(f, Yt):                                 # (nrhs, N) -> (nrhs, N)
        return f(Yt.t()).t().contiguous()

    iters = 0
    relres = None
    outer = 0
    for outer in range(maxit + 1):
        R = Bt - op(matvec, Xt) if outer else Bt.clone()
        rn = _norms(R)
        if segments is not None:
            relres = (_seg_norms(R, segments) / bseg).max(dim=1).values.cpu()
        else:
            relres = (rn / bnorm).cpu()
        if bool((relres <= tol).all()) or outer == maxit:
            break
        # a column whose residual is exactly zero (a zero right-hand side, or
        # one already solved) keeps a zero direction: its g, and so its
        # update y, stay zero while the other columns iterate
        V[:, 0, :] = R / torch.where(rn > 0, rn, torch.ones_like(rn))[:, None]
        Rh = np.zeros((nrhs, restart, restart), hdt)           # rotated H, upper triangular
        cs = np.zeros((nrhs, restart))
        sn = np.zeros((nrhs, restart), hdt)
        g = np.zeros((nrhs, restart + 1), hdt)
        g[:, 0] = rn.cpu().numpy()
        j = 0
        for j in range(restart):
            if isinstance(precond, Factor):
                Z[:, j, :].copy_(V[:, j, :])
                rskelf_sv_(precond, Z[:, j, :])
            elif hasattr(precond, "solve_"):             # a captured solve (graph.SolveGraph)
                Z[:, j, :].copy_(V[:, j, :])
                precond.solve_(Z[:, j, :])
            else:
                Z[:, j, :] = op(precond, V[:, j, :])
            w = op(inner_matvec or matvec, Z[:, j, :])
            # (nrhs, j + 2): h, ||w||
            hdev = _cgs2_torch(V, j + 1, w) if cplx else _cgs2(V, j + 1, w)
            hn = hdev[:, -1]
            V[:, j + 1, :] = w / torch.where(hn.real > 0, hn, torch.ones_like(hn))[:, None]
            h = hdev.cpu().numpy()
            iters += 1
            if cplx:
                # complex Givens rotations (zrotg): c real, s complex,
                # [c s; -conj(s) c] [a; b] = [a / |a| r; 0]
                for i in range(j):
                    t = cs[:, i] * h[:, i] + sn[:, i] * h[:, i + 1]
                    h[:, i + 1] = -np.conj(sn[:, i]) * h[:, i] + cs[:, i] * h[:, i + 1]
                    h[:, i] = t
                a, b = h[:, j], h[:, j + 1]
                aa = np.abs(a)
                r = np.sqrt(aa ** 2 + np.abs(b) ** 2)
                ph = np.where(aa > 0, a / np.where(aa > 0, aa, 1.0), 1.0)
                safe = np.where(r > 0, r, 1.0)
                cs[:, j] = np.where(r > 0, aa / safe, 1.0)
                sn[:, j] = np.where(r > 0, ph * np.conj(b) / safe, 0.0)
                Rh[:, : j + 1, j] = h[:, : j + 1]
                Rh[:, j, j] = ph * r
                g[:, j + 1] = -np.conj(sn[:, j]) * g[:, j]
                g[:, j] = cs[:, j] * g[:, j]
                h = np.abs(h)                              # only |h_{j+1}| is read below
            else:
                for i in range(j):                             # earlier rotations
                    t = cs[:, i] * h[:, i] + sn[:, i] * h[:, i + 1]
                    h[:, i + 1] = -sn[:, i] * h[:, i] + cs[:, i] * h[:, i + 1]
                    h[:, i] = t
                r = np.hypot(h[:, j], h[:, j + 1])
                safe = np.where(r > 0, r, 1.0)
                cs[:, j] = np.where(r > 0, h[:, j] / safe, 1.0)
                sn[:, j] = np.where(r > 0, h[:, j + 1] / safe, 0.0)
                Rh[:, : j + 1, j] = h[:, : j + 1]
                Rh[:, j, j] = r
                g[:, j + 1] = -sn[:, j] * g[:, j]
                g[:, j] = cs[:, j] * g[:, j]
            est = np.abs(g[:, j + 1])                          # ||g - H y|| of the LS solution
            if bool((est / inner_ref <= 0.5 * tol).all()) or bool((h[:, j + 1] == 0).all()):
                break
        y = np.zeros((nrhs, j + 1), hdt)                       # back substitution
        for i in range(j, -1, -1):
            d = Rh[:, i, i]
            y[:, i] = np.where(d != 0, (g[:, i] - (Rh[:, i, i + 1: j + 1] * y[:, i + 1:]).sum(1))
                               / np.where(d != 0, d, 1.0), 0.0)
        yd = torch.from_numpy(y).to(B2.device)[:, None, :]     # (nrhs, 1, j+1)
        Xt = Xt + torch.bmm(yd, Z[:, : j + 1, :])[:, 0, :]
    info = dict(iters=iters, outer=outer, relres=relres.tolist())
    X = Xt.t()
    return (X.reshape(-1) if vec else X), info
