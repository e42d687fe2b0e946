"""Krylov refinement with an rskelf preconditioner.

The reference solves with rskelf as a preconditioner inside MATLAB's gmres
(rskelf/test/ie_square3.m:97: gmres(mv,B,32,1e-12,32,@(x)rskelf_sv(F,x))).
Here the preconditioner is usually a low-precision (FP32) factorization and
the operator an exact FP64 product, so this is flexible GMRES (the
preconditioned directions are kept, so a preconditioner that is linear only
up to FP32 rounding does not limit the attainable FP64 residual), run on
all right-hand sides at once and stopped on the true relative residual
||b - A x|| / ||b|| of every column.
"""
import torch


def fgmres(matvec, B, precond, tol=1e-10, restart=16, maxit=8):
    """Solve A X = B.  matvec, precond: callables on (N, nrhs) FP64 CUDA
    tensors.  Returns (X, info) with info = dict(iters, outer, relres)."""
    vec = B.dim() == 1
    B2 = B.reshape(B.shape[0], -1).to(torch.float64)
    N, nrhs = B2.shape
    bnorm = torch.linalg.vector_norm(B2, dim=0)
    bnorm = torch.where(bnorm > 0, bnorm, torch.ones_like(bnorm))
    X = torch.zeros_like(B2)
    iters = 0
    relres = None
    outer = 0
    for outer in range(maxit + 1):
        R = B2 - matvec(X) if outer else B2.clone()
        rn = torch.linalg.vector_norm(R, dim=0)
        relres = (rn / bnorm).cpu()
        if bool((relres <= tol).all()) or outer == maxit:
            break
        V = [R / rn]
        Z = []
        H = torch.zeros(nrhs, restart + 1, restart, dtype=torch.float64)
        g = torch.zeros(nrhs, restart + 1, dtype=torch.float64)
        g[:, 0] = rn.cpu()
        y = None
        for j in range(restart):
            z = precond(V[j])
            Z.append(z)
            w = matvec(z)
            hcol = torch.zeros(nrhs, j + 2, dtype=torch.float64, device=B2.device)
            for i in range(j + 1):          # modified Gram-Schmidt, per column
                h = (V[i] * w).sum(0)
                hcol[:, i] = h
                w = w - V[i] * h
            hn = torch.linalg.vector_norm(w, dim=0)
            hcol[:, j + 1] = hn
            H[:, : j + 2, j] = hcol.cpu()
            V.append(w / torch.where(hn > 0, hn, torch.ones_like(hn)))
            iters += 1
            Hj = H[:, : j + 2, : j + 1]
            y = torch.linalg.lstsq(Hj, g[:, : j + 2, None]).solution[..., 0]
            est = torch.linalg.vector_norm(g[:, : j + 2] - (Hj @ y[..., None])[..., 0], dim=1)
            if bool((est / bnorm.cpu() <= 0.5 * tol).all()) or bool((hn == 0).all()):
                break
        Zm = torch.stack(Z)                 # (j+1, N, nrhs)
        yd = y.to(B2.device).t()            # (j+1, nrhs)
        X = X + (Zm * yd[:, None, :]).sum(0)
    info = dict(iters=iters, outer=outer, relres=relres.tolist())
    return (X.reshape(-1) if vec else X), info
