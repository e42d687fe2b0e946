"""FP64 CPU restatement of FLAM's rskelf path -- the oracle (test infrastructure).

Each function states which reference file and lines it follows.  Indices are
0-based here; the reference is 1-based MATLAB.  Only tests, the smoke check
and the CPU-baseline leg of the benchmark use this module.

Pinning (what checks each function; see tests/test_oracle_*.py):
  hypoct          README per-level block counts (README.md:99-106), partition /
                  adjacency invariants checked by brute force on small trees.
  interp_decomp   rank = #{|R_jj| > tol |R_11|} of textbook QRCP, the ID error
                  bound against the SVD, max|T| <= Tmax after refinement, the
                  empty / m > 8n special cases.
  rskelf          closed-form inverse of the unit-circle matrix, dense LU at
                  small N, README rank column within a few percent.
  rskelf_sv/_mv   sv(mv(x)) = x, adjoint identities, 't' = conj of 'c',
                  agreement of 'p'/'h' with 'n' on SPD input.
"""
import math
import warnings

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sps


class FlamError(ValueError):
    """Argument error carrying the reference's error identifier."""

    def __init__(self, ident, msg):
        super().__init__(f"{ident}: {msg}")
        self.ident = ident


def chksymm(symm):
    """core/chksymm.m:8-13."""
    s = str(symm).lower()
    if s not in ("n", "s", "h", "p"):
        raise FlamError("FLAM:chksymm:invalidSymm",
                        "Symmetry parameter must be one of 'N', 'S', 'H', or 'P'.")
    return s


def chktrans(trans):
    """core/chktrans.m:8-13."""
    t = str(trans).lower()
    if t not in ("n", "t", "c"):
        raise FlamError("FLAM:chktrans:invalidTrans",
                        "Transpose parameter must be one of 'N', 'T', or 'C'.")
    return t


def mround(v):
    """MATLAB round(): halves go away from zero (numpy rounds them to even)."""
    v = np.asarray(v, dtype=float)
    return np.sign(v) * np.floor(np.abs(v) + 0.5)


# --------------------------------------------------------------------------
# tree
# --------------------------------------------------------------------------
class Tree:
    """Adaptive 2^d-tree as produced by core/hypoct.m.

    nlvl   number of levels (root = level 1 in the reference, index 0 here)
    lvp    box-index offsets per level, length nlvl+1
    l      (d, nlvl) box side lengths per level
    ctr    list of (d,) centres; xi list of point-index arrays (leaves only)
    prnt   parent box (-1 for root); chld list of child lists; nbor list of
           neighbour lists.
    """

    def __init__(self):
        self.nlvl = 0
        self.lvp = []
        self.l = None
        self.ctr = []
        self.xi = []
        self.prnt = []
        self.chld = []
        self.nbor = []

    @property
    def nbox(self):
        return len(self.xi)

    def level_of(self, box):
        """0-based level index of a box (lvp[lv] <= box < lvp[lv+1])."""
        return int(np.searchsorted(self.lvp, box, side="right") - 1)


def hypoct(x, occ, lvlmax=np.inf, ext=None):
    """Build the tree; follows core/hypoct.m:51-171.

    Long sides (>= max side / sqrt 2) are bisected (hypoct.m:81-82); a box
    with more than occ points is split, a point going to the upper child
    in a dimension iff x > ctr strictly (hypoct.m:90-93).  Neighbours are
    siblings, adjacent coarser leaves and adjacent children of the parent's
    neighbours (hypoct.m:141-170).
    """
    if not occ >= 0:
        raise FlamError("FLAM:hypoct:invalidOcc", "Leaf occupancy must be nonnegative.")
    if not lvlmax >= 1:
        raise FlamError("FLAM:hypoct:invalidLvlmax", "Maximum tree depth must be at least 1.")
    x = np.asarray(x, dtype=float)
    d, n = x.shape
    if ext is None or np.size(ext) == 0:
        ext = np.stack([x.min(axis=1), x.max(axis=1)], axis=1)
    ext = np.asarray(ext, dtype=float).reshape(d, 2)
    l = ext[:, 1] - ext[:, 0]
    t = Tree()
    t.ctr = [0.5 * (ext[:, 0] + ext[:, 1])]
    t.xi = [np.arange(n)]
    t.prnt = [-1]
    t.chld = [[]]
    t.nbor = [[]]
    t.lvp = [0, 1]
    lsz = [l.copy()]
    nlvl = 1
    weights = 2 ** np.arange(d)
    while nlvl < lvlmax:
        nbox_before = len(t.xi)
        ldiv = l >= l.max() / math.sqrt(2.0)
        l = l.copy()
        l[ldiv] *= 0.5
        for p in range(t.lvp[nlvl - 1], t.lvp[nlvl]):
            pts = t.xi[p]
            if len(pts) <= occ:
                continue
            c = t.ctr[p]
            code = weights @ ((x[:, pts] > c[:, None]) & ldiv[:, None]).astype(np.int64)
            for u in np.unique(code):
                bits = (u >> np.arange(d)) & 1
                t.ctr.append(c + l * ldiv * (bits - 0.5))
                t.xi.append(pts[code == u])
                t.prnt.append(p)
                t.chld.append([])
                t.nbor.append([])
                t.chld[p].append(len(t.xi) - 1)
            t.xi[p] = np.zeros(0, dtype=np.int64)
        if len(t.xi) <= nbox_before:
            break
        nlvl += 1
        t.lvp.append(len(t.xi))
        lsz.append(l.copy())
    t.nlvl = nlvl
    t.lvp = np.asarray(t.lvp, dtype=np.int64)
    t.l = np.stack(lsz, axis=1)

    for lv in range(1, nlvl):
        lc = t.l[:, lv]
        for i in range(t.lvp[lv], t.lvp[lv + 1]):
            c = t.ctr[i]
            p = t.prnt[i]
            nb = [j for j in t.chld[p] if j != i]
            for j in t.nbor[p]:
                if len(t.xi[j]) == 0:
                    continue
                jl = t.l[:, t.level_of(j)]
                dist = mround((np.abs(c - t.ctr[j]) - 0.5 * (lc + jl)) / lc)
                if dist.max() <= 0:
                    nb.append(j)
            cand = [q for j in t.nbor[p] for q in t.chld[j]]
            if cand:
                cc = np.stack([t.ctr[q] for q in cand], axis=1)
                dist = mround(np.abs(c[:, None] - cc) / lc[:, None]).max(axis=0)
                nb.extend(q for q, dd in zip(cand, dist) if dd <= 1)
            t.nbor[i] = nb
    return t


# --------------------------------------------------------------------------
# interpolative decomposition
# --------------------------------------------------------------------------
def _retriangularize(R, k):
    """Rotate the rows of R so that its first k columns become upper
    triangular.  Stands in for the qrupdate/Givens/Householder steps of the
    reference: T = R11 \\ R12, the column norms of R22 and the row norms of
    inv(R11) are all invariant under such rotations."""
    Q, _ = np.linalg.qr(R[:, :k], mode="complete")
    Rn = Q.conj().T @ R
    Rn[:, :k] = np.triu(Rn[:, :k])
    return Rn


def interp_decomp(A, rank_or_tol, Tmax=2.0, rrqr_iter=np.inf):
    """Column ID  A(:,rd) ~ A(:,sk) T; follows core/id.m:44-248 (no fixed cols).

    Returns (sk, rd, T, niter) with 0-based index arrays.
    """
    if not rank_or_tol >= 0:
        raise FlamError("FLAM:id:invalidRankOrTol", "Rank or tolerance must be nonnegative.")
    if not Tmax >= 1:
        raise FlamError("FLAM:id:invalidTmax", "Interpolation matrix entry bound must be >= 1.")
    if not rrqr_iter >= 0:
        raise FlamError("FLAM:id:invalidRRQRIter",
                        "Maximum number of RRQR iterations must be nonnegative.")
    A = np.asarray(A)
    m, n = A.shape
    niter = 0
    if m == 0 or n == 0:                                   # id.m:64-68
        return (np.zeros(0, dtype=np.int64), np.arange(n),
                np.zeros((0, n), dtype=A.dtype), niter)
    tol = math.fmod(rank_or_tol, 1.0)                      # id.m:71-73
    kmax = int(math.floor(rank_or_tol))
    if kmax == 0 or kmax > n:
        kmax = n
    if m > 8 * n:                                          # id.m:102
        A = sla.qr(A, mode="r")[0][:n, :]
    R, p = sla.qr(A, mode="r", pivoting=True)              # id.m:105
    R = R[: min(R.shape), :]
    tol_abs = abs(R[0, 0]) * tol                           # id.m:106-107
    k = int(np.count_nonzero(np.abs(np.diag(R)) > tol_abs))  # id.m:108-111
    R = R[:k, :].copy()
    k = min(k, kmax)

    def t_of(Rm, kk):
        if kk == 0:
            return np.zeros((0, n), dtype=Rm.dtype)
        return sla.solve_triangular(Rm[:kk, :kk], Rm[:kk, kk:])

    def r2_of(Rm, kk):
        Ri = sla.solve_triangular(Rm[:kk, :kk], np.eye(kk, dtype=Rm.dtype))
        return np.sum(np.abs(Ri) ** 2, axis=1)

    # Gu-Eisenstat refinement, one-loop variant (id.m:116-236).  R keeps
    # R12 in R-space; T is formed on demand.
    if math.isfinite(Tmax) and rrqr_iter > 0 and 0 < k < n:
        f2 = Tmax ** 2
        conv = False
        while niter < rrqr_iter:
            T = t_of(R, k)
            c2 = np.sum(np.abs(R[k:, k:]) ** 2, axis=0)
            r2 = r2_of(R, k)
            score = np.abs(T) ** 2 + r2[:, None] * c2[None, :]
            flat = score.ravel(order="F")
            idx = int(np.argmax(flat))
            if flat[idx] <= f2:
                conv = True
                break
            niter += 1
            i, j = idx % k, idx // k
            q = np.arange(n)
            if i < k - 1:
                q[[i, k - 1]] = q[[k - 1, i]]
            if j > 0:
                q[[k, k + j]] = q[[k + j, k]]
            q[[k - 1, k]] = q[[k, k - 1]]
            p = p[q]
            R = _retriangularize(R[:, q], k)
            # rank reduction if a skeleton is within tol of the others' span
            rr = 1.0 / np.sqrt(r2_of(R, k))
            i = int(np.argmin(rr))
            if rr[i] > tol_abs:
                continue
            q = np.arange(n)
            if i < k - 1:
                q[[i, k - 1]] = q[[k - 1, i]]
            p = p[q]
            R = _retriangularize(R[:, q], k)
            k -= 1
            R = R[:k, :].copy()
            if k == 0:
                break
        if not conv:
            warnings.warn("FLAM:id:maxIterCount: Maximum number of RRQR iterations reached.")
    T = t_of(R, k)
    return p[:k].astype(np.int64), p[k:].astype(np.int64), T, niter


# --------------------------------------------------------------------------
# factorization
# --------------------------------------------------------------------------
def _defaults(opts):
    o = dict(lvlmax=np.inf, ext=None, Tmax=2.0, rrqr_iter=np.inf, symm="n",
             stop=np.inf, verb=0, order="sequential")
    if opts:
        o.update(opts)
    o["symm"] = chksymm(o["symm"])
    if o["order"] not in ("sequential", "level"):
        raise ValueError("order must be 'sequential' or 'level'")
    st = o["stop"]
    o["stopf"] = st if callable(st) else (lambda lvl, l, s=st: lvl >= s)
    return o


def rskelf(Afun, x, occ, rank_or_tol, pxyfun=None, opts=None):
    """Recursive skeletonization factorization; follows rskelf/rskelf.m:109-309.

    opts['order'] = 'sequential' reproduces the reference's in-level box
    order dependence (a box sees already-compressed neighbours' skeletons,
    rskelf.m:186 vs :260); 'level' snapshots every neighbour's point list at
    the start of the level, the convention of a level-parallel device
    implementation.

    Returns a dict with keys N, nlvl, lvp, factors (list of dicts sk, rd, T,
    L, U, p, E, F), symm, stats (per-level rows: lvl, nblk, nrem1, nrem2)
    and, for partial factorizations, Si and S.
    """
    o = _defaults(opts)
    symm = o["symm"]
    x = np.asarray(x, dtype=float)
    N = x.shape[1]
    t = hypoct(x, occ, o["lvlmax"], o["ext"])
    nbox = t.nbox
    nonempty = np.array([len(v) > 0 for v in t.xi])
    pblk = np.zeros(t.nlvl + 1, dtype=np.int64)
    for lv in range(t.nlvl):
        pblk[lv + 1] = pblk[lv] + int(nonempty[t.lvp[lv]:t.lvp[lv + 1]].sum())

    dtype = np.result_type(Afun(np.arange(1), np.arange(1)).dtype, np.float64)
    xi = [v.copy() for v in t.xi]
    M = [None] * nbox
    Mi = [None] * nbox
    rem = np.ones(N, dtype=bool)
    P = np.zeros(N, dtype=np.int64)
    factors = []
    lvp = [0]
    stats = []
    for lvl in range(t.nlvl, 0, -1):
        li = lvl - 1
        l = t.l[:, li]
        if o["stopf"](t.nlvl - lvl, l):
            break
        nrem1 = int(rem.sum())
        boxes = range(t.lvp[li], t.lvp[li + 1])
        for i in boxes:                                     # rskelf.m:179-181
            xi[i] = np.concatenate([xi[i]] + [xi[c] for c in t.chld[i]]).astype(np.int64)
        view = list(xi) if o["order"] == "level" else xi
        for i in boxes:
            slf = xi[i]
            nbr = (np.concatenate([view[j] for j in t.nbor[i]]).astype(np.int64)
                   if t.nbor[i] else np.zeros(0, dtype=np.int64))
            ns = len(slf)
            Mloc = np.zeros((ns, ns), dtype=dtype)         # rskelf.m:190-199
            Mi[i] = slf
            if lvl < t.nlvl:
                P[slf] = np.arange(ns)
                for c in t.chld[i]:
                    kk = P[xi[c]]
                    Mloc[np.ix_(kk, kk)] = M[c]
                    M[c] = None
                    Mi[c] = None
            Kpxy = np.zeros((0, ns), dtype=dtype)
            if lvl > 2:                                     # rskelf.m:202-206
                if pxyfun is None:
                    nbr = np.setdiff1d(np.flatnonzero(rem), slf)
                else:
                    Kpxy, nbr = pxyfun(x, slf, nbr, l, t.ctr[i])
            K = Afun(nbr, slf)                              # rskelf.m:209-212
            if symm == "n":
                K = np.vstack([K, Afun(slf, nbr).conj().T])
            K = np.vstack([K, Kpxy])
            sk, rd, T, _ = interp_decomp(K, rank_or_tol, o["Tmax"], o["rrqr_iter"])
            if len(rd) == 0:                                # rskelf.m:216
                M[i] = Mloc
                continue
            K = Afun(slf, slf) + Mloc                       # rskelf.m:219-223
            if symm == "s":
                K[rd, :] -= T.T @ K[sk, :]
            else:
                K[rd, :] -= T.conj().T @ K[sk, :]
            K[:, rd] -= K[:, sk] @ T
            Krr = K[np.ix_(rd, rd)]
            Ksr = K[np.ix_(sk, rd)]
            if symm in ("n", "s"):                          # rskelf.m:224-227
                pp, L, U = sla.lu(Krr, p_indices=True)
                perm = np.argsort(pp)                       # K(rd(perm),rd) = L U
                E = sla.solve_triangular(U, Ksr.T, trans="T").T if len(sk) else Ksr.copy()
                G = sla.solve_triangular(L, K[np.ix_(rd[perm], sk)], lower=True,
                                         unit_diagonal=True)
                X = E @ G
            elif symm == "h":                               # rskelf.m:228-233
                lu_, D, perm = sla.ldl(Krr, lower=True, hermitian=True)
                L = lu_[perm]
                rd = rd[perm]
                T = T[:, perm]
                Ksr = K[np.ix_(sk, rd)]
                U = D
                E = sla.solve_triangular(L, Ksr.T.conj(), lower=True, unit_diagonal=True).conj().T
                E = np.linalg.solve(U, E.T).T
                perm = None
                G = None
                X = E @ (U @ E.conj().T)
            else:                                           # rskelf.m:234-237
                L = sla.cholesky(Krr, lower=True)
                E = sla.solve_triangular(L, Ksr.conj().T, lower=True).conj().T
                U = None
                perm = None
                G = None
                X = E @ E.conj().T
            M[i] = Mloc[np.ix_(sk, sk)] - X                 # rskelf.m:240-246
            Mi[i] = slf[sk]
            factors.append(dict(sk=slf[sk], rd=slf[rd], T=T, L=L, U=U, p=perm, E=E, F=G))
            xi[i] = slf[sk]                                 # rskelf.m:259-261
            rem[slf[rd]] = False
        lvp.append(len(factors))
        nblk = int(pblk[li] + t.lvp[li + 1] - t.lvp[li])
        stats.append(dict(lvl=lvl, nblk=nblk, nrem1=nrem1, nrem2=int(rem.sum())))
    F = dict(N=N, nlvl=t.nlvl, lvp=np.asarray(lvp, dtype=np.int64), factors=factors,
             symm=symm, stats=stats, tree=t)
    Sn = int(rem.sum())
    if Sn:                                                  # rskelf.m:279-308
        Si = np.flatnonzero(rem)
        P[Si] = np.arange(Sn)
        I, J, V = [], [], []
        for i in range(nbox):
            if M[i] is None:
                continue
            ii = P[Mi[i]]
            gi, gj = np.meshgrid(ii, ii, indexing="ij")
            I.append(gi.ravel(order="F"))
            J.append(gj.ravel(order="F"))
            V.append(M[i].ravel(order="F"))
        if I:
            F["S"] = sps.csc_matrix((np.concatenate(V), (np.concatenate(I), np.concatenate(J))),
                                    shape=(Sn, Sn))
        else:
            F["S"] = sps.csc_matrix((Sn, Sn), dtype=dtype)
        F["Si"] = Si
    return F


# --------------------------------------------------------------------------
# sweeps (rskelf/sv/*.m, rskelf/mv/*.m)
# --------------------------------------------------------------------------
def _lsolve(L, B, unit=False):
    return sla.solve_triangular(L, B, lower=True, unit_diagonal=unit)


def _usolve(U, B):
    return sla.solve_triangular(U, B, lower=False)


def _sv_up(F, X, symm, trans):
    for f in F["factors"][: F["lvp"][-1]]:
        sk, rd = f["sk"], f["rd"]
        T = f["T"]
        if symm == "p" or symm == "h":                      # sv_p.m:9-15, sv_h.m:9-16
            X[rd] -= T.conj().T @ X[sk]
            X[rd] = _lsolve(f["L"], X[rd], unit=(symm == "h"))
            X[sk] -= f["E"] @ X[rd]
            if symm == "h":
                X[rd] = np.linalg.solve(f["U"], X[rd])
        elif trans == "n":                                  # sv_nn.m:9-15, sv_sn.m
            X[rd] -= (T.T if symm == "s" else T.conj().T) @ X[sk]
            X[rd] = _lsolve(f["L"], X[rd[f["p"]]], unit=True)
            X[sk] -= f["E"] @ X[rd]
        else:                                               # sv_nc.m:9-15, sv_sc.m
            X[rd] -= T.conj().T @ X[sk]
            X[rd] = sla.solve_triangular(f["U"], X[rd], trans="C")
            X[sk] -= f["F"].conj().T @ X[rd]
    return X


def _sv_down(F, X, symm, trans):
    for f in reversed(F["factors"][: F["lvp"][-1]]):
        sk, rd = f["sk"], f["rd"]
        T = f["T"]
        if symm == "p" or symm == "h":                      # sv_p.m:19-27
            X[rd] -= f["E"].conj().T @ X[sk]
            X[rd] = sla.solve_triangular(f["L"], X[rd], lower=True, trans="C",
                                         unit_diagonal=(symm == "h"))
            X[sk] -= T @ X[rd]
        elif trans == "n":                                  # sv_nn.m:19-27
            X[rd] -= f["F"] @ X[sk]
            X[rd] = _usolve(f["U"], X[rd])
            X[sk] -= T @ X[rd]
        else:                                               # sv_nc.m:19-27, sv_sc.m
            X[rd] -= f["E"].conj().T @ X[sk]
            X[rd[f["p"]]] = sla.solve_triangular(f["L"], X[rd], lower=True, trans="C",
                                                 unit_diagonal=True)
            X[sk] -= (T.conj() if symm == "s" else T) @ X[rd]
    return X


def _mv_up(F, X, symm, trans):
    for f in F["factors"][: F["lvp"][-1]]:
        sk, rd = f["sk"], f["rd"]
        T = f["T"]
        if symm == "p" or symm == "h":                      # mv_p.m:9-15, mv_h.m
            X[sk] += T @ X[rd]
            X[rd] = _tri_mul(f["L"], X[rd], lower=True, conjt=True, unit=(symm == "h"))
            X[rd] += f["E"].conj().T @ X[sk]
            if symm == "h":
                X[rd] = f["U"] @ X[rd]
        elif trans == "n":                                  # mv_nn.m:9-15
            X[sk] += T @ X[rd]
            X[rd] = np.triu(f["U"]) @ X[rd]
            X[rd] += f["F"] @ X[sk]
        else:                                               # mv_nc.m:9-15, mv_sc.m
            X[sk] += (T.conj() if symm == "s" else T) @ X[rd]
            X[rd] = _tri_mul(f["L"], X[rd[f["p"]]], lower=True, conjt=True, unit=True)
            X[rd] += f["E"].conj().T @ X[sk]
    return X


def _mv_down(F, X, symm, trans):
    for f in reversed(F["factors"][: F["lvp"][-1]]):
        sk, rd = f["sk"], f["rd"]
        T = f["T"]
        if symm == "p" or symm == "h":                      # mv_p.m:19-27
            X[sk] += f["E"] @ X[rd]
            X[rd] = _tri_mul(f["L"], X[rd], lower=True, conjt=False, unit=(symm == "h"))
            X[rd] += T.conj().T @ X[sk]
        elif trans == "n":                                  # mv_nn.m:19-27, mv_sn.m
            X[sk] += f["E"] @ X[rd]
            X[rd[f["p"]]] = _tri_mul(f["L"], X[rd], lower=True, conjt=False, unit=True)
            X[rd] += (T.T if symm == "s" else T.conj().T) @ X[sk]
        else:                                               # mv_nc.m:19-27
            X[sk] += f["F"].conj().T @ X[rd]
            X[rd] = np.triu(f["U"]).conj().T @ X[rd]
            X[rd] += T.conj().T @ X[sk]
    return X


def _tri_mul(L, B, lower, conjt, unit):
    Lt = np.tril(L) if lower else np.triu(L)
    if unit:
        Lt = Lt - np.diag(np.diag(Lt)) + np.eye(Lt.shape[0], dtype=Lt.dtype)
    if conjt:
        Lt = Lt.conj().T
    return Lt @ B


def _prep(F, X):
    X = np.asarray(X)
    vec = X.ndim == 1
    dt = np.result_type(X.dtype, np.float64, *(f["E"].dtype for f in F["factors"][:1]))
    Xw = np.array(X, dtype=dt, copy=True)
    if vec:
        Xw = Xw[:, None]
    return Xw, vec


def rskelf_sv(F, X, trans="n", mode=3):
    """Apply F^{-1}; follows rskelf/rskelf_sv.m:13-36 and rskelf/sv/*.m."""
    trans = chktrans(trans)
    if trans == "t":                                        # rskelf_sv.m:22
        return np.conj(rskelf_sv(F, np.conj(X), "c", mode))
    Xw, vec = _prep(F, X)
    symm = F["symm"]
    if mode & 1:
        Xw = _sv_up(F, Xw, symm, trans)
    if mode & 2:
        Xw = _sv_down(F, Xw, symm, trans)
    return Xw[:, 0] if vec else Xw


def rskelf_mv(F, X, trans="n", mode=3):
    """Apply F; follows rskelf/rskelf_mv.m:13-36 and rskelf/mv/*.m."""
    trans = chktrans(trans)
    if trans == "t":
        return np.conj(rskelf_mv(F, np.conj(X), "c", mode))
    Xw, vec = _prep(F, X)
    symm = F["symm"]
    if mode & 1:
        Xw = _mv_up(F, Xw, symm, trans)
    if mode & 2:
        Xw = _mv_down(F, Xw, symm, trans)
    return Xw[:, 0] if vec else Xw


def _logdet_ldl(D):
    """log det of the block-diagonal D of an L D L' factorization; follows
    core/logdet_ldl.m: 1 x 1 blocks by their diagonal, 2 x 2 blocks by
    their determinant (complex logarithm)."""
    D = np.asarray(D)
    n = D.shape[0]
    off = np.zeros(n, dtype=bool)
    off[:-1] |= D[np.arange(n - 1), np.arange(1, n)] != 0
    off[1:] |= D[np.arange(1, n), np.arange(n - 1)] != 0
    ld = np.sum(np.log(np.diag(D)[~off].astype(complex)))
    idx = np.flatnonzero(off)
    for i in range(0, len(idx), 2):
        ld += np.log(complex(np.linalg.det(D[np.ix_(idx[i:i + 2], idx[i:i + 2])])))
    return ld


def _detperm(p):
    """Sign of a permutation vector (compat/detperm.m): (-1)^(n - cycles)."""
    p = np.asarray(p)
    seen = np.zeros(len(p), dtype=bool)
    cycles = 0
    for i in range(len(p)):
        if not seen[i]:
            cycles += 1
            j = i
            while not seen[j]:
                seen[j] = True
                j = p[j]
    return -1.0 if (len(p) - cycles) % 2 else 1.0


def rskelf_logdet(F):
    """log det of the factored matrix with 0 <= imag < 2 pi; follows
    rskelf/rskelf_logdet.m:11-30."""
    ld = 0j
    for f in F["factors"]:
        if F["symm"] == "p":
            ld += 2 * np.sum(np.log(np.diag(f["L"]).astype(complex)))
        elif F["symm"] == "h":
            ld += _logdet_ldl(f["U"])
        else:
            ld += np.sum(np.log(np.diag(f["U"]).astype(complex))) + np.log(complex(_detperm(f["p"])))
    return complex(ld.real, math.fmod(ld.imag, 2 * math.pi) % (2 * math.pi))


def rskelf_cholmv(F, X, trans="n"):
    """Apply the generalized Cholesky factor C of F = C C' ('p' only);
    follows rskelf/rskelf_cholmv.m:15-52: trans 'n' is the downward half of
    rskelf_mv, 'c' the upward half."""
    if F["symm"] != "p":
        raise FlamError("FLAM:rskelf_cholmv:invalidSymm", "Symmetry parameter must be 'P'.")
    trans = chktrans(trans)
    if trans == "t":
        return np.conj(rskelf_cholmv(F, np.conj(X), "c"))
    return rskelf_mv(F, X, "n", 2 if trans == "n" else 1)


def rskelf_cholsv(F, X, trans="n"):
    """Apply C^{-1} ('p' only); follows rskelf/rskelf_cholsv.m:15-52: trans
    'n' is the upward half of rskelf_sv, 'c' the downward half."""
    if F["symm"] != "p":
        raise FlamError("FLAM:rskelf_cholsv:invalidSymm", "Symmetry parameter must be 'P'.")
    trans = chktrans(trans)
    if trans == "t":
        return np.conj(rskelf_cholsv(F, np.conj(X), "c"))
    return rskelf_sv(F, X, "n", 1 if trans == "n" else 2)


# --------------------------------------------------------------------------
# harness
# --------------------------------------------------------------------------
def snorm(n, mv, mva=None, tol=1e-2, herm=False, niter_max=32, rng=None):
    """Power-method spectral norm estimate; follows core/snorm.m:31-72
    (seeded here; the reference draws from an unseeded rand)."""
    rng = np.random.default_rng(0) if rng is None else rng
    x = rng.random(n)
    xnorm = np.linalg.norm(x)
    s = xnorm
    for _ in range(niter_max):
        x = x / xnorm
        s_prev = s
        x = mv(x) if herm else mva(mv(x))
        xnorm = np.linalg.norm(x)
        s = xnorm if herm else math.sqrt(xnorm)
        if xnorm <= np.finfo(float).eps:
            return s
        if abs(s - s_prev) <= s_prev * tol:
            return s
    warnings.warn("FLAM:snorm:maxIterCount: Maximum number of iterations reached.")
    return s


def rank_table(F):
    """The reference's verbose per-level table (rskelf.m:264-272) as rows
    (lvl, nblk, start npts, end npts)."""
    return [(s["lvl"], s["nblk"], s["nrem1"], s["nrem2"]) for s in F["stats"]]


# --------------------------------------------------------------------------
# refinement (the CPU leg of the benchmark)
# --------------------------------------------------------------------------
def fgmres(matvec, B, precond, tol=1e-10, restart=16, maxit=8):
    """Flexible GMRES with the factor as right preconditioner, all columns of
    B at once, stopped on every column's true relative residual -- the
    refinement the device runs (the reference preconditions MATLAB's gmres
    with rskelf_sv, rskelf/test/ie_square3.m:97): Arnoldi with classical
    Gram-Schmidt twice, Givens rotations on the Hessenberg columns, the
    preconditioned directions kept.  Real data.  Returns (X, iters, relres)."""
    B = np.asarray(B, dtype=float)
    vec = B.ndim == 1
    B2 = B.reshape(B.shape[0], -1)
    N, nrhs = B2.shape
    bn = np.linalg.norm(B2, axis=0)
    bn = np.where(bn > 0, bn, 1.0)
    X = np.zeros_like(B2)
    iters = 0
    rel = None
    for outer in range(maxit + 1):
        R = B2 - matvec(X) if outer else B2.copy()
        rn = np.linalg.norm(R, axis=0)
        rel = rn / bn
        if np.all(rel <= tol) or outer == maxit:
            break
        V = np.zeros((restart + 1, N, nrhs))
        Z = np.zeros((restart, N, nrhs))
        V[0] = R / np.where(rn > 0, rn, 1.0)
        H = np.zeros((restart + 1, restart, nrhs))
        cs = np.zeros((restart, nrhs))
        sn = np.zeros((restart, nrhs))
        g = np.zeros((restart + 1, nrhs))
        g[0] = rn
        j = 0
        for j in range(restart):
            Z[j] = precond(V[j])
            w = matvec(Z[j])
            h = np.zeros((j + 2, nrhs))
            for _ in range(2):                          # classical Gram-Schmidt, twice
                c = np.einsum("inr,nr->ir", V[: j + 1], w)
                w = w - np.einsum("inr,ir->nr", V[: j + 1], c)
                h[: j + 1] += c
            h[j + 1] = np.linalg.norm(w, axis=0)
            V[j + 1] = w / np.where(h[j + 1] > 0, h[j + 1], 1.0)
            iters += 1
            for i in range(j):
                t = cs[i] * h[i] + sn[i] * h[i + 1]
                h[i + 1] = -sn[i] * h[i] + cs[i] * h[i + 1]
                h[i] = t
            r = np.hypot(h[j], h[j + 1])
            safe = np.where(r > 0, r, 1.0)
            cs[j] = np.where(r > 0, h[j] / safe, 1.0)
            sn[j] = np.where(r > 0, h[j + 1] / safe, 0.0)
            H[: j + 1, j] = h[: j + 1]
            H[j, j] = r
            g[j + 1] = -sn[j] * g[j]
            g[j] = cs[j] * g[j]
            if np.all(np.abs(g[j + 1]) / bn <= 0.5 * tol):
                break
        y = np.zeros((j + 1, nrhs))
        for i in range(j, -1, -1):
            d = H[i, i]
            y[i] = np.where(d != 0, (g[i] - np.einsum("kr,kr->r", H[i, i + 1: j + 1], y[i + 1:]))
                            / np.where(d != 0, d, 1.0), 0.0)
        X = X + np.einsum("jnr,jr->nr", Z[: j + 1], y)
    return (X[:, 0] if vec else X), iters, rel
