"""Gauss-Newton reconstruction on the GPU (drop-in for redbirdpy.recon on the hot path).

``runrecon`` keeps the reference's control flow (recon.py:39-335): sync
properties, forward-solve every wavelength, build Jacobians, form the
normal equations and update the unknowns.  The dense algebra stays on the
device:

* the Jacobian of every wavelength is written node-major into one block
  JT (nn x nwv*npairs) by the Jacobian kernel;
* for chromophore unknowns the flattened Jacobian [E[w,c] J_w] is never
  built: its Gram is (E E^T)[w,w'] * (J_w J_w'^T), one Gram over nn;
* the regularised system s^2 J J^T + lambda I is diagonally normalised
  (recon.py:784-789) and solved by a float64 Cholesky; with
  ``precision='mixed'`` the Gram is a bf16x3 MFMA product of the float64
  Jacobian and the solution is refined against float64 products J (J^T y),
  so the step is that of the float64 Jacobian (a float32 Jacobian would
  perturb the reconstructed parameters by ~1e-6 relative).
"""

__all__ = ["runrecon", "reginv", "reginvover", "reginvunder", "matreform", "matflat", "syncprop", "prior"]

import time
from collections.abc import Mapping

import numpy as np
import torch

from . import _gpu
from .forward import forward_device, _active_pairs
from .utility import sdmap, meshinterp
from .property import updateprop, extinction, CHROMOPHORES


# ---------------------------------------------------------------------------
# small host helpers with the reference's semantics
# ---------------------------------------------------------------------------
def _normalize_diag(A):
    """Scale to unit diagonal: (D A D, d) with d = 1/sqrt(|diag|+1e-16) (recon.py:784-789)."""
    if torch.is_tensor(A):
        di = 1.0 / torch.sqrt(torch.abs(torch.diagonal(A)) + 1e-16)
        return di[:, None] * A * di[None, :], di
    di = 1.0 / np.sqrt(np.abs(np.diag(A)) + 1e-16)
    return (di[:, None] * di[None, :]) * A, di


def _flatten_detphi(detphi, sd, wavelengths, rfcw):
    """Measurements flattened wavelength by wavelength (recon.py:792-815)."""
    if isinstance(detphi, np.ndarray):
        return detphi.flatten()
    out = []
    for wv in wavelengths:
        p = detphi.get(wv, detphi) if isinstance(detphi, dict) else detphi
        if isinstance(p, dict):
            for md in rfcw:
                out.extend(p.get(md, {}).get("detphi", p).flatten())
        else:
            out.extend(np.asarray(p).flatten())
    return np.array(out)


def matreform(Amat, ymeas, ymodel, form="complex"):
    """Real-valued forms of the complex system (recon.py:512-567)."""
    rhs = ymeas - ymodel
    if form == "complex":
        return Amat, rhs, 1
    if form in ("real", "reim"):
        if not np.isreal(rhs).all() and not np.isreal(Amat).all():
            if form == "reim":
                A = np.block([[np.real(Amat), -np.imag(Amat)], [np.imag(Amat), np.real(Amat)]])
            else:
                A = np.vstack([np.real(Amat), np.imag(Amat)])
            return A, np.concatenate([np.real(rhs), np.imag(rhs)]), 2
        return np.real(Amat), np.real(rhs), 1
    if form == "logphase":
        t = np.conj(ymodel) / np.abs(ymodel * ymodel)
        t = t[:, None] * Amat if Amat.ndim == 2 else t * Amat
        lr = np.log(np.abs(ymeas)) - np.log(np.abs(ymodel))
        if np.isreal(ymodel).all():
            return np.real(t), lr, 1
        return np.vstack([np.real(t), np.imag(t)]), np.concatenate([lr, np.angle(ymeas) - np.angle(ymodel)]), 2
    raise ValueError("Unknown form: %s" % form)


def matflat(Amat, weight=None):
    """Flatten a dict of Jacobian blocks (recon.py:570-593)."""
    if isinstance(Amat, np.ndarray):
        return Amat
    if isinstance(Amat, dict):
        keys = list(Amat.keys())
        w = np.ones(len(keys)) if weight is None else weight
        first = Amat[keys[0]]
        if isinstance(first, dict):
            inner = list(first.keys())
            return np.vstack([np.hstack([Amat[k][wv] * w[j] for j, k in enumerate(keys)]) for wv in inner])
        return np.hstack([Amat[k] * w[i] for i, k in enumerate(keys)])
    return Amat


def prior(seg, priortype, params=None):
    """Structure-prior regularisation matrix (recon.py:596-666).

    1-D labels: 'laplace' gives L = I with -1/|g| between distinct members of
    each label group g of more than one member; 'helmholtz' uses
    -1/(|g| + beta).  A 2-D composition (n x nc) with 'comp' links members
    whose compositions differ by dval = sum |s_i - s_j| < alpha nc with
    -alpha - dval/nc, normalised by beta sqrt(|rowsum_i| |rowsum_j|), plus I
    (sparse).  Anything else gives None.
    """
    if not priortype:
        return None
    params = params or {}
    seg = np.asarray(seg)
    if seg.ndim == 1:
        if priortype not in ("laplace", "helmholtz"):
            return None
        _, inv = np.unique(seg, return_inverse=True)
        c = np.bincount(inv)[inv].astype(float)  # size of each member's group
        extra = 0.0 if priortype == "laplace" else params.get("beta", 1.0)
        same = inv[:, None] == inv[None, :]
        L = np.where(same & (c[:, None] > 1), -1.0 / (c[:, None] + extra), 0.0)
        np.fill_diagonal(L, 1.0)
        return L
    if priortype == "comp" and seg.ndim == 2:
        from scipy import sparse as sp

        alpha = params.get("alpha", 0.1)
        beta = params.get("beta", 1.0)
        n, nc = seg.shape
        dval = np.abs(seg[:, None, :] - seg[None, :, :]).sum(axis=2)
        link = (dval < alpha * nc) & ~np.eye(n, dtype=bool)
        W = np.where(link, -alpha - dval / nc, 0.0)
        rs = np.abs(W.sum(axis=1))
        W = np.where(link, W / (beta * np.sqrt(rs[:, None] * rs[None, :] + 1e-16)), 0.0)
        return (sp.csr_matrix(W) + sp.eye(n)).tocsr()
    return None


def _dual(recon):
    return "mapid" in recon and "mapweight" in recon and "elem" in recon


def syncprop(cfg, recon):
    """Recon unknowns onto the forward mesh (recon.py:669-781).

    Label tables (< 50 rows) and same-mesh arrays are copied; on a dual mesh
    (recon['mapid'] / ['mapweight'] / ['elem']) node arrays are interpolated
    to the forward nodes with meshinterp.
    """
    if "param" in recon:
        keys = list(recon["param"].keys())
        first = recon["param"][keys[0]]
        n = len(first) if hasattr(first, "__len__") else 1
        if n < 50:
            cfg["param"] = {k: (v.copy() if hasattr(v, "copy") else v) for k, v in recon["param"].items()}
        else:
            cfg.setdefault("param", {})
            for k in keys:
                cfg["param"][k] = (meshinterp(recon["param"][k], recon["mapid"], recon["mapweight"], recon["elem"])
                                   if _dual(recon) else recon["param"][k].copy())
    elif "prop" in recon:
        if isinstance(recon["prop"], dict):
            first = next(iter(recon["prop"].values()))
            if first.shape[0] >= 50 and _dual(recon):
                cfg["prop"] = {k: meshinterp(v, recon["mapid"], recon["mapweight"], recon["elem"])
                               for k, v in recon["prop"].items()}
            else:
                cfg["prop"] = {k: v.copy() for k, v in recon["prop"].items()}
        elif recon["prop"].shape[0] >= 50 and _dual(recon):
            cfg["prop"] = meshinterp(recon["prop"], recon["mapid"], recon["mapweight"], recon["elem"])
        else:
            cfg["prop"] = recon["prop"].copy()
    return cfg, recon


# ---------------------------------------------------------------------------
# regularised solves on the GPU (recon.py:338-509)
# ---------------------------------------------------------------------------
def _chol_solve_normalized(H, rhs):
    """y = H^-1 rhs via the unit-diagonal normalisation and a float64 Cholesky."""
    Hn, di = _normalize_diag(H)
    L = _gpu.cholesky(Hn)
    z = torch.cholesky_solve((di * rhs).reshape(-1, 1), L).reshape(-1)
    return di * z


def _dense_inputs(Amat, rhs):
    """Device copies of a dense A and rhs (float64, or complex128 for complex input)."""
    dev = _gpu.device()
    cplx = np.iscomplexobj(Amat) or np.iscomplexobj(rhs)
    dt = np.complex128 if cplx else np.float64
    A = torch.as_tensor(np.ascontiguousarray(Amat, dtype=dt), device=dev)
    b = torch.as_tensor(np.ascontiguousarray(np.asarray(rhs, dtype=dt).ravel()), device=dev)
    return A, b, cplx


def _drop_zero(A, b):
    """Columns and rows of A that are all zero (recon.py:375-390, :441-455): (A', b', kept columns)."""
    n0 = A.shape[1]
    keep = torch.nonzero(torch.sum(torch.abs(A), 0) != 0).reshape(-1)
    if keep.numel() < n0:
        A = A[:, keep]
    rows = torch.sum(torch.abs(A), 1) != 0
    if int(rows.sum()) < A.shape[0]:
        A, b = A[rows], b[rows]
    return A, b, keep, n0


def _solve_sym(H, g, cplx):
    """H^-1 g in the unit-diagonal form (recon.py:784-789): Cholesky for real SPD H, an LU solve for
    the complex symmetric H of complex data."""
    if not cplx:
        return _chol_solve_normalized(H, g)
    Hn, di = _normalize_diag(H)
    return di * torch.linalg.solve(Hn, (di * g).reshape(-1, 1)).reshape(-1)


def _spans(ncols, nx, blocks):
    """Column ranges an R^-1 (or L^T L) of size nx applies to: all columns, or the unknown blocks of
    that size (recon.py:458-470)."""
    if nx == ncols:
        return [(0, nx)]
    if blocks is None:
        return []
    cum = np.cumsum([0] + [int(v[1]) for v in blocks.values()])
    return [(int(cum[i]), int(cum[i + 1])) for i in range(len(cum) - 1) if cum[i + 1] - cum[i] == nx]


def reginvunder(Amat, rhs, lambda_, invR=None, blocks=None, **kwargs):
    """x = R^-1 A'^T (A' A'^T + lambda I)^-1 rhs with A' = A R^-1 (recon.py:429-509).

    invR (the inverse of the prior's R factor) transforms all columns when
    its size matches, else each unknown block of its size listed in
    ``blocks`` ({name: (rows, cols)}); without it, x = A^T (A A^T + lambda I)^-1 rhs.
    Zero columns and rows are dropped first, as in the reference.
    """
    A, b, cplx = _dense_inputs(Amat, rhs)
    A, b, keep, n0 = _drop_zero(A, b)
    spans = []
    if invR is not None:
        R = torch.as_tensor(np.asarray(invR, dtype=np.float64), device=A.device).to(A.dtype)
        spans = _spans(A.shape[1], R.shape[0], blocks)
        A = A.clone()
        for a, e in spans:
            A[:, a:e] = A[:, a:e] @ R
    if cplx:
        H = A @ A.t()  # unconjugated, as the reference's Amat @ Amat.T
        H.diagonal().add_(lambda_)
        res = A.t() @ _solve_sym(H, b, True)
    else:
        AT = A.t().contiguous()  # (N x m) node-major, the Gram kernel's layout
        H = _gpu.gram(AT)
        H.diagonal().add_(lambda_)
        y = _chol_solve_normalized(H, b)
        res = _gpu.jt_rows(AT, 1, AT.shape[1], y)[:, 0]
    for a, e in spans:
        res[a:e] = R @ res[a:e]
    if keep.numel() < n0:
        full = torch.zeros(n0, dtype=res.dtype, device=res.device)
        full[keep] = res
        res = full
    return res.cpu().numpy()


def reginvover(Amat, rhs, lambda_, LTL=None, blocks=None, **kwargs):
    """x = (A^T A + lambda (I | L^T L))^-1 A^T rhs (recon.py:363-426).

    An L^T L of the Hessian's size is added whole, a smaller one to each
    diagonal block of its size; zero columns (and the matching L^T L rows)
    and zero rows are dropped first, as in the reference.
    """
    A, b, cplx = _dense_inputs(Amat, rhs)
    A, b, keep, n0 = _drop_zero(A, b)
    if LTL is not None:
        LTL = np.asarray(LTL.toarray() if hasattr(LTL, "toarray") else LTL, dtype=np.float64)
        if keep.numel() < n0 and LTL.shape[0] > keep.numel():
            li = keep.cpu().numpy()
            li = li[li < LTL.shape[0]]
            LTL = LTL[np.ix_(li, li)]
    if cplx:
        H = A.t() @ A
        g = A.t() @ b
    else:
        H = _gpu.gram(A)
        g = _gpu.jt_cols(A, 1, A.shape[1], b.reshape(-1, 1))  # A^T b
    if LTL is None:
        H.diagonal().add_(lambda_)
    else:
        Lt = torch.as_tensor(LTL, device=A.device).to(H.dtype)
        if Lt.shape[0] == H.shape[0]:
            H = H + lambda_ * Lt
        else:
            nx = Lt.shape[0]
            for i in range(0, H.shape[0], nx):
                e = min(i + nx, H.shape[0])
                H[i:e, i:e] += lambda_ * Lt[: e - i, : e - i]
    x = _solve_sym(H, g, cplx)
    if keep.numel() < n0:
        full = torch.zeros(n0, dtype=x.dtype, device=x.device)
        full[keep] = x
        x = full
    return x.cpu().numpy()


def reginv(Amat, rhs, lambda_, Areg=None, blocks=None, **kwargs):
    """Over- or under-determined branch by shape (recon.py:338-360)."""
    Areg = Areg or {}
    if Amat.shape[0] >= Amat.shape[1]:
        return reginvover(Amat, rhs, lambda_, Areg.get("ltl"), blocks, **kwargs)
    return reginvunder(Amat, rhs, lambda_, Areg.get("lir"), blocks, **kwargs)


# ---------------------------------------------------------------------------
# Gauss-Newton loop
# ---------------------------------------------------------------------------
class GNStep:
    """One normal-equation solve in Kronecker form (recon.py:429-509 with blockscale, :235-240).

    Unknown blocks c (chromophores, or a single 'mua' block) share the node
    dimension; row blocks w (wavelengths) hold npairs rows each.  The
    flattened Jacobian is J[(w,p), (c,n)] = E[w,c] Jw[p,n], so
    J J^T = (E E^T)[w,w'] * (Jw Jw'^T): one Gram over the nodes.

    fp64 : float64 Jacobian, float64 Gram, Cholesky of the diagonally
           normalised s^2 J J^T + lambda I.
    mixed: float64 Jacobian, bf16x3 MFMA Gram (its hi/lo bf16 split) whose
           diagonal is replaced by the float64 column norms, float64
           Cholesky of it and a float32 inverse of the factor as a
           preconditioner, then iterative refinement
           with float64 products J (J^T v) until the residual is below
           ``reftol``.
    Returns dmu (nn x nc) = s^2 J^T y.
    """

    def __init__(self, JT, npairs, E, lam, precision="fp64", reftol=1e-10, maxref=20, scale2=None):
        self.JT = JT
        self.scale2 = scale2  # 1/||J||_F^2 taken before a prior transform of JT (recon.py:235)
        self.np = npairs
        self.E = _gpu.h2d(E, JT.device, np.float64)
        self.lam = float(lam)
        self.precision = precision
        self.reftol = reftol
        self.maxref = maxref
        self.info = {}

    def _hv(self, v, s2, W):
        """H v = s^2 W o (J J^T) v + lambda v with float64 accumulation."""
        nw = W.shape[0]
        if _gpu.jt_hv_ok(self.JT, nw, self.np):
            return _gpu.jt_hv(self.JT, W, self.np, v, float(s2), self.lam)
        t = _gpu.jt_rows(self.JT, nw, self.np, v)  # (nn, nw)
        z = t @ W  # W symmetric
        u = _gpu.jt_cols(self.JT, nw, self.np, z)
        return s2 * u + self.lam * v

    def solve(self, misfit):
        JT, npairs, E = self.JT, self.np, self.E
        nw, nc = E.shape
        W = E @ E.t()
        Wk = torch.repeat_interleave(torch.repeat_interleave(W, npairs, 0), npairs, 1)
        mk = getattr(self, "mark", None) or (lambda name, t: t)
        tq = time.perf_counter()
        if self.precision == "fp64":
            colsq = _gpu.jt_cols(JT, nw, npairs)  # exact ||J_w row||^2, float64
            G = _gpu.gram(JT, None, "fp64")
        else:
            G, colsq = _gpu.gram(JT, None, "bf16x3", want_colsq=True)
            G.diagonal().copy_(colsq)
        wdiag = torch.repeat_interleave(torch.diagonal(W), npairs)
        s2 = 1.0 / torch.sum(wdiag * colsq).item() if self.scale2 is None else self.scale2  # 1/||J_flat||_F^2
        tq = mk("gn_gram", tq)
        H = s2 * (G * Wk)
        del G
        y = self._factor_refine(H, misfit, lambda v: self._hv(v, s2, W), mk, tq,
                                exact=lambda: s2 * (_gpu.gram(JT, None, "fp64") * Wk))
        T = _gpu.jt_rows(JT, nw, npairs, y)  # (nn, nw) = J_w^T y_w
        return s2 * (T @ E)

    def _factor_refine(self, H, misfit, hv, mk, tq, exact=None):
        """y = (H + lambda I)^-1 misfit: Cholesky of the unit-diagonal form; the
        mixed path refines with the float64 operator hv.  When the bf16x3 Gram
        is not positive definite (its ~4e-6 entry error can exceed a tiny
        lambda on a rank-deficient J J^T), ``exact()`` gives the float64 Gram
        form of H to factor instead."""
        H.diagonal().add_(self.lam)
        Hn, di = _normalize_diag(H)
        try:
            L = _gpu.cholesky(Hn)
        except torch.linalg.LinAlgError:
            if self.precision == "fp64" or exact is None:
                raise
            H = exact()
            H.diagonal().add_(self.lam)
            Hn, di = _normalize_diag(H)
            L = _gpu.cholesky(Hn)
            self.info["fp64_gram_fallback"] = True
        if self.precision == "fp64":

            def hsolve(r):
                return di * torch.cholesky_solve((di * r).reshape(-1, 1), L).reshape(-1)

        else:
            # refinement applies the factor ~5 times: an explicit L^-1 (one
            # triangular solve with m right-hand sides) turns each application
            # into two dense products instead of two sequential triangular solves.
            # It is a preconditioner (the residuals are float64), so float32
            # suffices: its ~1e-7 error is below the bf16x3 Gram's ~4e-6.
            L32 = L.float()
            Li = torch.linalg.solve_triangular(L32, torch.eye(L.shape[0], dtype=L32.dtype, device=L.device),
                                               upper=False)
            del L32

            def hsolve(r):
                return di * (Li.t() @ (Li @ (di * r).float())).double()

        tq = mk("gn_cholesky", tq)

        y = hsolve(misfit)
        nref = 0
        if self.precision != "fp64":
            rn = torch.linalg.norm(misfit)
            for nref in range(1, self.maxref + 1):
                res = misfit - hv(y)
                rel = (torch.linalg.norm(res) / rn).item()
                if rel <= self.reftol:
                    break
                y = y + hsolve(res)
            self.info["refine_relres"] = rel
            if rel > self.reftol:
                import warnings

                warnings.warn("Gauss-Newton step: iterative refinement stopped at relative residual %.3e after "
                              "%d passes (target %.1e)" % (rel, self.maxref, self.reftol), _gpu.SolverWarning,
                              stacklevel=3)
        self.info["refinements"] = nref
        mk("gn_refine", tq)
        return y


class GNStepReIm(GNStep):
    """'reim' rows for complex data (recon.py:537-541 with :243-251).

    The real system is A = [[Jr, -Ji], [Ji, Jr]] on the unknowns [x; y] (the
    real and imaginary parts of a complex update).  The flattened Jacobian is
    J[(w,p),(c,n)] = E[w,c] J_w[p,n] (wavelength row blocks w, unknown blocks
    c); the reference's update parsing walks the unknown blocks over the
    first N = nc*nn entries of [x; y], so every block receives its own x (the
    real part of its complex Tikhonov step) and y is never read.  With the
    node-major [Re | Im] Jacobian JT (K x 2m, m = nw*npairs), its Gram
    G = [[Grr, Gri], [Gir, Gii]] and Wk = (E E^T)[w,w'] over the pair blocks:
    A A^T = [[Wk o (Grr + Gii), Wk o (Gri - Gir)], [Wk o (Gir - Gri), Wk o (Grr + Gii)]],
    ||A||_F^2 = 2 ||J||_F^2, and x_c = s^2 sum_w E[w,c] (Jr_w^T y1_w + Ji_w^T y2_w).
    """

    def __init__(self, JT, npairs, E, lam, precision="fp64", reftol=1e-10, maxref=20):
        super().__init__(JT, npairs, E, lam, precision, reftol, maxref)

    def _halves(self, v):
        """(nn, nw) = Jr_w^T v1_w + Ji_w^T v2_w for v = [v1; v2] (each nw*npairs)."""
        nw = self.E.shape[0]
        t = _gpu.jt_rows(self.JT, 2 * nw, self.np, v)
        return t[:, :nw] + t[:, nw:]

    def _hv(self, v, s2, W):
        """(s^2 A A^T + lambda) v with float64 accumulation (four reads of J)."""
        m = v.shape[0] // 2
        nw = W.shape[0]
        za = self._halves(v) @ W  # Jr^T v1 + Ji^T v2, mixed over wavelengths by E E^T
        zb = self._halves(torch.cat([v[m:], -v[:m]])) @ W  # Jr^T v2 - Ji^T v1
        u1 = _gpu.jt_cols(self.JT, 2 * nw, self.np, torch.cat([za, -zb], 1))  # [Jr za; -Ji zb]
        u2 = _gpu.jt_cols(self.JT, 2 * nw, self.np, torch.cat([zb, za], 1))  # [Jr zb; Ji za]
        return s2 * torch.cat([u1[:m] + u1[m:], u2[:m] + u2[m:]]) + self.lam * v

    def _gram_blocks(self):
        """Gram blocks rr, ri, ir, ii of the [Re | Im] Jacobian and the Kronecker weights Wk."""
        JT, E = self.JT, self.E
        m = JT.shape[1] // 2
        if self.precision == "fp64":
            colsq = _gpu.jt_cols(JT, 2 * E.shape[0], self.np)
            G = _gpu.gram(JT, None, "fp64")
        else:
            G, colsq = _gpu.gram(JT, None, "bf16x3", want_colsq=True)
            G.diagonal().copy_(colsq)
        W = E @ E.t()
        Wk = torch.repeat_interleave(torch.repeat_interleave(W, self.np, 0), self.np, 1)
        return G, colsq, W, Wk, m

    def solve(self, misfit):
        JT, E = self.JT, self.E
        mk = getattr(self, "mark", None) or (lambda name, t: t)
        tq = time.perf_counter()
        G, colsq, W, Wk, m = self._gram_blocks()
        wdiag = torch.repeat_interleave(torch.diagonal(W), self.np)
        s2 = 1.0 / (2.0 * torch.sum(wdiag * (colsq[:m] + colsq[m:])).item())
        tq = mk("gn_gram", tq)

        def assemble(Gm):
            rr, ri, ir, ii = Gm[:m, :m], Gm[:m, m:], Gm[m:, :m], Gm[m:, m:]
            d, o = Wk * (rr + ii), Wk * (ri - ir)
            return s2 * torch.cat([torch.cat([d, o], 1), torch.cat([-o, d], 1)], 0)

        H = assemble(G)
        del G
        y = self._factor_refine(H, misfit, lambda v: self._hv(v, s2, W), mk, tq,
                                exact=lambda: assemble(_gpu.gram(JT, None, "fp64")))
        return s2 * (self._halves(y) @ E)  # (nn, nc) = the x of every block


class GNStepComplex(GNStepReIm):
    """reform 'complex' (recon.py:192-198, :235-240, :474-501): the complex
    normal equations of the unconjugated Gauss-Newton step, float64.

    blockscale s = 1/sqrt(sum J^2) with the complex square (the reference's
    np.sqrt(np.sum(Jflat**2)) on complex Jflat, principal branch), H = s^2 J J^T
    + lambda I (complex symmetric: J^T, not J^H), y = D solve(D H D, D r) with
    D = 1/sqrt(|diag H| + 1e-16), and dmu = s^2 J^T y, complex.  From the
    [Re | Im] Gram blocks: J J^T = Wk o ((Grr - Gii) + i (Gri + Gir)).
    """

    def solve(self, misfit):
        JT, E = self.JT, self.E
        self.precision = "fp64"
        G, colsq, W, Wk, m = self._gram_blocks()
        wdiag = torch.repeat_interleave(torch.diagonal(W), self.np)
        sumsq = complex(torch.sum(wdiag * (colsq[:m] - colsq[m:])).item(),
                        2.0 * torch.sum(wdiag * torch.diagonal(G[:m, m:])).item())
        s2 = 1.0 / sumsq  # (1/sqrt(sum J^2))^2
        rr, ri, ir, ii = G[:m, :m], G[:m, m:], G[m:, :m], G[m:, m:]
        H = s2 * torch.complex(Wk * (rr - ii), Wk * (ri + ir))
        del G
        H.diagonal().add_(self.lam)
        Hn, di = _normalize_diag(H)
        r = misfit.to(torch.complex128)
        y = di * torch.linalg.solve(Hn, di * r)
        yr, yi = y.real.contiguous(), y.imag.contiguous()
        re = self._halves(torch.cat([yr, -yi]))  # Jr^T yr - Ji^T yi
        im = self._halves(torch.cat([yi, yr]))  # Jr^T yi + Ji^T yr
        self.info["refinements"] = 0
        return s2 * (torch.complex(re, im) @ E.to(torch.complex128))


def _solve_over(JG, npairs, E, lam, misfit, LTL=None):
    """reginvover (recon.py:363-426) on the flattened Jacobian, for few unknowns
    (labels or a small mesh): J[(b,p),(c,n)] = E[b,c] JG[n, b*np+p], blockscale
    s = 1/||J||_F, zero columns and rows dropped, (J^T J + lam (I | L^T L)) x = J^T r
    in the unit-diagonal form; returns s * x as (n x nc)."""
    N, M = JG.shape
    nb, nc = E.shape
    A = (JG.t().reshape(nb, npairs, 1, N) * E[:, None, :, None]).reshape(nb * npairs, nc * N)
    s = 1.0 / torch.sqrt(torch.sum(A * A))
    A = A * s
    keep = torch.nonzero(torch.sum(torch.abs(A), 0) != 0).reshape(-1)
    n0 = A.shape[1]
    if keep.numel() < n0:
        A = A[:, keep]
        if LTL is not None and LTL.shape[0] > keep.numel():
            li = keep.cpu().numpy()
            li = li[li < LTL.shape[0]]
            LTL = LTL[np.ix_(li, li)]
    rows = torch.sum(torch.abs(A), 1) != 0
    A, r = A[rows], misfit[rows]
    g = A.t() @ r
    H = A.t() @ A
    if LTL is None:
        H.diagonal().add_(lam)
    else:
        Lt = torch.as_tensor(LTL, dtype=torch.float64, device=A.device)
        if Lt.shape[0] == H.shape[0]:
            H = H + lam * Lt
        else:
            nx = Lt.shape[0]
            for i in range(0, H.shape[0], nx):
                e = min(i + nx, H.shape[0])
                H[i:e, i:e] += lam * Lt[: e - i, : e - i]
    Hn, di = _normalize_diag(H)
    x = di * torch.linalg.solve(Hn, di * g)
    if keep.numel() < n0:
        full = torch.zeros(n0, dtype=x.dtype, device=x.device)
        full[keep] = x
        x = full
    return (s * x).reshape(nc, N).t()


class JacobianDict(Mapping):
    """runrecon's ``Jmua`` (recon.py:155-179, returned at :335): the node Jacobians of the
    last iteration's forward fields, keyed like the reference -- chromophore -> the
    wavelength-stacked E[w, c] J_w (jacchrome, forward.py:572-589) for multispectral
    unknowns, else 'mua' (one wavelength) or the wavelength.  Each entry is a fresh
    host array (pairs x nodes, user node order; complex for frequency-domain data),
    built on access from the device-resident fields by the Jacobian kernel, so the
    several-GB matrices exist on the host only when asked for.
    """

    def __init__(self, mesh, X, k, wvs, pairs, cplx, chs=None, E=None):
        self._mesh, self._X, self._k = mesh, X, k
        self._wvs, self._pairs, self._cplx = list(wvs), pairs, cplx
        self._chs = list(chs) if chs else None
        self._E = None if E is None else np.asarray(E, dtype=float)
        self._keys = self._chs or (["mua"] if self._wvs == [""] else self._wvs)

    def _jw(self, i):
        s, d = self._pairs[i]
        dt = torch.complex128 if self._cplx else torch.float64
        JT, _ = _gpu.jac_node(self._mesh, self._X[:, i * self._k : (i + 1) * self._k], s, d, dt)
        return self._mesh.user_rows(JT).t().cpu().numpy()

    def __getitem__(self, key):
        if key not in self._keys:
            raise KeyError(key)
        if self._chs:
            c = self._chs.index(key)
            return np.vstack([self._E[i, c] * self._jw(i) for i in range(len(self._wvs))])
        return self._jw(0 if key == "mua" else self._wvs.index(key))

    def __iter__(self):
        return iter(self._keys)

    def __len__(self):
        return len(self._keys)


# parameters runrecon updates (recon.py:278); others (aa3) keep their Jacobian block but no update
_UPDATED_PARAMS = ("hbo", "hbr", "water", "lipids", "scatamp", "scatpow")


def _chromophores(cfg):
    if isinstance(cfg.get("prop"), dict) and "param" in cfg:
        return [k for k in cfg["param"].keys() if k in CHROMOPHORES]
    return []


def runrecon(cfg, recon, detphi0, sd=None, **kwargs):
    """Iterative Gauss-Newton reconstruction (recon.py:39-335).

    Supported on the GPU: a single-wavelength node-based mua unknown
    (recon['prop'] with nn rows), and multispectral chromophore unknowns
    (recon['param'] node arrays); CW or frequency-domain data with reform
    'real' ([Re; Im] rows for complex data, recon.py:532-546), 'reim'
    ([[Re, -Im], [Im, Re]] rows, :537-541, node unknowns on one mesh),
    'complex' (complex normal equations, :192-198, float64) or 'logphase'
    (log-amplitude and phase rows, :548-565); priors (lmat, ltl, or
    prior= with recon['seg'], :127-133, :225-233) and label compression by a 1-D
    recon['seg'] (:210-218), with the overdetermined step when the unknowns are
    few (:363-426); a dual mesh (recon['node'],
    ['elem'], ['mapid'], ['mapweight']) for a single unknown block, the
    Jacobian moved onto the recon nodes as in recon.py:200-205.  Extra
    keywords: precision ('fp64' | 'mixed'), extin (dict wavelength ->
    {chromophore: value} replacing the extinction table), solver tolerance
    ``solvertol`` (default 1e-12, mixed 1e-10).  Log-amplitude rows need
    each detector value to relative accuracy: a value at r times the largest
    one carries ~solvertol/r relative error (an iterative solve from zero
    leaves values below ~solvertol of the peak at exactly 0, where
    'logphase' is undefined).
    Returns (recon, resid, cfg, updates, Jmua, detphi, phi) like the reference;
    Jmua is a JacobianDict with the reference's keys (chromophores, or 'mua',
    or wavelengths) whose entries are materialised on access (they can be
    several GB); ``return_jac=True`` materialises them all into a dict.
    """
    maxiter = kwargs.get("maxiter", 5)
    lam = kwargs.get("lambda_", recon.get("lambda", 0.05))
    report = kwargs.get("report", True)
    tol = kwargs.get("tol", 0)
    reform = kwargs.get("reform", "real")
    rfcw = kwargs.get("rfcw", [1])
    precision = kwargs.get("precision", "fp64")
    extin = kwargs.get("extin", None)
    stol = kwargs.get("solvertol", 1e-12 if precision == "fp64" else 1e-10)
    return_jac = kwargs.get("return_jac", False)
    warm = kwargs.get("warmstart", True)  # start each forward solve from the last iteration's fields
    optodes = kwargs.get("optodes", None)
    if isinstance(rfcw, int):
        rfcw = [rfcw]
    if reform not in ("real", "reim", "complex", "logphase"):
        raise ValueError("Unknown form: %s" % reform)  # matreform's error (recon.py:567)
    if rfcw != [1]:
        # the reference passes runforward's per-mode dict of fields to jac (recon.py:151-158,
        # forward.py:513), which fails there; one mode per reconstruction here
        raise NotImplementedError("runrecon reconstructs one mode at a time (rfcw=[1])")
    Aregu = {}  # recon.py:127-133: an lmat keyword, else ltl, else a prior built from recon['seg']
    if "lmat" in kwargs:
        Aregu["lmat"] = np.asarray(kwargs["lmat"], dtype=float)
    elif "ltl" in kwargs:
        Aregu["ltl"] = np.asarray(kwargs["ltl"], dtype=float)
    elif kwargs.get("prior") and "seg" in recon:
        lm = prior(recon["seg"], kwargs["prior"], kwargs)
        if lm is not None:
            Aregu["lmat"] = lm.toarray() if hasattr(lm, "toarray") else np.asarray(lm, dtype=float)
    if sd is None:
        sd = sdmap(cfg)
    if "prop" in recon and isinstance(recon["prop"], np.ndarray) and recon["prop"].ndim == 1:
        recon["prop"] = recon["prop"].reshape(1, -1)
    resid = np.zeros(maxiter)
    updates = []
    X0 = None
    detphi_out = phi_out = Jm_out = None
    timing = kwargs.get("timing", None)  # dict: per-phase seconds (synchronising) when given
    dual = "node" in recon and "elem" in recon and "mapid" in recon
    remap = None  # device CSR of the dual-mesh Jacobian map, built once

    def mark(name, t_prev):
        if timing is None:
            return t_prev
        torch.cuda.synchronize()
        t = time.perf_counter()
        timing[name] = timing.get(name, 0.0) + (t - t_prev)
        return t

    for it in range(maxiter):
        t0 = time.time()
        tp = time.perf_counter()
        if "param" in recon or "prop" in recon:
            cfg, recon = syncprop(cfg, recon)
        if "param" in cfg and isinstance(cfg.get("prop"), dict):
            cfg["prop"] = updateprop(cfg, extin=extin)
        tp = mark("props", tp)
        res = forward_device(cfg, rfcw, precision, stol, None, X0 if warm else None, optodes=optodes,
                             where="runrecon iteration %d" % (it + 1))
        optodes = res.optodes  # reused next iteration where the moved optode positions are unchanged
        X0 = res.X
        tp = mark("forward", tp)
        if timing is not None:
            timing.setdefault("forward_iters", []).append(res.info.iters)
        wvs = list(cfg["prop"].keys()) if isinstance(cfg.get("prop"), dict) else [""]
        chs = _chromophores(cfg)
        mesh = _gpu.mesh_for(cfg)
        detphi = {wv: res.detvals[i].cpu().numpy() for i, wv in enumerate(wvs)}
        if wvs == [""]:
            detphi = detphi[""]
        y0 = _flatten_detphi(detphi0, sd, wvs, rfcw)
        y = _flatten_detphi(detphi, sd, wvs, rfcw)
        cplx = res.is_complex
        # rows of the real system: (part, wavelength, pair) with part Re [, Im]
        # (recon.py:532-565); 'logphase' scales each complex row by
        # conj(y)/|y|^2 before splitting
        nparts = 2 if cplx else 1
        rowscale = None
        if reform == "logphase":
            rowscale = np.conj(y) / np.abs(y * y)
            misfit = np.log(np.abs(y0)) - np.log(np.abs(y))
            if cplx:
                misfit = np.concatenate([misfit, np.angle(y0) - np.angle(y)])
        elif cplx and reform == "complex":
            misfit = y0 - y  # the reference skips matreform for 'complex' (recon.py:192-198)
        elif cplx:
            misfit = np.concatenate([np.real(y0 - y), np.imag(y0 - y)])
        else:
            misfit = np.real(y0 - y)
        npairs = None
        JT = None
        nw = len(wvs)
        jdt = torch.float64  # both precisions: the refined step is that of this Jacobian
        for i, wv in enumerate(wvs):
            sdw = sd.get(wv, sd) if isinstance(sd, dict) else sd
            s, d = _active_pairs(sdw)
            if JT is None:
                npairs = len(s)
                JT = torch.empty((mesh.nn, nparts * nw * npairs), dtype=jdt, device=mesh.dev)
            elif len(s) != npairs:
                raise NotImplementedError("all wavelengths must have the same active pairs")
            blk = slice(i * npairs, (i + 1) * npairs)
            if not cplx and rowscale is None:
                _gpu.jac_node(mesh, res.phi(i), s, d, jdt, out=JT[:, blk])
                continue
            Jw, _ = _gpu.jac_node(mesh, res.phi(i), s, d, torch.complex128 if cplx else jdt)
            if rowscale is not None:
                sc = _gpu.h2d(rowscale[blk] if cplx else np.real(rowscale[blk]), mesh.dev)
                Jw = Jw * sc[None, :]
            JT[:, blk] = Jw.real
            if cplx:
                JT[:, nw * npairs + i * npairs : nw * npairs + (i + 1) * npairs] = Jw.imag
            del Jw
        tp = mark("jacobian", tp)
        resid[it] = np.sum(np.abs(misfit))
        if chs:
            if extin is None:
                E = extinction(wvs, chs)[0]
            else:
                E = np.array([[extin[wv][c] for c in chs] for wv in wvs], dtype=float)
            keys = chs
        else:
            if len(wvs) != 1:
                raise NotImplementedError("multi-wavelength mua reconstruction needs chromophore params")
            E = np.ones((1, 1))
            keys = ["mua"]
        cform = cplx and reform in ("reim", "complex")  # Re/Im parts of complex rows and unknowns
        if nparts == 2 and not cform:
            E = np.vstack([E, E])  # the Im row blocks carry the same extinction weights
        JG = JT
        if dual:
            if len(keys) != 1:
                # the reference maps one block of forward-node columns (recon.py:842-874)
                raise NotImplementedError("dual-mesh reconstruction supports a single unknown block")
            nrec = np.asarray(recon["node"]).shape[0]
            if remap is None:
                remap = _gpu.remap_operator(recon["mapid"], recon["mapweight"], recon["elem"], nrec,
                                            perm_rows=mesh.iperm, dev=mesh.dev)
            JG = _gpu.jt_remap(remap, JT, nrec)
        # label compression (recon.py:210-218): node columns summed per label
        labelled = False
        if "seg" in recon and np.ndim(recon["seg"]) == 1:
            seg = np.asarray(recon["seg"])
            labs = np.unique(seg)
            if len(seg) == JG.shape[0] and len(labs) < 50:
                seg_rows = seg if dual else mesh.mesh_nodes(seg)  # JG rows are in device order
                S = torch.as_tensor(seg_rows[None, :] == labs[:, None], dtype=torch.float64, device=mesh.dev)
                JG = S @ JG  # (nlabels x M), one-hot product: a fixed summation order
                labelled = True
        ncol = JG.shape[0] * E.shape[1]
        over = JG.shape[1] >= ncol  # reginv's branch by the flattened shape (recon.py:356)
        if it == 0 and "lmat" in Aregu and "ltl" not in Aregu:  # recon.py:225-233
            if over:
                Aregu["ltl"] = Aregu["lmat"].T @ Aregu["lmat"]
            else:
                from scipy.linalg import qr

                Aregu["lir"] = np.linalg.inv(np.triu(qr(Aregu["lmat"], mode="r")[0]))
        if cform:
            # 'reim' ([[Jr, -Ji], [Ji, Jr]] rows) and 'complex' (complex normal equations) keep
            # the node unknowns of every block on one mesh
            if dual or labelled or Aregu or over:
                raise NotImplementedError("reform '%s' is supported for node unknowns on one mesh, without "
                                          "priors or label compression" % reform)
        misfit_t = _gpu.h2d(misfit, mesh.dev, np.complex128 if np.iscomplexobj(misfit) else np.float64)
        step = None
        if over:
            tp = mark("misfit", tp)
            dmu = _solve_over(JG, npairs, _gpu.h2d(E, mesh.dev, np.float64), lam,
                              misfit_t, Aregu.get("ltl"))
        else:
            invR = None
            lir = Aregu.get("lir")
            if lir is not None and lir.shape[0] == JG.shape[0]:
                # recon.py:458-470 transforms the unknown blocks of the prior's size;
                # every block here has JG.shape[0] columns (a prior of another size
                # transforms nothing, as in the reference)
                invR = torch.as_tensor(lir, dtype=torch.float64, device=mesh.dev)
                if not (dual or labelled) and mesh.reordered:
                    perm = torch.as_tensor(mesh.perm, device=mesh.dev)
                    invR = invR[perm][:, perm]  # into the device node order of JG's rows
                # the blockscale is that of the untransformed J (recon.py:235-236)
                Et = _gpu.h2d(E, mesh.dev, np.float64)
                wdiag = torch.repeat_interleave(torch.diagonal(Et @ Et.t()), npairs)
                s2pre = 1.0 / torch.sum(wdiag * _gpu.jt_cols(JG, E.shape[0], npairs)).item()
                JG = invR.t() @ JG  # A' = A R^-1 per block: node-major J' = R^-T J
            else:
                s2pre = None
            if reform == "complex" and cplx:
                step = GNStepComplex(JG, npairs, E, lam, "fp64")
            elif reform == "reim" and cplx:
                step = GNStepReIm(JG, npairs, E, lam, precision, reftol=kwargs.get("reftol", 1e-10))
            else:
                step = GNStep(JG, npairs, E, lam, precision, reftol=kwargs.get("reftol", 1e-10), scale2=s2pre)
            step.mark = (lambda name, t: mark(name, t)) if timing is not None else None
            tp = mark("misfit", tp)
            dmu = step.solve(misfit_t)
            if invR is not None:
                dmu = invR @ dmu  # back to the unknowns (recon.py:487-499)
        dmu = (dmu if (dual or labelled) else mesh.user_rows(dmu)).cpu().numpy()
        del JG
        tp = mark("gn_solve", tp)
        if timing is not None and step is not None:
            timing.setdefault("refinements", []).append(step.info.get("refinements", 0))
        update = {}
        for c, key in enumerate(keys):
            dx = dmu[:, c]
            update[key] = dx
            if key == "mua":  # recon.py:255-276
                if "prop" in recon and isinstance(recon["prop"], np.ndarray):
                    pr = recon["prop"]
                    if pr.shape[0] < len(dx) and pr.shape[0] < 50:  # label table, node update
                        for li in range(min(pr.shape[0], len(dx))):
                            pr[li, 0] += dx[li]
                    else:
                        pr[:, 0] += dx
            elif key in _UPDATED_PARAMS:
                tgt = recon if ("param" in recon and key in recon["param"]) else cfg
                if "param" not in tgt or key not in tgt["param"]:
                    continue
                pv = tgt["param"][key]
                npar = len(pv) if hasattr(pv, "__len__") else 1
                if npar < len(dx) and npar < 50:
                    if npar == 1:
                        if hasattr(pv, "__len__"):
                            tgt["param"][key][0] += dx[0]
                        else:
                            tgt["param"][key] += dx[0]
                    else:
                        for li in range(min(npar, len(dx))):
                            tgt["param"][key][li] += dx[li]
                else:
                    tgt["param"][key] = pv + dx
        updates.append(update)
        detphi_out = detphi
        if report:
            rel = resid[it] / resid[0] if it > 0 else 1.0
            print("iter [%4d]: residual=%.6e, relres=%.6e lambda=%.6e (time=%.2f s)"
                  % (it + 1, resid[it], rel, lam, time.time() - t0))
        if it > 0 and abs(resid[it] - resid[it - 1]) / resid[0] < tol:
            resid = resid[: it + 1]
            break
    pairs = [_active_pairs(sd.get(wv, sd) if isinstance(sd, dict) else sd) for wv in wvs]
    Jm_out = JacobianDict(mesh, X0, res.k, wvs, pairs, res.is_complex, chs or None, E[: len(wvs)] if chs else None)
    if return_jac:
        Jm_out = dict(Jm_out.items())
    tp = time.perf_counter()
    Xh = _gpu.to_host(mesh.user_rows(X0))
    mark("outputs", tp)
    k = res.k
    phi_out = {wv: Xh[:, i * k : (i + 1) * k] for i, wv in enumerate(wvs)}
    if wvs == [""]:
        phi_out = phi_out[""]
    recon["lambda"] = lam
    return recon, resid, cfg, updates, Jm_out, detphi_out, phi_out
