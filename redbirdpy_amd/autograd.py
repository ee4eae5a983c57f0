"""Detector values as a differentiable torch function of nodal mua.

``detval_mua(cfg, mua)`` runs the GPU forward (assembly, multi-RHS solve,
femgetdet) and returns the (ndet x nsrc) detector values as a tensor that
autograd can differentiate with respect to the nodal absorption ``mua``.
The backward is hand-written: by reciprocity the detector columns of the
forward solve are the adjoint fields, so the vector-Jacobian product is the
Jacobian contraction g[n] = sum_{d,s} w[d,s] J[(s,d), n] over the stored
fields (rb_jac_contract_grid), without forming J.

The forward follows the derivative model of redbirdpy's ``jac``
(forward.py:456-569): element mua is the mean of its four nodal values and
the diffusion coefficient is held fixed (mus' moves opposite to mua), with
the optodes kept where the base properties put them; so the gradient is
exactly sum_p w_p Jnode[p, :] with Jnode from ``jac``.  Frequency-domain
(complex) detector values follow torch's convention for a real input:
grad = Re(sum_p conj(w_p) J_p).
"""

__all__ = ["detval_mua", "DetvalMua", "MuaForward"]

import numpy as np
import torch

from . import _gpu
from .forward import forward_device, locate_optodes


class MuaForward:
    """Fixed part of the map mua -> detector values: mesh, base optics, located optodes."""

    def __init__(self, cfg, precision="fp64", tol=1e-12, maxiter=None):
        if isinstance(cfg.get("prop"), dict):
            raise NotImplementedError("detval_mua takes a single-wavelength cfg")
        self.cfg = dict(cfg)
        self.precision = precision
        self.tol = tol
        self.maxiter = maxiter
        self.mesh = _gpu.mesh_for(cfg)
        props, _, _ = _gpu.props_for(cfg)
        kind, plen, arrays = _gpu.expand_props(cfg, props)
        ne = cfg["elem"].shape[0]
        e0 = np.asarray(cfg["elem"])[:, :4].astype(np.int64) - 1
        if arrays[0] == "label":
            tab, seg = arrays[1], np.asarray(arrays[2]).astype(np.int64)
            mua0, musp0, nref = tab[seg, 0], tab[seg, 1], np.full(ne, tab[0, 2])
        else:
            mua0, musp0, nref = arrays[1], arrays[2], np.asarray(arrays[3], dtype=float)
            if len(mua0) != ne:  # node rows: element values are the means over the four nodes
                nref = nref[e0].mean(1) if len(nref) == len(mua0) else np.full(ne, nref.ravel()[0])
                mua0, musp0 = mua0[e0].mean(1), musp0[e0].mean(1)
            elif len(nref) != ne:
                nref = np.full(ne, nref.ravel()[0])
        self.mutot = mua0 + musp0  # held fixed: D = 1 / (3 (mua + mus'))
        self.nref = np.asarray(nref, dtype=float)
        self.e0 = torch.as_tensor(e0, device=self.mesh.dev)
        self.optodes = locate_optodes(cfg)

    def run(self, mua):
        """Forward at nodal mua (user numbering, device tensor): (detval, fields, src cols, det cols)."""
        mua_e = mua.detach().to(torch.float64)[self.e0].mean(1).cpu().numpy()
        c = dict(self.cfg)
        c["prop"] = np.column_stack([mua_e, self.mutot - mua_e, np.zeros_like(mua_e), self.nref])
        c.pop("seg", None)
        res = forward_device(c, [1], self.precision, self.tol, self.maxiter, optodes=self.optodes,
                             keep_optodes=True)
        op = res.optodes[""]
        ts = op.ns + op.ws
        S = np.arange(ts)
        D = np.arange(ts, ts + op.nd + op.wd)
        return res.detvals[0], res.X, S, D


class DetvalMua(torch.autograd.Function):
    """(ndet x nsrc) detector values of the fixed problem at nodal mua, with the adjoint backward."""

    @staticmethod
    def forward(ctx, mua, model):
        det, X, S, D = model.run(mua)
        ctx.model = model
        ctx.fields = X
        ctx.S, ctx.D = S, D
        return det

    @staticmethod
    def backward(ctx, grad):
        m = ctx.model
        w = torch.conj_physical(grad) if grad.is_complex() else grad
        g = _gpu.jac_contract(m.mesh, ctx.fields, ctx.S, ctx.D, w.reshape(-1))
        g = m.mesh.user_rows(g.real if g.is_complex() else g)
        return g.to(torch.float64), None


def detval_mua(cfg, mua, precision="fp64", tol=1e-12, model=None):
    """Detector values (ndet x nsrc) as a differentiable function of nodal mua (nn tensor on the GPU).

    cfg: a meshprep'ed single-wavelength cfg (its properties give the base
    scattering, refractive index and optode positions).  Pass ``model`` (a
    MuaForward) to reuse the set-up across calls.
    """
    model = model or MuaForward(cfg, precision, tol)
    return DetvalMua.apply(mua, model)
