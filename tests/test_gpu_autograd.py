"""The differentiable forward (redbirdpy_amd.autograd): gradients against rb.jac,
the oracle's jac and central finite differences of the same forward map."""

import numpy as np
import pytest

from conftest import box_cfg, grid_optodes

pytestmark = pytest.mark.gpu


def _setup(omega, tol=1e-12):
    import torch
    import redbirdpy_amd as rb

    p1 = (16, 16, 8)
    cfg = box_cfg(p1, 2.0, src=grid_optodes(p1, 2), det=grid_optodes(p1, 3, margin=3.0), omega=omega)
    cfg, sd = rb.meshprep(cfg)
    nn = cfg["node"].shape[0]
    model = rb.autograd.MuaForward(cfg, tol=tol)
    mua = torch.full((nn,), 0.01, dtype=torch.float64, device="cuda", requires_grad=True)
    return rb, cfg, sd, model, mua


@pytest.mark.parametrize("omega", [0.0, 2 * np.pi * 1e8])
def test_detval_mua_forward_and_gradient_match_jac(omega):
    import torch
    from oracle import rb_oracle as orc

    rb, cfg, sd, model, mua = _setup(omega)
    det = rb.autograd.detval_mua(cfg, mua, model=model)
    # at the base properties the map is the plain forward
    d0, phi = rb.runforward(cfg)
    assert np.max(np.abs(det.detach().cpu().numpy() - d0)) <= 1e-10 * np.max(np.abs(d0))
    rng = np.random.default_rng(17)
    w = rng.standard_normal(det.shape)
    if det.is_complex():
        w = w + 1j * rng.standard_normal(det.shape)
        loss = (torch.as_tensor(w.conj(), device="cuda") * det).real.sum()
    else:
        loss = (torch.as_tensor(w, device="cuda") * det).sum()
    loss.backward()
    g = mua.grad.cpu().numpy()
    # sum_p w_p J[p, :] with p = (s, d) rows of sd (row = d * Nsrc + s)
    Jn, _ = rb.jac(sd, phi, cfg["deldotdel"], cfg["elem"], cfg["evol"], return_elem=False)
    ref = np.real(w.conj().ravel() @ Jn) if np.iscomplexobj(w) else w.ravel() @ Jn
    assert np.max(np.abs(g - ref)) <= 1e-10 * np.max(np.abs(ref))
    oc, osd = orc.meshprep(box_cfg((16, 16, 8), 2.0, src=grid_optodes((16, 16, 8), 2),
                                   det=grid_optodes((16, 16, 8), 3, margin=3.0), omega=omega))
    od, ophi = orc.runforward(oc)
    Jo, _ = orc.jac(osd, ophi, oc["deldotdel"], oc["elem"], oc["evol"])
    refo = np.real(w.conj().ravel() @ Jo) if np.iscomplexobj(w) else w.ravel() @ Jo
    assert np.max(np.abs(g - refo)) <= 1e-8 * np.max(np.abs(refo))


@pytest.mark.parametrize("omega", [0.0, 2 * np.pi * 1e8])
def test_detval_mua_gradient_matches_finite_differences(omega):
    """Central differences of the same map (step 1e-6 relative) on 16 seeded nodes.

    The iterative solves carry an error of about tol * |detval| that does not
    vary smoothly with mua: a 1e-6 relative step needs tol well below 1e-12
    for the difference quotient to resolve the gradient, so they run to 1e-14."""
    import torch

    rb, cfg, sd, model, mua = _setup(omega, tol=1e-14)
    rng = np.random.default_rng(3)
    w = torch.as_tensor(rng.standard_normal((model.run(mua)[0].shape)), device="cuda")

    def loss_of(m):
        d = rb.autograd.detval_mua(cfg, m, model=model)
        return ((w * d).real if d.is_complex() else w * d).sum()

    loss_of(mua).backward()
    g = mua.grad.cpu().numpy()
    h = 1e-6 * 0.01
    for n in rng.choice(mua.shape[0], 16, replace=False):
        vals = []
        for sgn in (+1, -1):
            m = mua.detach().clone()
            m[n] += sgn * h
            vals.append(loss_of(m).item())
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(g[n] - fd) <= 1e-5 * np.max(np.abs(g)) + 1e-6 * abs(fd), (n, g[n], fd)
