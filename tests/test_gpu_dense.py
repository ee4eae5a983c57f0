"""GPU Gram (fp64 FMA, bf16 / bf16x3 MFMA) and Jacobian contraction against float64 references."""

import numpy as np
import pytest
import torch

from conftest import box_cfg, grid_optodes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("K,m", [(5000, 300), (1031, 128), (64, 70)])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_gram_modes(K, m, dtype):
    from redbirdpy_amd import _gpu

    rng = np.random.default_rng(K + m)
    # rows spanning orders of magnitude, like Jacobian rows
    J = rng.standard_normal((K, m)) * np.exp(rng.uniform(-6, 0, m))[None, :]
    JT = torch.as_tensor(J, device="cuda").to(dtype)
    ref = J.astype(np.float64) if dtype == torch.float64 else JT.double().cpu().numpy()
    G_ref = ref.T @ ref
    d = torch.as_tensor(1 / np.sqrt(np.diag(G_ref)), device="cuda")
    Gn_ref = G_ref * np.outer(d.cpu().numpy(), d.cpu().numpy())
    G64 = _gpu.gram(JT, d, "fp64").cpu().numpy()
    assert np.max(np.abs(G64 - Gn_ref)) < 1e-12
    G3 = _gpu.gram(JT, d, "bf16x3").cpu().numpy()
    assert np.max(np.abs(G3 - Gn_ref)) < 2e-5  # ~16-bit products, unit diagonal
    G1 = _gpu.gram(JT, d, "bf16").cpu().numpy()
    assert np.max(np.abs(G1 - Gn_ref)) < 2e-2
    assert np.array_equal(G3, G3.T)
    # the exact float64 column sums of squares taken from the split pass
    G3b, colsq = _gpu.gram(JT, d, "bf16x3", want_colsq=True)
    assert torch.equal(G3b, torch.as_tensor(G3, device="cuda"))
    np.testing.assert_allclose(colsq.cpu().numpy(), np.diag(G_ref), rtol=1e-13)


def test_jac_contract_matches_dense():
    import redbirdpy_amd as rb
    from redbirdpy_amd import _gpu

    p1 = (16, 16, 8)
    cfg = box_cfg(p1, 2.0, src=grid_optodes(p1, 2), det=grid_optodes(p1, 3, margin=3.0))
    cfg, sd = rb.meshprep(cfg)
    det, phi = rb.runforward(cfg)
    Jn, _ = rb.jac(sd, phi, cfg["deldotdel"], cfg["elem"], cfg["evol"], return_elem=False)
    rng = np.random.default_rng(1)
    w = rng.standard_normal(Jn.shape[0])
    mesh = _gpu.mesh_for(cfg)
    S = np.arange(4)
    D = np.arange(4, 13)
    phid = mesh.mesh_rows(torch.as_tensor(phi, device="cuda"))
    g = mesh.user_rows(_gpu.jac_contract(mesh, phid, S, D, torch.as_tensor(w, device="cuda")))
    np.testing.assert_allclose(g.cpu().numpy(), Jn.T @ w, rtol=1e-11, atol=1e-14 * np.abs(Jn).max())


def _hv_ref(J, W, npairs, v, s2, lam):
    K, M = J.shape
    nb = M // npairs
    t = (J * v[None, :]).reshape(K, nb, npairs).sum(-1)
    z = t @ W
    u = (J.reshape(K, nb, npairs) * z[:, :, None]).sum(0).reshape(-1)
    return s2 * u + lam * v


@pytest.mark.parametrize("K,nb,npairs", [(3001, 4, 1024), (2500, 3, 64), (777, 16, 64), (4099, 2, 1024),
                                         (1500, 4, 128), (900, 8, 1024)])
def test_jt_hv_matches_two_pass(K, nb, npairs):
    """The one-pass s2 J W J^T v + lam v (refinement operator) against float64 numpy and the two-pass kernels."""
    from redbirdpy_amd import _gpu

    rng = np.random.default_rng(K + nb)
    M = nb * npairs
    J = rng.standard_normal((K, M)) * np.exp(rng.uniform(-4, 0, M))[None, :]
    E = rng.uniform(0.1, 1.0, (nb, 3))
    W = E @ E.T
    v = rng.standard_normal(M)
    s2, lam = 0.37, 1e-3
    JT = torch.as_tensor(J, device="cuda")
    Wd = torch.as_tensor(W, device="cuda")
    vd = torch.as_tensor(v, device="cuda")
    assert _gpu.jt_hv_ok(JT, nb, npairs)
    out = _gpu.jt_hv(JT, Wd, npairs, vd, s2, lam)
    ref = _hv_ref(J, W, npairs, v, s2, lam)
    scale = s2 * np.abs(J).T @ (np.abs(J) @ np.abs(v)) * W.max() + lam * np.abs(v)
    assert np.all(np.abs(out.cpu().numpy() - ref) <= 1e-12 * scale)
    # the two-pass route (rows then columns)
    t = _gpu.jt_rows(JT, nb, npairs, vd)
    two = s2 * _gpu.jt_cols(JT, nb, npairs, t @ Wd) + lam * vd
    assert np.all(np.abs(out.cpu().numpy() - two.cpu().numpy()) <= 1e-12 * scale)
    # deterministic
    assert torch.equal(out, _gpu.jt_hv(JT, Wd, npairs, vd, s2, lam))


def test_jt_hv_rejects_unsupported_shapes():
    from redbirdpy_amd import _gpu

    JT = torch.zeros((10, 3 * 50), dtype=torch.float64, device="cuda")
    assert not _gpu.jt_hv_ok(JT, 3, 50)
    with pytest.raises(RuntimeError):
        _gpu.jt_hv(JT, torch.eye(3, dtype=torch.float64, device="cuda"), 50, torch.zeros(150, device="cuda"), 1.0, 0.0)


@pytest.mark.parametrize("n", [64, 200, 1000, 4096])
def test_cholesky_matches_lapack(n):
    """Blocked HIP Cholesky (rb_potrf) against LAPACK on a normalised Gram-like SPD matrix."""
    from redbirdpy_amd import _gpu

    rng = np.random.default_rng(n)
    X = rng.standard_normal((n, n + 50)) * np.exp(rng.uniform(-3, 0, n))[:, None]
    H = X @ X.T + 1e-3 * np.eye(n)
    d = 1 / np.sqrt(np.diag(H))
    H = H * np.outer(d, d)
    Hd = torch.as_tensor(H, device="cuda")
    Hd_up = Hd.clone()
    iu = torch.triu_indices(n, n, 1, device="cuda")
    Hd_up[iu[0], iu[1]] = 7.0  # the factor reads only the lower triangle
    L = _gpu.cholesky(Hd).cpu().numpy()
    Lref = np.linalg.cholesky(H)
    assert np.array_equal(L, np.tril(L))
    assert np.max(np.abs(L - Lref)) < 1e-10 * np.max(np.abs(Lref))
    assert np.max(np.abs(L @ L.T - H)) < 1e-13 * n
    assert np.array_equal(_gpu.cholesky(Hd_up).cpu().numpy(), L)
    # the solve the Gauss-Newton step makes with it
    b = rng.standard_normal(n)
    x = torch.cholesky_solve(torch.as_tensor(b, device="cuda").reshape(-1, 1), torch.as_tensor(L, device="cuda"))
    np.testing.assert_allclose(x.cpu().numpy().ravel(), np.linalg.solve(H, b), rtol=1e-8, atol=1e-10)


def test_cholesky_reports_first_bad_pivot():
    from redbirdpy_amd import _gpu

    H = np.eye(150)
    H[70, 70] = -1.0
    with pytest.raises(torch.linalg.LinAlgError, match="order 71"):
        _gpu.cholesky(torch.as_tensor(H, device="cuda"))
