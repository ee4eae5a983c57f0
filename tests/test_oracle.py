"""Pin the CPU oracle against what the mathematics fixes (no GPU).

Each check is chosen so that a plausible slip in the oracle (a dropped
term, a wrong weight, sign or index, a transposed operand) fails it:
closed forms of the Kuhn grid, brute-force FEM integrals by quadrature,
finite differences, reciprocity, library routines (lstsq, quad).
"""

import numpy as np
import pytest
from scipy import sparse, integrate

from oracle import rb_oracle as orc
from redbirdpy_amd.mesh import kuhn_box
from conftest import box_cfg, grid_optodes


# ---------------------------------------------------------------------------
# mesh integer outputs
# ---------------------------------------------------------------------------
def _kuhn_counts(nx, ny, nz):
    nn = nx * ny * nz
    ne = 6 * (nx - 1) * (ny - 1) * (nz - 1)
    nf = 4 * ((nx - 1) * (ny - 1) + (nx - 1) * (nz - 1) + (ny - 1) * (nz - 1))
    # edges along the 7 lattice directions of the Kuhn subdivision
    E = (nx - 1) * ny * nz + nx * (ny - 1) * nz + nx * ny * (nz - 1)
    E += (nx - 1) * (ny - 1) * nz + (nx - 1) * ny * (nz - 1) + nx * (ny - 1) * (nz - 1)
    E += (nx - 1) * (ny - 1) * (nz - 1)
    return nn, ne, nf, nn + 2 * E


@pytest.mark.parametrize("dims,h", [((4, 6, 2), 1.0), ((60, 60, 30), 2.0)])
def test_kuhn_closed_forms(dims, h):
    node, elem = kuhn_box([0, 0, 0], dims, h)
    nx, ny, nz = [int(round(d / h)) + 1 for d in dims]
    nn, ne, nf, nnz = _kuhn_counts(nx, ny, nz)
    cfg, _ = orc.meshprep({"node": node, "elem": elem, "srcpos": [[1, 1, 0]], "srcdir": [[0, 0, 1]],
                           "detpos": [[2, 2, 0]], "prop": np.array([[0, 0, 1, 1], [0.01, 1, 0, 1.37]]),
                           "seg": np.ones(len(elem), int)})
    assert cfg["node"].shape[0] == nn and cfg["elem"].shape[0] == ne
    assert cfg["face"].shape[0] == nf
    assert len(cfg["rows"]) + nn == nnz
    assert np.all(cfg["evol"] > 0)
    np.testing.assert_allclose(cfg["evol"], h ** 3 / 6, rtol=1e-12)
    np.testing.assert_allclose(cfg["area"], h ** 2 / 2, rtol=1e-12)
    assert cfg["evol"].sum() == pytest.approx(np.prod(dims), rel=1e-12)
    assert cfg["nvol"].sum() == pytest.approx(np.prod(dims), rel=1e-12)
    # all faces point outward: (face centroid - box centre) . normal > 0
    f0 = cfg["face"] - 1
    p = cfg["node"][f0]
    nrm = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    cen = p.mean(axis=1) - np.array(dims) / 2
    assert np.all(np.sum(nrm * cen, axis=1) > 0)


def test_volface_single_and_pair():
    node = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1]], float)
    one = np.array([[1, 2, 3, 4]])
    f = orc.volface(orc.meshreorient(node, one))
    assert f.shape == (4, 3)
    two = orc.meshreorient(node, np.array([[1, 2, 3, 4], [2, 3, 4, 5]]))
    f2 = orc.volface(two)
    assert f2.shape == (6, 3)  # the shared face (2,3,4) is interior
    assert not any(sorted(x) == [2, 3, 4] for x in f2.tolist())
    # lexicographic order of sorted triples
    keys = [tuple(sorted(x)) for x in f2.tolist()]
    assert keys == sorted(keys)


def test_reorient_fixes_negative_volume():
    node = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    e = np.array([[2, 1, 3, 4]])
    r = orc.meshreorient(node, e)
    assert r.tolist() == [[1, 2, 3, 4]]


def test_femnz_brute_force():
    node, elem = kuhn_box([0, 0, 0], (3, 2, 2), 1.0)
    elem = orc.meshreorient(node, elem)
    rows, cols, cnt = orc.femnz(elem, node.shape[0])
    ref = set()
    for t in elem - 1:
        for i in t:
            for j in t:
                if i != j:
                    ref.add((i, j))
    assert set(zip(rows.tolist(), cols.tolist())) == ref
    assert cnt.sum() == len(ref)


# ---------------------------------------------------------------------------
# deldotdel and femlhs against brute-force integrals
# ---------------------------------------------------------------------------
def test_deldotdel_kuhn_closed_form():
    h = 2.0
    node, elem = kuhn_box([0, 0, 0], (2, 2, 2), h)
    elem = orc.meshreorient(node, elem)
    cfg = {"node": node, "elem": elem, "evol": orc.elemvolume(node, elem)}
    dd, _ = orc.deldotdel(cfg)
    # every Kuhn path tet has V grad.grad = (h/6) * a permutation of the path
    # stencil; its multiset of entries is fixed
    for row in dd * 6 / h:
        assert sorted(np.round(row, 12)) == sorted([1, -1, 0, 0, 2, -1, 0, 2, -1, 1])


def _p1_grads(P):
    """Gradients of the 4 barycentric functions from the inverse vertex matrix."""
    M = np.column_stack([np.ones(4), P])
    return np.linalg.inv(M)[1:, :].T  # (4, 3)


# degree-3 exact rule on the reference tet (Keast, 5 points), weights sum to 1
_KEAST = [(-0.8, (0.25, 0.25, 0.25, 0.25))] + [
    (0.45, tuple(0.5 if j == i else 1.0 / 6 for j in range(4))) for i in range(4)
]


def _brute_femlhs(cfg, props_node=None, omega=0.0):
    """A = sum_e [D_e V grad.grad + int mua phi phi + i w/c int n phi phi] + Robin, by quadrature."""
    node = cfg["node"]
    e0 = cfg["elem"][:, :4] - 1
    nn = node.shape[0]
    A = np.zeros((nn, nn), dtype=complex)
    reff = cfg["reff"]
    seg = cfg.get("seg")
    for e, t in enumerate(e0):
        P = node[t]
        V = abs(np.linalg.det(P[1:] - P[0])) / 6
        G = _p1_grads(P)
        if props_node is None:
            pr = cfg["prop"][seg[e]] if cfg["prop"].shape[0] < len(e0) else cfg["prop"][e]
            mu_q = lambda lam: pr[0]
            n_q = lambda lam: cfg["prop"][seg[0]][3] if cfg["prop"].shape[0] < len(e0) else pr[3]
            D = 1 / (3 * (pr[0] + pr[1] * (1 - pr[2])))
        else:
            pn = props_node[t]
            mu_q = lambda lam: lam @ pn[:, 0]
            n_q = lambda lam: lam @ pn[:, 3]
            D = np.mean(1 / (3 * (pn[:, 0] + pn[:, 1] * (1 - pn[:, 2]))))
        for a in range(4):
            for b in range(4):
                mass = sum(w * mu_q(np.array(l)) * l[a] * l[b] for w, l in _KEAST) * V
                fdm = sum(w * n_q(np.array(l)) * l[a] * l[b] for w, l in _KEAST) * V
                A[t[a], t[b]] += D * V * (G[a] @ G[b]) + mass + 1j * omega / orc.C0 * fdm
    c = (1 - reff) / (2 * (1 + reff))
    tri = [((2 / 3, 1 / 6, 1 / 6), 1 / 3), ((1 / 6, 2 / 3, 1 / 6), 1 / 3), ((1 / 6, 1 / 6, 2 / 3), 1 / 3)]
    for f in cfg["face"] - 1:
        P = node[f]
        area = 0.5 * np.linalg.norm(np.cross(P[1] - P[0], P[2] - P[0]))
        for a in range(3):
            for b in range(3):
                A[f[a], f[b]] += c * area * sum(w * l[a] * l[b] for l, w in tri)
    return A


def _small_mesh_cfg(seed=0, omega=0.0):
    rng = np.random.default_rng(seed)
    node, elem = kuhn_box([0, 0, 0], (3, 2, 2), 1.0)
    node = node + rng.uniform(-0.15, 0.15, node.shape)  # distorted tets
    cfg = {"node": node, "elem": elem, "srcpos": [[1, 1, 0]], "srcdir": [[0, 0, 1]], "detpos": [[2, 1, 0]],
           "prop": np.array([[0, 0, 1, 1], [0.02, 1.1, 0.1, 1.4], [0.05, 0.7, 0, 1.33]]),
           "seg": rng.integers(1, 3, len(elem)), "omega": omega}
    cfg, _ = orc.meshprep(cfg)
    return cfg


@pytest.mark.parametrize("omega", [0.0, 2 * np.pi * 1e8])
def test_femlhs_label_branch_vs_quadrature(omega):
    cfg = _small_mesh_cfg(1, omega)
    A = orc.femlhs(cfg, cfg["deldotdel"]).toarray()
    Ab = _brute_femlhs(cfg, omega=omega)
    assert np.max(np.abs(A - Ab)) < 1e-13 * np.max(np.abs(Ab))


@pytest.mark.parametrize("omega", [0.0, 2 * np.pi * 1e8])
def test_femlhs_node_branch_vs_quadrature(omega):
    cfg = _small_mesh_cfg(2, omega)
    rng = np.random.default_rng(3)
    nn = cfg["node"].shape[0]
    pn = np.column_stack([rng.uniform(0.005, 0.05, nn), rng.uniform(0.5, 1.5, nn), rng.uniform(0, 0.2, nn),
                          rng.uniform(1.3, 1.5, nn)])
    cfg["prop"] = pn
    cfg.pop("seg")
    A = orc.femlhs(cfg, cfg["deldotdel"]).toarray()
    Ab = _brute_femlhs(cfg, props_node=pn, omega=omega)
    assert np.max(np.abs(A - Ab)) < 1e-13 * np.max(np.abs(Ab))


def test_femlhs_elem_branch_vs_quadrature():
    cfg = _small_mesh_cfg(4)
    rng = np.random.default_rng(5)
    ne = cfg["elem"].shape[0]
    cfg["prop"] = np.column_stack([rng.uniform(0.005, 0.05, ne), rng.uniform(0.5, 1.5, ne), np.zeros(ne),
                                   rng.uniform(1.3, 1.5, ne)])
    A = orc.femlhs(cfg, cfg["deldotdel"]).toarray()
    Ab = _brute_femlhs(cfg)
    assert np.max(np.abs(A - Ab)) < 1e-13 * np.max(np.abs(Ab))


def test_femlhs_stiffness_is_7point_stencil():
    h = 2.0
    cfg = box_cfg((8, 8, 8), h, prop=np.array([[0, 0, 1, 1], [0.0, 1.0, 0, 1.0]]))
    cfg, _ = orc.meshprep(cfg)
    cfg["reff"] = 1.0  # kills the Robin term
    A = orc.femlhs(cfg, cfg["deldotdel"])
    D = 1 / 3.0
    nx = 5
    centre = 2 + nx * (2 + nx * 2)
    row = A.getrow(centre).toarray().ravel()
    nz = {int(j): row[j] for j in np.flatnonzero(np.abs(row) > 1e-14)}
    assert nz[centre] == pytest.approx(6 * D * h)
    for off in (1, nx, nx * nx):
        assert nz[centre + off] == pytest.approx(-D * h)
        assert nz[centre - off] == pytest.approx(-D * h)
    assert len(nz) == 7
    assert abs(A.sum(axis=1)).max() < 1e-12  # stiffness rows sum to zero


# ---------------------------------------------------------------------------
# solve, readout, Jacobian
# ---------------------------------------------------------------------------
def _fd_cfg(omega=0.0):
    p1 = (16, 16, 8)
    cfg = box_cfg(p1, 2.0, src=grid_optodes(p1, 2), det=grid_optodes(p1, 2, margin=3.0), omega=omega)
    return orc.meshprep(cfg)


@pytest.mark.parametrize("omega", [0.0, 2 * np.pi * 1e8])
def test_femsolve_residual_and_reciprocity(omega):
    cfg, sd = _fd_cfg(omega)
    A = orc.femlhs(cfg, cfg["deldotdel"])
    # symmetric up to summation order (complex symmetric, not Hermitian)
    assert abs(A - A.T).max() <= 1e-15 * abs(A).max()
    rhs, loc, bary, _ = orc.femrhs(cfg, sd)
    x, flag = orc.femsolve(A, rhs)
    r = A @ x - rhs.toarray()
    assert np.linalg.norm(r) / np.linalg.norm(rhs.toarray()) < 1e-12
    det = orc.femgetdet(x, cfg, rhs, loc, bary)
    ns = 4
    # detval[d, s] = rhs_d^T A^-1 rhs_s = phi_s at detector d = phi_d at source s
    R = rhs.toarray()
    recip = R[:, :ns].T @ x[:, ns:]  # (src x det): rhs_s^T phi_d
    np.testing.assert_allclose(det, recip.T, rtol=1e-10)
    if omega == 0:
        assert np.all(det > 0)


def test_bary_and_tsearch():
    cfg, sd = _fd_cfg()
    rhs, loc, bary, opt = orc.femrhs(cfg, sd)
    assert np.all(~np.isnan(loc))
    np.testing.assert_allclose(bary.sum(axis=1), 1.0, rtol=0, atol=1e-12)
    assert np.all(bary >= -1e-10)
    e0 = cfg["elem"][:, :4] - 1
    for i in range(len(loc)):
        p = bary[i] @ cfg["node"][e0[int(loc[i]) - 1]]
        np.testing.assert_allclose(p, opt[i], atol=1e-12)


def test_jac_matches_finite_differences():
    """Adjoint Jacobian == central differences of detval under the jac model.

    The model: element mua is the mean of nodal mua and D is held fixed, so
    bumping node n by delta adds delta/4 to mua of every tet holding n and
    removes delta/4 from mus' there (mua + mus' unchanged).
    """
    cfg, sd = _fd_cfg()
    ne = cfg["elem"].shape[0]
    base = np.tile(np.array([0.01, 1.0, 0.0, 1.37]), (ne, 1))
    cfg["prop"] = base
    A = orc.femlhs(cfg, cfg["deldotdel"])
    rhs, loc, bary, _ = orc.femrhs(cfg, sd)
    phi, _ = orc.femsolve(A, rhs)
    Jn, Je = orc.jac(sd, phi, cfg["deldotdel"], cfg["elem"], cfg["evol"])
    assert np.all(Jn <= 1e-10)  # CW Jacobian is non-positive (test_integration.py:236-244)
    np.testing.assert_allclose(Jn.sum(axis=1), Je.sum(axis=1), rtol=1e-12)
    rng = np.random.default_rng(3)
    e0 = cfg["elem"][:, :4] - 1
    delta = 1e-4 * 0.01  # central differences: O(delta^2) truncation, roundoff ~ eps*max|det|/delta
    for n in rng.choice(cfg["node"].shape[0], 6, replace=False):
        hold = np.any(e0 == n, axis=1)
        vals = []
        for sgn in (+1, -1):
            p = base.copy()
            p[hold, 0] += sgn * delta / 4
            p[hold, 1] -= sgn * delta / 4
            cfg["prop"] = p
            Ax = orc.femlhs(cfg, cfg["deldotdel"])
            xx, _ = orc.femsolve(Ax, rhs)
            vals.append(orc.femgetdet(xx, cfg, rhs, loc, bary).ravel())
        fd = (vals[0] - vals[1]) / (2 * delta)
        np.testing.assert_allclose(Jn[:, n], fd, rtol=1e-6, atol=1e-7 * np.abs(Jn).max())


def test_jac_is_mass_matrix_form():
    """Jelem[p, e] = -phi_s^T M_e phi_d with the exact P1 element mass V(1+delta)/20."""
    cfg, sd = _fd_cfg()
    A = orc.femlhs(cfg, cfg["deldotdel"])
    rhs, _, _, _ = orc.femrhs(cfg, sd)
    phi, _ = orc.femsolve(A, rhs)
    _, Je = orc.jac(sd, phi, cfg["deldotdel"], cfg["elem"], cfg["evol"])
    e0 = cfg["elem"][:, :4] - 1
    M = (np.ones((4, 4)) + np.eye(4)) / 20
    for e in (0, 17, 301):
        for p in (0, 5, 11):
            s, d = int(sd[p, 0]), int(sd[p, 1])
            v = -cfg["evol"][e] * phi[e0[e], s] @ M @ phi[e0[e], d]
            assert Je[p, e] == pytest.approx(v, rel=1e-12)


# ---------------------------------------------------------------------------
# regularised inverses, reflection, extinction, sd map
# ---------------------------------------------------------------------------
def test_reginv_matches_lstsq():
    rng = np.random.default_rng(9)
    for m, n in ((6, 20), (25, 8)):
        A = rng.standard_normal((m, n))
        b = rng.standard_normal(m)
        lam = 0.3
        x = orc.reginv(A, b, lam)
        aug = np.vstack([A, np.sqrt(lam) * np.eye(n)])
        xl = np.linalg.lstsq(aug, np.concatenate([b, np.zeros(n)]), rcond=None)[0]
        np.testing.assert_allclose(x, xl, rtol=1e-10, atol=1e-12)


def test_matreform_reim_is_the_complex_tikhonov_update():
    """'reim' rows [[Re, -Im], [Im, Re]] are the real form of the complex map z -> J z, so the
    underdetermined step x = A^T (A A^T + lam I)^-1 r is [Re z; Im z] with
    z = J^H (J J^H + lam I)^-1 r; 'real' rows ([Re; Im], real unknown) give another step."""
    rng = np.random.default_rng(12)
    m, n = 7, 30
    J = rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))
    y0 = rng.standard_normal(m) + 1j * rng.standard_normal(m)
    y = rng.standard_normal(m) + 1j * rng.standard_normal(m)
    lam = 0.2
    A, rhs, nb = orc.matreform(J, y0, y, "reim")
    assert nb == 2 and A.shape == (2 * m, 2 * n)
    x = orc.reginv(A, rhs, lam)
    r = y0 - y
    z = J.conj().T @ np.linalg.solve(J @ J.conj().T + lam * np.eye(m), r)
    np.testing.assert_allclose(x[:n], z.real, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(x[n:], z.imag, rtol=1e-10, atol=1e-12)
    Ar, rr, _ = orc.matreform(J, y0, y, "real")
    assert Ar.shape == (2 * m, n)
    assert not np.allclose(orc.reginv(Ar, rr, lam), z.real, rtol=1e-3)


def test_reginv_with_priors_closed_forms():
    """Overdetermined: (A^T A + lam L^T L) x = A^T b is the least-squares solution of
    [A; sqrt(lam) L] x = [b; 0] (an L^T L of one block's size repeats down the diagonal).
    Underdetermined with R^-1 from L = QR: x = P A^T (A P A^T + lam I)^-1 b, P = (L^T L)^-1,
    applied per unknown block of R^-1's size."""
    from scipy.linalg import qr

    rng = np.random.default_rng(21)
    m, n = 40, 6
    A = rng.standard_normal((m, n))
    b = rng.standard_normal(m)
    L = np.eye(n) + 0.3 * rng.standard_normal((n, n))
    lam = 0.5
    x = orc.reginvover(A, b, lam, L.T @ L)
    xl = np.linalg.lstsq(np.vstack([A, np.sqrt(lam) * L]), np.concatenate([b, np.zeros(n)]), rcond=None)[0]
    np.testing.assert_allclose(x, xl, rtol=1e-10, atol=1e-12)
    Lb = np.eye(3) + 0.3 * rng.standard_normal((3, 3))
    x = orc.reginvover(A, b, lam, Lb.T @ Lb)
    Lf = np.kron(np.eye(2), Lb)
    xl = np.linalg.lstsq(np.vstack([A, np.sqrt(lam) * Lf]), np.concatenate([b, np.zeros(n)]), rcond=None)[0]
    np.testing.assert_allclose(x, xl, rtol=1e-10, atol=1e-12)
    # underdetermined
    m, n = 5, 12
    A = rng.standard_normal((m, n))
    b = rng.standard_normal(m)
    L = np.eye(n) + 0.2 * rng.standard_normal((n, n))
    _, R = qr(L)
    x = orc.reginvunder(A, b, lam, np.linalg.inv(np.triu(R)))
    P = np.linalg.inv(L.T @ L)
    np.testing.assert_allclose(x, P @ A.T @ np.linalg.solve(A @ P @ A.T + lam * np.eye(m), b), rtol=1e-9)
    Lb = np.eye(4) + 0.2 * rng.standard_normal((4, 4))
    _, R = qr(Lb)
    blocks = {"a": (m, 4), "b": (m, 4), "c": (m, 4)}
    x = orc.reginvunder(A, b, lam, np.linalg.inv(np.triu(R)), blocks)
    P = np.kron(np.eye(3), np.linalg.inv(Lb.T @ Lb))
    np.testing.assert_allclose(x, P @ A.T @ np.linalg.solve(A @ P @ A.T + lam * np.eye(m), b), rtol=1e-9)


def test_masksum_is_the_adjoint_of_label_expansion():
    """masksum(J, seg) v = J (v expanded to nodes by label) for every label vector v."""
    rng = np.random.default_rng(22)
    seg = rng.integers(3, 9, size=40)
    J = rng.standard_normal((7, 40))
    labs = np.unique(seg)
    Jc = orc.masksum(J, seg)
    assert Jc.shape == (7, len(labs))
    for _ in range(3):
        v = rng.standard_normal(len(labs))
        np.testing.assert_allclose(Jc @ v, J @ v[np.searchsorted(labs, seg)], rtol=1e-12)
    np.testing.assert_allclose(Jc.sum(axis=1), J.sum(axis=1), rtol=1e-12)


def test_getreff_against_quadrature():
    def exact(n):
        oc = np.arcsin(1 / n)

        def R(t):
            if t >= oc:
                return 1.0
            ci, ct = np.cos(t), np.sqrt(1 - (n * np.sin(t)) ** 2)
            return 0.5 * ((n * ct - ci) / (n * ct + ci)) ** 2 + 0.5 * ((n * ci - ct) / (n * ci + ct)) ** 2

        rp = 2 * integrate.quad(lambda t: np.sin(t) * np.cos(t) * R(t), 0, np.pi / 2, points=[oc])[0]
        rj = 3 * integrate.quad(lambda t: np.sin(t) * np.cos(t) ** 2 * R(t), 0, np.pi / 2, points=[oc])[0]
        return (rp + rj) / (2 - rp + rj)

    for n in (1.33, 1.37, 1.45):
        assert orc.getreff(n, 1.0) == pytest.approx(exact(n), abs=3e-3)
    assert orc.getreff(1.0, 1.0) == 0.0


def test_extinction_table_entries_and_interpolation():
    ex, tab = orc.extinction([800, 850, 805, 1050, 240], ["hbo", "hbr"])
    c = 2.303e-7
    assert ex[0, 0] == pytest.approx(816 * c) and ex[0, 1] == pytest.approx(761.72 * c)
    assert ex[1, 0] == pytest.approx(1058 * c) and ex[1, 1] == pytest.approx(691.32 * c)
    assert ex[2, 0] == pytest.approx(0.5 * (816 + 864) * c)
    assert ex[3, 0] == pytest.approx((1024 + 5 * (1024 - 1080)) * c)  # extrapolated
    assert ex[4, 1] == pytest.approx((112736 - (116296 - 112736)) * c)
    w, _ = orc.extinction([750], ["water", "lipids"])
    assert w[0, 0] == pytest.approx(0.0026) and w[0, 1] == pytest.approx(0.0005)


def test_sdmap_layout():
    cfg = {"srcpos": np.zeros((3, 3)), "detpos": np.zeros((2, 3)), "prop": np.zeros((2, 4))}
    sd = orc.sdmap(cfg)
    assert sd.shape == (6, 3)
    for r, (s, d, a) in enumerate(sd):
        assert s == r % 3 and d == 3 + r // 3 and a == 1


def test_updateprop_and_jacchrome():
    cfg = {"node": np.zeros((60, 3)), "elem": np.ones((70, 4), int),
           "prop": {"690": np.array([[0, 0, 1, 1], [0.01, 1, 0, 1.37]]),
                    "830": np.array([[0, 0, 1, 1], [0.01, 1, 0, 1.37]])},
           "param": {"hbo": np.full(60, 20.0), "hbr": np.full(60, 10.0)}}
    p = orc.updateprop(cfg)
    ex, _ = orc.extinction(["690", "830"], ["hbo", "hbr", "water", "lipids"])
    for i, wv in enumerate(["690", "830"]):
        mua = ex[i] @ np.array([20, 10, 0.23, 0.58])
        np.testing.assert_allclose(p[wv][:, 0], mua, rtol=1e-14)
        np.testing.assert_allclose(p[wv][:, 1:], [[1, 0, 1.37]] * 60)
    J = {"690": np.ones((2, 5)), "830": 2 * np.ones((2, 5))}
    Jc = orc.jacchrome(J, ["hbo", "hbr"])
    e2, _ = orc.extinction(["690", "830"], ["hbo", "hbr"])
    np.testing.assert_allclose(Jc["hbr"][:2], e2[0, 1])
    np.testing.assert_allclose(Jc["hbr"][2:], 2 * e2[1, 1])


# ---------------------------------------------------------------------------
# dual mesh (meshinterp, _remap_jacobian) and the frequency-domain reforms
# ---------------------------------------------------------------------------
def _dual_meshes():
    fine, _ = orc.meshprep(box_cfg((12, 12, 6), 1.0))
    cnode, celem = kuhn_box([0, 0, 0], [12, 12, 6], 3.0)
    loc, bary = orc.tsearchn(cnode, celem, fine["node"])
    return fine, cnode, celem, loc, bary


def test_meshinterp_reproduces_linear_fields():
    fine, cnode, celem, loc, bary = _dual_meshes()
    assert not np.any(np.isnan(loc))
    coef = np.array([[0.3, -1.2], [0.7, 0.1], [-0.4, 2.0]])
    v = cnode @ coef + np.array([1.5, -0.25])
    got = orc.meshinterp(v, loc, bary, celem)
    np.testing.assert_allclose(got, fine["node"] @ coef + np.array([1.5, -0.25]), rtol=1e-12, atol=1e-12)
    # points outside the source mesh keep toval (or 0)
    loc2 = loc.copy()
    loc2[:5] = np.nan
    got2 = orc.meshinterp(v[:, 0], loc2, bary, celem, toval=np.full(len(loc), 7.0))
    assert np.all(got2[:5] == 7.0)


def test_remap_jacobian_is_adjoint_of_meshinterp():
    fine, cnode, celem, loc, bary = _dual_meshes()
    rng = np.random.default_rng(11)
    J = rng.standard_normal((5, fine["node"].shape[0]))
    recon = {"node": cnode, "elem": celem, "mapid": loc, "mapweight": bary}
    Jr = orc.remap_jacobian(J, recon, fine)
    assert Jr.shape == (5, cnode.shape[0])
    for _ in range(3):
        v = rng.standard_normal(cnode.shape[0])
        np.testing.assert_allclose(Jr @ v, J @ orc.meshinterp(v, loc, bary, celem), rtol=1e-12)
    # a linear field's Jacobian sum: J_rec applied to ones equals J applied to ones
    np.testing.assert_allclose(Jr.sum(1), J.sum(1), rtol=1e-12)


def test_fd_jacobian_and_logphase_rows_match_finite_differences():
    """Complex Jacobian (FD) and matreform('logphase') rows vs central differences of
    detval and of (log|detval|, angle(detval)) under the jac derivative model."""
    cfg, sd = _fd_cfg(omega=2 * np.pi * 1e8)
    ne = cfg["elem"].shape[0]
    base = np.tile(np.array([0.01, 1.0, 0.0, 1.37]), (ne, 1))
    cfg["prop"] = base
    A = orc.femlhs(cfg, cfg["deldotdel"])
    rhs, loc, bary, _ = orc.femrhs(cfg, sd)
    phi, _ = orc.femsolve(A, rhs)
    y = orc.femgetdet(phi, cfg, rhs, loc, bary).ravel()
    Jn, _ = orc.jac(sd, phi, cfg["deldotdel"], cfg["elem"], cfg["evol"])
    assert np.iscomplexobj(Jn)
    Alp, rlp, nb = orc.matreform(Jn, y * 1.01, y, "logphase")
    assert nb == 2 and Alp.shape == (2 * len(y), Jn.shape[1])
    np.testing.assert_allclose(rlp[: len(y)], np.log(1.01), rtol=1e-12)
    np.testing.assert_allclose(rlp[len(y):], 0, atol=1e-12)
    Are, _, nb2 = orc.matreform(Jn, y * 1.01, y, "real")
    assert nb2 == 2
    e0 = cfg["elem"][:, :4] - 1
    delta = 1e-4 * 0.01
    for n in np.random.default_rng(5).choice(cfg["node"].shape[0], 4, replace=False):
        hold = np.any(e0 == n, axis=1)
        vals = []
        for sgn in (+1, -1):
            p = base.copy()
            p[hold, 0] += sgn * delta / 4
            p[hold, 1] -= sgn * delta / 4
            cfg["prop"] = p
            xx, _ = orc.femsolve(orc.femlhs(cfg, cfg["deldotdel"]), rhs)
            vals.append(orc.femgetdet(xx, cfg, rhs, loc, bary).ravel())
        fd = (vals[0] - vals[1]) / (2 * delta)
        np.testing.assert_allclose(Jn[:, n], fd, rtol=1e-6, atol=1e-7 * np.abs(Jn).max())
        np.testing.assert_allclose(Are[:, n], np.concatenate([fd.real, fd.imag]), rtol=1e-6,
                                   atol=1e-7 * np.abs(Jn).max())
        fla = (np.log(np.abs(vals[0])) - np.log(np.abs(vals[1]))) / (2 * delta)
        fph = (np.angle(vals[0]) - np.angle(vals[1])) / (2 * delta)
        np.testing.assert_allclose(Alp[:, n], np.concatenate([fla, fph]), rtol=1e-5,
                                   atol=1e-6 * np.abs(Alp).max())
