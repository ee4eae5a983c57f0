"""Pin the oracle's reconstruction and bookkeeping functions by what the mathematics fixes (no GPU).

Gauss-Newton step: the update of one runrecon iteration equals the textbook
Tikhonov step s * argmin ||s J x - r||^2 + lambda ||x||^2 (numpy lstsq on the
augmented system) with s = 1/||J||_F, J built from the separately pinned
jac (finite differences), extinction (table entries) and the measurement
order of _flatten_detphi (known answers below); node, label and chromophore
branches.  A plausible slip -- the blockscale applied once instead of twice,
an update added to the wrong rows, a transposed Kronecker block, a wrong
measurement order -- moves the update by O(1) and fails.

The product's host helpers with the same contract are checked against the
same known answers.
"""

import copy

import numpy as np
import pytest

from oracle import rb_oracle as orc
from redbirdpy_amd import recon as prec, utility as putil, property as pprop
from redbirdpy_amd.mesh import kuhn_box


# ---------------------------------------------------------------------------
# known answers for the small helpers
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("fn", [orc.normalize_diag, prec._normalize_diag])
def test_normalize_diag_known_answer(fn):
    """The reference's own known answer (test/test_recon.py:437-444)."""
    A = np.diag([4.0, 9.0, 16.0])
    An, di = fn(A)
    np.testing.assert_allclose(np.diag(An), [1, 1, 1])
    np.testing.assert_allclose(di, [0.5, 1 / 3, 0.25])
    B = np.array([[4.0, 2.0], [2.0, 9.0]])
    Bn, _ = fn(B)
    np.testing.assert_allclose(Bn, [[1.0, 2.0 / 6.0], [2.0 / 6.0, 1.0]], rtol=1e-15)


@pytest.mark.parametrize("fn", [orc.flatten_detphi, prec._flatten_detphi])
def test_flatten_detphi_order(fn):
    """Measurements flatten wavelength-major, then detector-major (row d*Ns + s of a
    (Ndet x Nsrc) block) -- the row order of sdmap, so misfit row r meets Jacobian row r."""
    Ns, Nd = 3, 2
    wvs = ["690", "750"]
    det = {w: np.array([[1000 * i + 10 * d + s for s in range(Ns)] for d in range(Nd)], float)
           for i, w in enumerate(wvs)}
    y = fn(det, None, wvs, [1])
    want = [1000 * i + 10 * d + s for i in range(2) for d in range(Nd) for s in range(Ns)]
    np.testing.assert_array_equal(y, want)
    sd = orc.sdmap({"srcpos": np.zeros((Ns, 3)), "detpos": np.zeros((Nd, 3)), "prop": np.zeros((2, 4))})
    for r in range(Ns * Nd):
        s, d = int(sd[r, 0]), int(sd[r, 1]) - Ns
        assert y[r] == 10 * d + s
    # one wavelength, plain array; and the mode nesting of a runforward with rfcw=[1, 2]
    np.testing.assert_array_equal(fn(det["690"], None, [""], [1]), want[:6])
    nested = {1: {"detphi": det["690"]}, 2: {"detphi": det["750"]}}
    np.testing.assert_array_equal(fn(nested, None, [""], [1, 2]), want)


@pytest.mark.parametrize("fn", [orc.matflat, prec.matflat])
def test_matflat_order(fn):
    a = np.arange(6.0).reshape(2, 3)
    b = 10 + np.arange(6.0).reshape(2, 3)
    np.testing.assert_array_equal(fn({"hbo": a, "hbr": b}), np.hstack([a, b]))
    np.testing.assert_array_equal(fn({"hbo": a, "hbr": b}, np.array([2.0, -1.0])), np.hstack([2 * a, -b]))
    nested = {"hbo": {"690": a, "750": 2 * a}, "hbr": {"690": b, "750": 3 * b}}
    np.testing.assert_array_equal(fn(nested), np.vstack([np.hstack([a, b]), np.hstack([2 * a, 3 * b])]))
    np.testing.assert_array_equal(fn(a), a)


# ---------------------------------------------------------------------------
# bulk properties and optode displacement
# ---------------------------------------------------------------------------
def _small_box():
    """A 27-node, 48-tet box; face[0, 0] is set to corner node 0 (1-based 1), held by a few tets."""
    node, elem = kuhn_box([0, 0, 0], [2, 2, 2], 1.0)
    holders = np.flatnonzero(np.any(elem[:, :4] == 1, axis=1))
    return node, elem, np.array([[1, 2, 4]]), holders


@pytest.mark.parametrize("getbulk", [orc.getbulk, pprop.getbulk])
def test_getbulk_branches(getbulk):
    node, elem, face, holders = _small_box()
    ne, nn = elem.shape[0], node.shape[0]
    table = np.array([[0, 0, 1, 1], [0.01, 1.0, 0, 1.37], [0.02, 5.0, 0.8, 1.4]])
    seg = np.ones(ne, int)
    seg[holders[0]] = 2  # the label of the first element holding face[0,0] decides
    cfg = {"node": node, "elem": elem, "face": face, "prop": table, "seg": seg}
    np.testing.assert_array_equal(getbulk(cfg), table[2])
    seg = np.full(ne, 2)
    seg[holders[0]] = 1
    cfg["seg"] = seg
    np.testing.assert_array_equal(getbulk(cfg), table[1])
    nseg = np.ones(nn, int)
    nseg[0] = 2  # node labels: the label of node face[0,0]
    cfg["seg"] = nseg
    np.testing.assert_array_equal(getbulk(cfg), table[2])
    nodeprop = np.column_stack([0.01 + np.arange(nn) * 0.001, np.ones(nn), np.zeros(nn), np.full(nn, 1.37)])
    cfg = {"node": node, "elem": elem, "face": np.array([[5, 2, 3]]), "prop": nodeprop}
    np.testing.assert_array_equal(getbulk(cfg), nodeprop[4])
    elemprop = np.column_stack([0.01 + np.arange(ne) * 0.001, np.ones(ne), np.zeros(ne), np.full(ne, 1.37)])
    cfg = {"node": node, "elem": elem, "face": face, "prop": elemprop}
    np.testing.assert_array_equal(getbulk(cfg), elemprop[holders[0]])
    cfg = {"node": node, "elem": elem, "prop": table, "bulk": {"mua": 0.02, "dcoeff": 0.5, "n": 1.4}}
    np.testing.assert_allclose(getbulk(cfg), [0.02, 1 / 1.5, 0.0, 1.4])


@pytest.mark.parametrize("mod", [orc, putil])
def test_getoptodes_shift_closed_form(mod):
    """Point optodes move by dir * l_tr, l_tr = 1/(mua + mus (1 - g)) of the bulk medium."""
    node, elem, face, holders = _small_box()
    table = np.array([[0, 0, 1, 1], [0.01, 1.0, 0, 1.37], [0.02, 5.0, 0.8, 1.4]])
    seg = np.ones(elem.shape[0], int)
    seg[holders[0]] = 2
    cfg = {"node": node, "elem": elem, "face": face, "prop": table, "seg": seg,
           "srcpos": np.array([[0.2, 0.2, 0.0], [0.3, 0.1, 0.0]]), "srcdir": np.array([[0.0, 0.0, 1.0]]),
           "detpos": np.array([[0.1, 0.1, 0.0]]), "detdir": np.array([[0.6, 0.0, 0.8]])}
    ltr = 1.0 / (0.02 + 5.0 * 0.2)
    src, det, ws, wd = mod.getoptodes(cfg)
    np.testing.assert_allclose(src, cfg["srcpos"] + [0, 0, ltr], rtol=1e-15)
    np.testing.assert_allclose(det, cfg["detpos"] + ltr * np.array([0.6, 0.0, 0.8]), rtol=1e-15)
    assert ws is None and wd is None
    assert mod.getltr(cfg) == pytest.approx(ltr, rel=1e-15)
    cfg["prop"] = {"690": table, "750": table * np.array([1, 0.5, 1, 1])}
    assert mod.getltr(cfg, "750") == pytest.approx(1.0 / (0.02 + 2.5 * 0.2), rel=1e-15)
    assert mod.getltr(cfg) == pytest.approx(ltr, rel=1e-15)  # first wavelength when none is named


def test_updateprop_branches_known_answers():
    """mua = sum_c E[c] C_c with the default water 0.23 / lipids 0.58 added; label params fill
    rows 1.. of the template; scatamp/scatpow give mus' = a * lambda^-b with g = 0."""
    node, elem = kuhn_box([0, 0, 0], [2, 2, 2], 1.0)
    tab = np.array([[0, 0, 1, 1], [0.01, 1, 0.5, 1.4], [0.02, 1, 0.5, 1.4]])
    for updateprop in (orc.updateprop, pprop.updateprop):
        cfg = {"node": node, "elem": elem, "prop": {"750": tab.copy(), "800": tab.copy()},
               "param": {"hbo": np.array([20.0, 30.0]), "hbr": np.array([10.0, 5.0]),
                         "scatamp": 2.0, "scatpow": 0.5}}
        p = updateprop(cfg)
        c = 2.303e-7
        e750 = np.array([518 * c, 1405.24 * c, 0.0026, 0.0005])
        for j, conc in enumerate([[20, 10, 0.23, 0.58], [30, 5, 0.23, 0.58]]):
            np.testing.assert_allclose(p["750"][1 + j, 0], e750 @ np.array(conc), rtol=1e-12)
            np.testing.assert_allclose(p["750"][1 + j, 1:3], [2.0 * 750 ** -0.5, 0.0], rtol=1e-14)
        np.testing.assert_array_equal(p["750"][0], tab[0])
        assert cfg["param"]["water"] == 0.23 and cfg["param"]["lipids"] == 0.58


# ---------------------------------------------------------------------------
# syncprop: one case per branch
# ---------------------------------------------------------------------------
def _dual():
    fine, _ = orc.meshprep(_box_cfg((12, 12, 6), 1.0))
    cnode, celem = kuhn_box([0, 0, 0], [12, 12, 6], 2.0)  # 196 recon nodes: node arrays, not labels
    loc, bary = orc.tsearchn(cnode, celem, fine["node"])
    return fine, cnode, celem, loc, bary


def _box_cfg(p1, h, omega=0.0):
    node, elem = kuhn_box([0, 0, 0], p1, h)
    return {"node": node, "elem": elem, "prop": np.array([[0, 0, 1, 1], [0.01, 1.0, 0, 1.37]]),
            "seg": np.ones(elem.shape[0], int),
            "srcpos": np.array([[2.0, 2.0, 0.0], [4.0, 4.0, 0.0]]), "srcdir": np.array([[0.0, 0.0, 1.0]]),
            "detpos": np.array([[2.0, 4.0, 0.0], [4.0, 2.0, 0.0]]), "detdir": np.array([[0.0, 0.0, 1.0]]),
            "omega": omega}


@pytest.mark.parametrize("syncprop", [orc.syncprop, prec.syncprop])
def test_syncprop_branches(syncprop):
    fine, cnode, celem, loc, bary = _dual()
    nn, nrec = fine["node"].shape[0], cnode.shape[0]
    lin = lambda X: 0.01 + 0.001 * X[:, 0] - 0.002 * X[:, 2]  # noqa: E731
    dual = {"node": cnode, "elem": celem, "mapid": loc, "mapweight": bary}
    # label params (< 50 entries): copied as they are, dual mesh or not
    cfg = copy.deepcopy(fine)
    r = dict(dual, param={"hbo": np.array([20.0, 30.0]), "hbr": np.array([1.0, 2.0])})
    syncprop(cfg, r)
    np.testing.assert_array_equal(cfg["param"]["hbo"], [20, 30])
    assert cfg["param"]["hbo"] is not r["param"]["hbo"]
    # node params on the recon mesh: interpolated to the forward nodes (a linear field exactly)
    cfg = copy.deepcopy(fine)
    r = dict(dual, param={"hbo": lin(cnode)})
    syncprop(cfg, r)
    np.testing.assert_allclose(cfg["param"]["hbo"], lin(fine["node"]), rtol=1e-12)
    # node params on the same mesh (no map): copied
    cfg = copy.deepcopy(fine)
    r = {"param": {"hbo": lin(fine["node"])}}
    syncprop(cfg, r)
    np.testing.assert_array_equal(cfg["param"]["hbo"], lin(fine["node"]))
    # node prop rows on the recon mesh: every column interpolated
    cfg = copy.deepcopy(fine)
    P = np.column_stack([lin(cnode), np.ones(nrec), np.zeros(nrec), np.full(nrec, 1.37)])
    syncprop(cfg, dict(dual, prop=P))
    assert cfg["prop"].shape == (nn, 4)
    np.testing.assert_allclose(cfg["prop"][:, 0], lin(fine["node"]), rtol=1e-12)
    np.testing.assert_allclose(cfg["prop"][:, 3], 1.37, rtol=1e-14)
    # label prop table with the dual-mesh keys present: copied, not interpolated
    cfg = copy.deepcopy(fine)
    T = np.array([[0.01, 1, 0, 1.37], [0.02, 1, 0, 1.37]])
    syncprop(cfg, dict(dual, prop=T))
    np.testing.assert_array_equal(cfg["prop"], T)
    # per-wavelength node props on the same mesh: copied per wavelength
    cfg = copy.deepcopy(fine)
    Pn = np.column_stack([lin(fine["node"]), np.ones(nn), np.zeros(nn), np.full(nn, 1.37)])
    syncprop(cfg, {"prop": {"690": Pn, "750": 2 * Pn}})
    np.testing.assert_array_equal(cfg["prop"]["750"], 2 * Pn)
    # per-wavelength node props on the recon mesh: interpolated per wavelength
    cfg = copy.deepcopy(fine)
    syncprop(cfg, dict(dual, prop={"690": P, "750": 2 * P}))
    np.testing.assert_allclose(cfg["prop"]["750"][:, 0], 2 * lin(fine["node"]), rtol=1e-12)


# ---------------------------------------------------------------------------
# runforward nesting
# ---------------------------------------------------------------------------
def test_runforward_nesting_two_wavelengths_two_modes():
    """rfcw=[1, 2] with two wavelengths: det[mode]['detphi'][wv]; mode 1 is the frequency-domain
    solve at omega, mode 2 the same wavelength with omega forced to 0 (forward.py:66-85,
    femlhs :123-124).  Each leaf equals an independent single-wavelength, single-mode run."""
    om = 2 * np.pi * 1e8
    base = _box_cfg((6, 6, 3), 1.0, omega=om)
    tabs = {"690": np.array([[0, 0, 1, 1], [0.01, 1.0, 0, 1.37]]),
            "750": np.array([[0, 0, 1, 1], [0.02, 0.8, 0, 1.37]])}
    cfg = copy.deepcopy(base)
    cfg["prop"] = {k: v.copy() for k, v in tabs.items()}
    cfg, sd = orc.meshprep(cfg)
    det, phi = orc.runforward(cfg, rfcw=[1, 2])
    assert set(det) == {1, 2} and set(det[1]["detphi"]) == {"690", "750"}
    for wv, tab in tabs.items():
        for md, omega in ((1, om), (2, 0.0)):
            one = copy.deepcopy(base)
            one["prop"] = tab.copy()
            one["omega"] = omega
            one, _ = orc.meshprep(one)
            d1, p1 = orc.runforward(one)
            np.testing.assert_allclose(det[md]["detphi"][wv], d1, rtol=1e-12)
            np.testing.assert_allclose(phi[md]["phi"][wv], p1, rtol=1e-10, atol=1e-14 * np.abs(p1).max())
            assert np.iscomplexobj(d1) == (md == 1)
    # one wavelength: the wavelength level is unwrapped; one mode: the mode level too
    one, _ = orc.meshprep(copy.deepcopy(base))
    d2, _ = orc.runforward(one, rfcw=[1, 2])
    assert isinstance(d2[2]["detphi"], np.ndarray) and d2[2]["detphi"].shape == (2, 2)
    d3, p3 = orc.runforward(one)
    assert isinstance(d3, np.ndarray) and d3.shape == (2, 2) and p3.shape == (one["node"].shape[0], 4)


# ---------------------------------------------------------------------------
# one Gauss-Newton iteration against the textbook Tikhonov step
# ---------------------------------------------------------------------------
def _tikhonov(J, r, lam):
    s = 1.0 / np.linalg.norm(J)
    n = J.shape[1]
    x = np.linalg.lstsq(np.vstack([s * J, np.sqrt(lam) * np.eye(n)]), np.concatenate([r, np.zeros(n)]),
                        rcond=None)[0]
    return s * x


def _spy_reginv(monkeypatch):
    seen = []
    real = orc.reginv

    def spy(A, rhs, lam, Areg=None, blocks=None):
        seen.append(np.array(A))
        return real(A, rhs, lam, Areg, blocks)

    monkeypatch.setattr(orc, "reginv", spy)
    return seen


def test_runrecon_node_step_is_tikhonov(monkeypatch):
    cfg0 = _box_cfg((6, 6, 3), 1.0)
    cfg, sd = orc.meshprep(copy.deepcopy(cfg0))
    nn = cfg["node"].shape[0]
    assert nn <= 200
    truth = np.column_stack([0.01 + 0.004 * (cfg["node"][:, 0] > 3), np.ones(nn), np.zeros(nn), np.full(nn, 1.37)])
    ct = copy.deepcopy(cfg)
    ct["prop"] = truth
    y0, _ = orc.runforward(ct)
    start = np.column_stack([np.full(nn, 0.01), np.ones(nn), np.zeros(nn), np.full(nn, 1.37)])
    # the expected step from the separately pinned parts
    ce = copy.deepcopy(cfg)
    ce["prop"] = start.copy()
    y, phi = orc.runforward(ce)
    J, _ = orc.jac(sd, phi, ce["deldotdel"], ce["elem"], ce["evol"])
    r = (y0 - y).ravel()
    lam = 0.05
    want = _tikhonov(J, r, lam)
    seen = _spy_reginv(monkeypatch)
    recon = {"prop": start.copy()}
    recon, resid, _, upd, Jm, _, _ = orc.runrecon(copy.deepcopy(cfg), recon, y0, sd, maxiter=1, lambda_=lam)
    assert J.shape[0] < J.shape[1]  # the underdetermined branch
    np.testing.assert_allclose(upd[0]["mua"], want, rtol=1e-8, atol=1e-12 * np.abs(want).max())
    np.testing.assert_allclose(recon["prop"][:, 0] - start[:, 0], want, rtol=1e-8,
                               atol=1e-12 * np.abs(want).max())
    np.testing.assert_array_equal(recon["prop"][:, 1:], start[:, 1:])
    assert resid[0] == pytest.approx(np.sum(np.abs(r)), rel=1e-12)
    A = seen[0]
    assert np.trace(A @ A.T) == pytest.approx(1.0, rel=1e-12)  # blockscale applied before the solve
    np.testing.assert_allclose(Jm["mua"], J, rtol=1e-12)


def test_runrecon_label_step_is_tikhonov(monkeypatch):
    """Label-compressed (overdetermined) step: J columns summed per node label, the update added to
    the label table's rows in label order (recon.py:210-218, :253-276)."""
    cfg0 = _box_cfg((6, 6, 3), 1.0)
    cfg, sd = orc.meshprep(copy.deepcopy(cfg0))
    nn = cfg["node"].shape[0]
    e0 = cfg["elem"][:, :4] - 1
    eseg = (cfg["node"][e0].mean(axis=1)[:, 0] > 3).astype(int)  # element labels 0 / 1
    nseg = (cfg["node"][:, 0] > 3).astype(int)  # node labels 0 / 1
    table = np.array([[0.01, 1.0, 0, 1.37], [0.01, 1.0, 0, 1.37]])
    ct = copy.deepcopy(cfg)
    ct["seg"] = eseg
    ct["prop"] = np.array([[0.01, 1.0, 0, 1.37], [0.016, 1.0, 0, 1.37]])
    y0, _ = orc.runforward(ct)
    ce = copy.deepcopy(cfg)
    ce["seg"] = eseg
    ce["prop"] = table.copy()
    y, phi = orc.runforward(ce)
    J, _ = orc.jac(sd, phi, ce["deldotdel"], ce["elem"], ce["evol"])
    S = (nseg[:, None] == np.array([0, 1])[None, :]).astype(float)
    Jl = J @ S
    lam = 1e-3
    want = _tikhonov(Jl, (y0 - y).ravel(), lam)
    seen = _spy_reginv(monkeypatch)
    c = copy.deepcopy(cfg)
    c["seg"] = eseg
    recon = {"prop": table.copy(), "seg": nseg}
    recon, _, _, upd, _, _, _ = orc.runrecon(c, recon, y0, sd, maxiter=1, lambda_=lam)
    assert seen[0].shape == (4, 2)
    np.testing.assert_allclose(upd[0]["mua"], want, rtol=1e-9)
    np.testing.assert_allclose(recon["prop"][:, 0] - table[:, 0], want, rtol=1e-9)
    assert want[1] > 0 and abs(want[1]) > 10 * abs(want[0])  # the step points at the perturbed label


def test_runrecon_chromophore_step_is_tikhonov_and_updates_whitelisted_keys(monkeypatch):
    """Multispectral step: rows (wavelength, pair), columns (chromophore, node) with block
    E[w, c] J_w; only hbo/hbr/water/lipids/scatamp/scatpow are updated (aa3 is not,
    recon.py:278-313)."""
    cfg0 = _box_cfg((6, 6, 3), 1.0)
    wvs = ["690", "750"]
    chs = ["hbo", "hbr", "aa3"]
    E = np.array([[2e-3, 1e-3, 5e-4], [1e-3, 3e-3, 1e-3]])
    extin = {w: {c: E[i, j] for j, c in enumerate(chs)} for i, w in enumerate(wvs)}
    cfg0["prop"] = {w: np.array([[0, 0, 1, 1], [0.01, 1.0, 0, 1.37]]) for w in wvs}
    cfg, sd = orc.meshprep(copy.deepcopy(cfg0))
    nn = cfg["node"].shape[0]
    bg = {"hbo": np.full(nn, 4.0), "hbr": np.full(nn, 2.0), "aa3": np.full(nn, 1.0)}
    tr = {k: v * (1 + 0.3 * (cfg["node"][:, 0] > 3)) for k, v in bg.items()}
    ct = copy.deepcopy(cfg)
    ct["param"] = copy.deepcopy(tr)
    ct["prop"] = orc.updateprop(ct, extin=extin)
    y0, _ = orc.runforward(ct)
    ce = copy.deepcopy(cfg)
    ce["param"] = copy.deepcopy(bg)
    ce["prop"] = orc.updateprop(ce, extin=extin)
    y, phi = orc.runforward(ce)
    Jw = [orc.jac(sd[w], phi[w], ce["deldotdel"], ce["elem"], ce["evol"])[0] for w in wvs]
    npair = Jw[0].shape[0]
    Jf = np.zeros((2 * npair, 3 * nn))
    for i in range(2):
        for j in range(3):
            Jf[i * npair : (i + 1) * npair, j * nn : (j + 1) * nn] = E[i, j] * Jw[i]
    r = np.concatenate([(y0[w] - y[w]).ravel() for w in wvs])
    lam = 1e-2
    want = _tikhonov(Jf, r, lam)
    _spy_reginv(monkeypatch)
    recon = {"param": copy.deepcopy(bg)}
    recon, _, _, upd, Jm, _, _ = orc.runrecon(copy.deepcopy(cfg), recon, y0, sd, maxiter=1, lambda_=lam,
                                              extin=extin)
    for j, c in enumerate(chs):
        np.testing.assert_allclose(upd[0][c], want[j * nn : (j + 1) * nn], rtol=1e-8,
                                   atol=1e-12 * np.abs(want).max())
    np.testing.assert_allclose(recon["param"]["hbo"] - bg["hbo"], want[:nn], rtol=1e-8,
                               atol=1e-12 * np.abs(want).max())
    np.testing.assert_array_equal(recon["param"]["aa3"], bg["aa3"])  # not in the update whitelist
    assert set(Jm) == set(chs) and Jm["hbr"].shape == (2 * npair, nn)


def test_runrecon_lmat_takes_precedence_over_ltl():
    """recon.py:127-133: an lmat keyword wins over an ltl keyword (if/elif)."""
    cfg0 = _box_cfg((6, 6, 3), 1.0)
    cfg, sd = orc.meshprep(copy.deepcopy(cfg0))
    e0 = cfg["elem"][:, :4] - 1
    eseg = (cfg["node"][e0].mean(axis=1)[:, 0] > 3).astype(int)
    nseg = (cfg["node"][:, 0] > 3).astype(int)
    ct = copy.deepcopy(cfg)
    ct["seg"] = eseg
    ct["prop"] = np.array([[0.01, 1.0, 0, 1.37], [0.016, 1.0, 0, 1.37]])
    y0, _ = orc.runforward(ct)
    L = np.array([[1.0, -0.9], [-0.9, 1.0]])
    out = {}
    for name, kw in (("lmat", {"lmat": L}), ("ltl", {"ltl": 5 * np.eye(2)}), ("both", {"lmat": L, "ltl": 5 * np.eye(2)})):
        c = copy.deepcopy(cfg)
        c["seg"] = eseg
        rec = {"prop": np.array([[0.01, 1.0, 0, 1.37], [0.01, 1.0, 0, 1.37]]), "seg": nseg}
        orc.runrecon(c, rec, y0, sd, maxiter=1, lambda_=1e-2, **kw)
        out[name] = rec["prop"][:, 0].copy()
    np.testing.assert_array_equal(out["both"], out["lmat"])
    assert not np.allclose(out["lmat"], out["ltl"], rtol=1e-3)
