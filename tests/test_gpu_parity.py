"""GPU parity of the HIP path against the CPU oracle (same seeded inputs).

Integer work (CSR pattern, connectivity, point location) must be
bit-exact; float64 values within the tolerances written in each test.
"""

import numpy as np
import pytest
from scipy import sparse

from conftest import box_cfg, grid_optodes

pytestmark = pytest.mark.gpu


def _both(make):
    import redbirdpy_amd as rb
    from oracle import rb_oracle as orc

    c1, sd1 = rb.meshprep(make())
    c2, sd2 = orc.meshprep(make())
    return rb, orc, c1, sd1, c2, sd2


def _relmax(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def test_meshprep_integer_bitexact():
    rb, orc, c1, sd1, c2, sd2 = _both(lambda: box_cfg((12, 10, 6), 2.0))
    assert np.array_equal(c1["elem"], c2["elem"])
    assert np.array_equal(c1["face"], c2["face"])
    assert np.array_equal(c1["idxcount"], c2["idxcount"])
    assert np.array_equal(c1["rows"], c2["rows"])
    assert np.array_equal(c1["cols"], c2["cols"])  # both sorted within each row
    assert np.array_equal(sd1, sd2)
    np.testing.assert_allclose(c1["evol"], c2["evol"], rtol=1e-14)
    np.testing.assert_allclose(c1["area"], c2["area"], rtol=1e-14)
    np.testing.assert_allclose(c1["nvol"], c2["nvol"], rtol=1e-13)
    np.testing.assert_allclose(c1["deldotdel"], c2["deldotdel"], rtol=1e-12, atol=1e-14)
    assert c1["reff"] == pytest.approx(c2["reff"], rel=1e-15)


@pytest.mark.parametrize("branch", ["label", "elem", "node"])
@pytest.mark.parametrize("omega", [0.0, 2 * np.pi * 1e8])
def test_femlhs_matches_oracle(branch, omega):
    def make():
        cfg = box_cfg((12, 10, 6), 2.0, omega=omega)
        rng = np.random.default_rng(7)
        ne = cfg["elem"].shape[0]
        nn = cfg["node"].shape[0]
        if branch == "elem":
            cfg["prop"] = np.column_stack([rng.uniform(0.005, 0.03, ne), rng.uniform(0.5, 1.5, ne),
                                           np.zeros(ne), rng.uniform(1.3, 1.45, ne)])
        elif branch == "node":
            cfg["prop"] = np.column_stack([rng.uniform(0.005, 0.03, nn), rng.uniform(0.5, 1.5, nn),
                                           np.zeros(nn), rng.uniform(1.3, 1.45, nn)])
            cfg.pop("seg")
        else:
            lab = np.ones(ne, dtype=int)
            lab[rng.random(ne) < 0.3] = 2
            cfg["seg"] = lab
            cfg["prop"] = np.array([[0, 0, 1, 1], [0.01, 1.0, 0, 1.37], [0.02, 1.2, 0, 1.37]])
        return cfg

    rb, orc, c1, sd1, c2, sd2 = _both(make)
    A1 = rb.femlhs(c1, c1["deldotdel"])
    A2 = orc.femlhs(c2, c2["deldotdel"])
    A2.sort_indices()
    assert np.array_equal(A1.indptr, A2.indptr)
    assert np.array_equal(A1.indices, A2.indices)
    assert np.iscomplexobj(A1.data) == np.iscomplexobj(A2.data)
    # entry-wise, relative to the summed magnitude of the contributions
    scale = np.max(np.abs(A2.data))
    assert np.max(np.abs(A1.data - A2.data)) <= 1e-14 * scale


def test_femlhs_multiwavelength():
    def make():
        cfg = box_cfg((12, 10, 6), 2.0)
        cfg["prop"] = {"690": np.array([[0, 0, 1, 1], [0.012, 1.1, 0, 1.37]]),
                       "830": np.array([[0, 0, 1, 1], [0.008, 0.9, 0, 1.37]])}
        return cfg

    rb, orc, c1, sd1, c2, sd2 = _both(make)
    for wv in ("690", "830"):
        A1 = rb.femlhs(c1, c1["deldotdel"], wavelength=wv)
        A2 = orc.femlhs(c2, c2["deldotdel"], wavelength=wv)
        assert np.max(np.abs((A1 - A2).data)) <= 1e-14 * np.max(np.abs(A2.data))


def test_tsearch_and_femrhs():
    def make():
        p1 = (20, 16, 8)
        return box_cfg(p1, 2.0, src=grid_optodes(p1, 3), det=grid_optodes(p1, 2) + [0.3, 0.7, 0])

    rb, orc, c1, sd1, c2, sd2 = _both(make)
    r1, l1, b1, o1 = rb.femrhs(c1, sd1)
    r2, l2, b2, o2 = orc.femrhs(c2, sd2)
    assert np.array_equal(np.isnan(l1), np.isnan(l2))
    assert np.array_equal(l1[~np.isnan(l1)], l2[~np.isnan(l2)])
    np.testing.assert_allclose(b1, b2, rtol=0, atol=1e-12)
    assert abs(r1 - r2).max() <= 1e-12
    np.testing.assert_allclose(o1, o2, rtol=1e-15)


@pytest.mark.parametrize("omega", [0.0, 2 * np.pi * 1e8])
def test_femsolve_vs_superlu(omega):
    rb, orc, c1, sd1, c2, sd2 = _both(lambda: box_cfg((16, 16, 8), 2.0, omega=omega))
    A = orc.femlhs(c2, c2["deldotdel"])
    rhs, _, _, _ = orc.femrhs(c2, sd2)
    xr, _ = orc.femsolve(A, rhs)
    rhs_d = rhs.toarray()
    rhs_d = np.column_stack([rhs_d, np.zeros(rhs_d.shape[0])])  # a zero column stays zero
    x, flag = rb.femsolve(A, rhs_d)
    assert flag == 0
    assert np.all(x[:, -1] == 0)
    err = np.linalg.norm(x[:, :-1] - xr) / np.linalg.norm(xr)
    assert err < 1e-10, err


def test_femsolve_mixed_precision():
    rb, orc, c1, sd1, c2, sd2 = _both(lambda: box_cfg((16, 16, 8), 2.0, omega=2 * np.pi * 1e8))
    A = orc.femlhs(c2, c2["deldotdel"])
    rhs, _, _, _ = orc.femrhs(c2, sd2)
    xr, _ = orc.femsolve(A, rhs)
    x, flag = rb.femsolve(A, rhs, precision="mixed", tol=1e-11)
    assert flag == 0
    err = np.linalg.norm(x - xr) / np.linalg.norm(xr)
    assert err < 1e-9, err


@pytest.mark.parametrize("omega", [0.0, 2 * np.pi * 1e8])
def test_runforward_matches_oracle(omega):
    def make():
        p1 = (20, 20, 10)
        return box_cfg(p1, 2.0, src=grid_optodes(p1, 2), det=grid_optodes(p1, 3, margin=3.0), omega=omega)

    rb, orc, c1, sd1, c2, sd2 = _both(make)
    d1, p1 = rb.runforward(c1)
    d2, p2 = orc.runforward(c2)
    assert d1.shape == d2.shape and p1.shape == p2.shape
    assert np.linalg.norm(p1 - p2) / np.linalg.norm(p2) < 1e-10
    assert np.max(np.abs(d1 - d2) / np.abs(d2)) < 1e-9
    # detector readout of the oracle field through the GPU kernel
    r2, l2, b2, _ = orc.femrhs(c2, sd2)
    g = rb.femgetdet(p2, c2, r2, l2, b2)
    np.testing.assert_allclose(g, d2, rtol=1e-13)


@pytest.mark.parametrize("omega", [0.0, 2 * np.pi * 1e8])
def test_jac_matches_oracle(omega):
    def make():
        p1 = (16, 16, 8)
        return box_cfg(p1, 2.0, src=grid_optodes(p1, 2), det=grid_optodes(p1, 3, margin=3.0), omega=omega)

    rb, orc, c1, sd1, c2, sd2 = _both(make)
    _, phi = orc.runforward(c2)
    sd = sd2.copy()
    sd[[1, 4], 2] = 0  # inactive pairs are dropped
    Jn1, Je1 = rb.jac(sd, phi, c1["deldotdel"], c1["elem"], c1["evol"])
    Jn2, Je2 = orc.jac(sd, phi, c2["deldotdel"], c2["elem"], c2["evol"])
    assert Jn1.shape == Jn2.shape and Je1.shape == Je2.shape
    assert np.linalg.norm(Jn1 - Jn2) / np.linalg.norm(Jn2) < 1e-12
    assert np.linalg.norm(Je1 - Je2) / np.linalg.norm(Je2) < 1e-12


def test_multiwavelength_forward():
    def make():
        p1 = (16, 16, 8)
        cfg = box_cfg(p1, 2.0, src=grid_optodes(p1, 2), det=grid_optodes(p1, 2, margin=3.0))
        cfg["prop"] = {"690": np.array([[0, 0, 1, 1], [0.012, 1.1, 0, 1.37]]),
                       "830": np.array([[0, 0, 1, 1], [0.008, 0.9, 0, 1.37]])}
        return cfg

    rb, orc, c1, sd1, c2, sd2 = _both(make)
    d1, p1 = rb.runforward(c1, sd=sd1)
    d2, p2 = orc.runforward(c2, sd=sd2)
    for wv in ("690", "830"):
        assert np.linalg.norm(p1[wv] - p2[wv]) / np.linalg.norm(p2[wv]) < 1e-10
        assert np.max(np.abs(d1[wv] - d2[wv]) / np.abs(d2[wv])) < 1e-9


@pytest.mark.parametrize("ncol", [3, 40, 64, 130])
def test_femsolve_long_rows_and_widths(ncol):
    """Rows longer than one wavefront, and every column-group width of the SpMM."""
    import redbirdpy_amd as rb
    from scipy.sparse.linalg import spsolve

    rng = np.random.default_rng(ncol)
    n = 3000
    A = sparse.random(n, n, density=0.002, random_state=1, format="csr")
    A = A + A.T
    # rows with ~150 entries, and one with ~900 (more than one staging window
    # of the streaming SpMM)
    for r, nz in ((5, 150), (17, 900), (2999, 150)):
        cols = rng.choice(n, nz, replace=False)
        B = sparse.csr_matrix((rng.uniform(-0.1, 0.1, nz), ([r] * nz, cols)), shape=(n, n))
        A = A + B + B.T
    A = A + sparse.diags(np.asarray(abs(A).sum(axis=1)).ravel() + 1.0)  # SPD by diagonal dominance
    b = rng.standard_normal((n, ncol))
    x, flag = rb.femsolve(A, b, tol=1e-13)
    assert flag == 0
    xr = spsolve(A.tocsc(), b)
    xr = xr.reshape(n, -1)
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-11
    xm, flag = rb.femsolve(A, b, tol=1e-12, precision="mixed")
    assert flag == 0
    assert np.linalg.norm(xm - xr) / np.linalg.norm(xr) < 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("cap", ["8", None])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_femsolve_sweep_many_chunk_groups(precision, cap, monkeypatch):
    """A capped SpMM grid gives each wavefront hundreds of row chunks (many
    row-pointer groups per wavefront) on the FD box at 2 mm spacing; uncapped,
    the grid is the default resident round (the mixed complex case is the
    benchmark's k_cg_spmm_sweep_ru<complex double, complex float, 1> kernel)."""
    import redbirdpy_amd as rb
    from scipy.sparse.linalg import spsolve
    from redbirdpy_amd import workloads

    cfg = workloads.fd_box(h=2.0)
    cfg, _ = rb.meshprep(cfg)
    A = rb.femlhs(cfg, cfg["deldotdel"])
    n = A.shape[0]
    rng = np.random.default_rng(7)
    b = rng.standard_normal((n, 64)) + 1j * rng.standard_normal((n, 64))
    xr = spsolve(A.tocsc(), b)
    if cap:
        monkeypatch.setenv("RB_SWEEP_MAX_GRID", cap)
    else:
        monkeypatch.delenv("RB_SWEEP_MAX_GRID", raising=False)
    x, flag = rb.femsolve(A, b, tol=1e-12, precision=precision)
    assert flag == 0
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-9


def test_setmesh_recomputes_mesh_fields():
    import redbirdpy_amd as rb
    from redbirdpy_amd.mesh import kuhn_box

    cfg, _ = rb.meshprep(box_cfg((12, 10, 6), 2.0))
    node, elem = kuhn_box([0, 0, 0], [16, 12, 8], 2.0)
    c2 = rb.setmesh(cfg, node, np.column_stack([elem, np.ones(len(elem), dtype=int)]))
    assert c2["node"] is node and c2["elem"].shape == elem.shape
    assert c2["face"].shape[0] == 4 * (8 * 6 + 8 * 4 + 6 * 4)  # Kuhn box boundary triangles
    assert c2["evol"].sum() == pytest.approx(16 * 12 * 8, rel=1e-12)
    assert np.array_equal(c2["seg"], np.ones(len(elem), dtype=int))
    assert cfg["face"].shape[0] != c2["face"].shape[0]  # the original is left alone


@pytest.mark.parametrize("nw", [2, 4, 8])
@pytest.mark.parametrize("cap", [None, "8"])
@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_runforward_multiwavelength_64col(nw, cap, precision, monkeypatch):
    """Real multi-wavelength solves with 64 columns each (32 sources + 32 detectors): the SpMM
    sweeps the wavelengths' systems in pairs (two columns per lane: 8-byte gathers in float32,
    16-byte in float64; each half of the wavefront scaled by its own system's values).  Detector
    values and fields against the oracle; the capped grid gives each wavefront many row-chunk
    groups."""
    import copy
    from oracle import rb_oracle as orc
    import redbirdpy_amd as rb

    xs, ys = np.meshgrid(np.linspace(3, 17, 8), np.linspace(3, 17, 4))
    src = np.column_stack([xs.ravel(), ys.ravel(), np.zeros(32)])
    cfg = box_cfg((20, 20, 10), 2.0, src=src, det=src + [1.0, 1.0, 0.0])
    cfg["prop"] = {str(650 + 25 * i): np.array([[0, 0, 1, 1], [0.004 + 0.003 * i, 1.0 - 0.05 * i, 0, 1.37]])
                   for i in range(nw)}
    if cap:
        monkeypatch.setenv("RB_SWEEP_MAX_GRID", cap)
    c1, _ = rb.meshprep(copy.deepcopy(cfg))
    c2, _ = orc.meshprep(copy.deepcopy(cfg))
    d1, f1 = rb.runforward(c1, precision=precision, tol=1e-12)
    d2, f2 = orc.runforward(c2)
    assert set(d1) == set(d2) and len(d2) == nw
    for wv in d2:
        assert d1[wv].shape == (32, 32)
        assert _relmax(d1[wv], d2[wv]) < 1e-8
        assert np.linalg.norm(f1[wv] - f2[wv]) / np.linalg.norm(f2[wv]) < 1e-9


def _star_split_cfg(omega=0.0):
    """A Kuhn box whose tets around the centre are split 1 -> 4 at their centroids:
    irregular connectivity (rows from 5 to ~40 entries) for pattern, assembly, solve
    and the node Jacobian's patch matrices."""
    from redbirdpy_amd.mesh import kuhn_box

    p1 = (16, 16, 8)
    node, elem = kuhn_box([0, 0, 0], p1, 2.0)
    cen = node[elem[:, :4] - 1].mean(axis=1)
    split = np.flatnonzero(np.sum((cen - [8.0, 8.0, 4.0]) ** 2, axis=1) < 12.0)
    keep = np.setdiff1d(np.arange(len(elem)), split)
    new_nodes = cen[split]
    ids = node.shape[0] + 1 + np.arange(len(split))
    quads = []
    for (a, b, c, d), x in zip(elem[split, :4], ids):
        quads += [(a, b, c, x), (a, b, x, d), (a, x, c, d), (x, b, c, d)]
    elem2 = np.vstack([elem[keep, :4], np.array(quads)]).astype(int)
    cfg = box_cfg(p1, 2.0, src=grid_optodes(p1, 2), det=grid_optodes(p1, 3, margin=3.0), omega=omega)
    cfg["node"] = np.vstack([node, new_nodes])
    cfg["elem"] = elem2
    cfg["seg"] = np.ones(len(elem2), dtype=int)
    return cfg


@pytest.mark.parametrize("omega", [0.0, 2 * np.pi * 1e8])
def test_irregular_mesh_forward_and_jacobian(omega):
    import copy

    rb, orc, c1, sd1, c2, sd2 = _both(lambda: copy.deepcopy(_star_split_cfg(omega)))
    assert np.array_equal(c1["face"], c2["face"]) and np.array_equal(c1["idxcount"], c2["idxcount"])
    counts = c1["idxcount"]
    assert counts.max() >= 3 * np.median(counts)  # the split nodes have long rows
    d1, p1 = rb.runforward(c1)
    d2, p2 = orc.runforward(c2)
    assert _relmax(d1, d2) < 1e-9
    assert np.linalg.norm(p1 - p2) / np.linalg.norm(p2) < 1e-10
    Jn1, Je1 = rb.jac(sd2, p2, c1["deldotdel"], c1["elem"], c1["evol"])
    Jn2, Je2 = orc.jac(sd2, p2, c2["deldotdel"], c2["elem"], c2["evol"])
    assert np.linalg.norm(Jn1 - Jn2) / np.linalg.norm(Jn2) < 1e-12
    assert np.linalg.norm(Je1 - Je2) / np.linalg.norm(Je2) < 1e-12


@pytest.mark.parametrize("precision,nw,ncol", [("mixed", 1, 272), ("fp64", 3, 64), ("mixed", 3, 64)])
def test_row_group_spmm_chunks_past_the_last_column(precision, nw, ncol, monkeypatch):
    """Solves too large for the sweep's buffer descriptor run the row-per-group SpMM and float64
    residual kernels over 64-column chunks rounded up to a power of two: one system of 272
    columns (chunks 5-7 lie past the last column) and three 64-column systems (chunk 3 past it)
    have no values there.  RB_SPMM_FLAT_MAX_BYTES forces that path on a small box; results against
    the oracle's SuperLU solves."""
    import copy
    from oracle import rb_oracle as orc
    import redbirdpy_amd as rb

    monkeypatch.setenv("RB_SPMM_FLAT_MAX_BYTES", "4096")
    ns = ncol // 2
    xs, ys = np.meshgrid(np.linspace(3, 17, 8), np.linspace(3, 17, (ns + 7) // 8))
    src = np.column_stack([xs.ravel(), ys.ravel(), np.zeros(xs.size)])[:ns]
    cfg = box_cfg((20, 20, 10), 2.0, src=src, det=src + [1.0, 1.0, 0.0])
    if nw > 1:
        cfg["prop"] = {str(650 + 25 * i): np.array([[0, 0, 1, 1], [0.004 + 0.003 * i, 1.0, 0, 1.37]])
                       for i in range(nw)}
    c1, _ = rb.meshprep(copy.deepcopy(cfg))
    c2, _ = orc.meshprep(copy.deepcopy(cfg))
    d1, f1 = rb.runforward(c1, precision=precision, tol=1e-12)
    d2, f2 = orc.runforward(c2)
    if nw == 1:
        d1, d2, f1, f2 = {"": d1}, {"": d2}, {"": f1}, {"": f2}
    for wv in d2:
        assert f1[wv].shape[1] == ncol
        assert _relmax(d1[wv], d2[wv]) < 1e-8
        assert np.linalg.norm(f1[wv] - f2[wv]) / np.linalg.norm(f2[wv]) < 1e-9


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_wide_solve_runs_as_column_groups_that_fit_the_sweep(precision, monkeypatch):
    """A one-system solve whose float64 vectors exceed the sweep's descriptor limit runs as
    64-column groups (the limit lowered so 272 columns of a 726-node box do not fit and 64 do):
    femsolve and the batch path's solve_system both, against the oracle's SuperLU solves."""
    import copy
    from oracle import rb_oracle as orc
    import redbirdpy_amd as rb
    from redbirdpy_amd import _gpu
    from redbirdpy_amd.forward import BatchForward

    xs, ys = np.meshgrid(np.linspace(3, 17, 8), np.linspace(3, 17, 17))
    src = np.column_stack([xs.ravel(), ys.ravel(), np.zeros(xs.size)])
    cfg = box_cfg((20, 20, 10), 2.0, src=src, det=src + [1.0, 1.0, 0.0])
    c1, _ = rb.meshprep(copy.deepcopy(cfg))
    c2, _ = orc.meshprep(copy.deepcopy(cfg))
    n = c1["node"].shape[0]
    monkeypatch.setenv("RB_SPMM_FLAT_MAX_BYTES", str(n * 8 * 100))
    assert _gpu.sweep_columns(n, 272, 1, False) == 64
    d1, f1 = rb.runforward(c1, precision=precision, tol=1e-12)
    d2, f2 = orc.runforward(c2)
    assert f1.shape[1] == 272
    assert _relmax(d1, d2) < 1e-8
    assert np.linalg.norm(f1 - f2) / np.linalg.norm(f2) < 1e-9
    res = BatchForward([c1]).solve(precision, 1e-12)
    assert _relmax(res[0].detval, d2) < 1e-8
    assert np.linalg.norm(res[0].phi - f2) / np.linalg.norm(f2) < 1e-9
