"""The LDS-staged SpMM plan (rb_spmm_plan_count / rb_spmm_plan_fill, host code, no GPU).

The plan alone must reproduce q = A p: for every block, row r's records name
its columns through the block's union list (LDS byte offset / 512 = union
slot) and its values through the CSR slot map; the row's own slot is where
the p.q partial reads p.  Checked against scipy's A @ p on a Morton-ordered
Kuhn box and on an irregular (centroid-split) mesh, with the block limits the
kernel's LDS buffers rely on.
"""

import ctypes

import numpy as np
import pytest
from scipy import sparse

from redbirdpy_amd import _abi, _gpu

RMAX = 64
LIMITS = [(128, 768)]  # the kernel's (RB_LDS_UMAX, RB_LDS_RECMAX)


def _pattern(node, elem):
    perm = _gpu.morton_order(node)
    iperm = np.empty_like(perm)
    iperm[perm] = np.arange(len(perm))
    e = iperm[np.asarray(elem)[:, :4].astype(np.int64) - 1]
    r = np.repeat(e, 4, axis=1).ravel()
    c = np.tile(e, (1, 4)).ravel()
    A = sparse.csr_matrix((np.ones(len(r)), (r, c)), shape=(len(perm),) * 2)
    A.sum_duplicates()
    A.sort_indices()
    rng = np.random.default_rng(0)
    A.data = rng.uniform(-1, 1, A.nnz)
    return A


def _plan(A, limits):
    p = _gpu.spmm_plan(A.indptr, A.indices, "cpu", limits)
    assert p is not None
    brow, bu, ul, ri, rm, lo, _ = [t.numpy() for t in p.tensors]
    return p.struct.nblk, p.struct.nrec, brow, bu, ul, ri, rm, lo


def _check(A, limits):
    UMAX, RECMAX = limits
    n = A.shape[0]
    nblk, nrec, brow, bu, ul, ri, rm, lo = _plan(A, limits)
    assert brow[0] == 0 and brow[nblk] == n and np.all(np.diff(brow[: nblk + 1]) >= 1)
    assert np.all(np.diff(brow[: nblk + 1]) <= RMAX)
    assert np.all(np.diff(bu[: nblk + 1]) <= UMAX)
    assert ri[2 * n] == nrec and np.all(np.diff(ri[0 : 2 * n + 2 : 2]) % 8 == 0)
    real = rm[:nrec] >= 0
    assert np.array_equal(np.sort(rm[:nrec][real]), np.arange(A.nnz))  # every CSR entry exactly once
    assert np.all(rm[nrec:] == -1) and np.all(lo[nrec:] == 0) and np.all(ul[bu[nblk]:] == 0)
    P = np.random.default_rng(1).standard_normal((n, 3))
    Q = np.zeros_like(P)
    vals = np.where(rm >= 0, A.data[np.maximum(rm, 0)], 0.0)
    for b in range(nblk):
        U = ul[bu[b] : bu[b + 1]]
        assert np.all(np.diff(U) > 0)
        assert ri[2 * brow[b + 1]] - ri[2 * brow[b]] <= RECMAX
        for r in range(brow[b], brow[b + 1]):
            assert U[ri[2 * r + 1] // 512] == r  # the row's own p
            recs = np.arange(ri[2 * r], ri[2 * r + 2])
            assert np.all(lo[recs] % 512 == 0) and np.all(lo[recs] // 512 < len(U))
            Q[r] = vals[recs] @ P[U[lo[recs] // 512]]
    ref = A @ P
    assert np.max(np.abs(Q - ref)) <= 1e-12 * np.max(np.abs(ref))
    return nblk, len(ul) - _abi.RB_LDS_UPAD


@pytest.mark.parametrize("limits", LIMITS)
def test_plan_reproduces_spmm_on_kuhn_box(limits):
    from redbirdpy_amd.mesh import kuhn_box

    node, elem = kuhn_box([0, 0, 0], [40, 40, 20], 2.0)
    A = _pattern(node, elem)
    nblk, nunion = _check(A, limits)
    assert nunion / A.shape[0] < 6.0  # each p row is staged a few times, not once per entry (~14.5)


@pytest.mark.parametrize("limits", LIMITS)
def test_plan_reproduces_spmm_on_irregular_mesh(limits):
    from redbirdpy_amd.mesh import kuhn_box

    node, elem = kuhn_box([0, 0, 0], [16, 16, 8], 2.0)
    cen = node[elem[:, :4] - 1].mean(axis=1)
    split = np.flatnonzero(np.sum((cen - [8.0, 8.0, 4.0]) ** 2, axis=1) < 12.0)
    keep = np.setdiff1d(np.arange(len(elem)), split)
    ids = node.shape[0] + 1 + np.arange(len(split))
    quads = []
    for (a, b, c, d), x in zip(elem[split, :4], ids):
        quads += [(a, b, c, x), (a, b, x, d), (a, x, c, d), (x, b, c, d)]
    elem2 = np.vstack([elem[keep, :4], np.array(quads)])
    A = _pattern(np.vstack([node, cen[split]]), elem2)
    assert np.diff(A.indptr).max() > 16  # rows longer than one 16-record pair of groups
    _check(A, limits)


def test_plan_refuses_a_row_beyond_the_limits():
    n = 400
    A = sparse.identity(n, format="csr") + sparse.csr_matrix((np.ones(300), ([5] * 300, np.arange(100, 400))),
                                                             shape=(n, n))
    A.sort_indices()
    L = _abi.load()
    rp, ci = A.indptr.astype(np.int32), A.indices.astype(np.int32)
    a = ctypes.c_int64(0)
    rc = L.rb_spmm_plan_count(n, rp.ctypes.data_as(ctypes.c_void_p), ci.ctypes.data_as(ctypes.c_void_p), 128, 768,
                              ctypes.byref(a), ctypes.byref(a), ctypes.byref(a))
    assert rc == -3  # RB_ERR_SIZE: the solve keeps the gather SpMM
    rc = L.rb_spmm_plan_count(n, rp.ctypes.data_as(ctypes.c_void_p), ci.ctypes.data_as(ctypes.c_void_p), 100, 768,
                              ctypes.byref(a), ctypes.byref(a), ctypes.byref(a))
    assert rc == -1  # no kernel for these limits
    assert _gpu.spmm_plan(A.indptr, A.indices, "cpu") is None


def test_sweep_columns_groups_of_64_under_the_descriptor_limit(monkeypatch):
    """Columns per solve call: the whole width while its float64 vectors fit 2 GiB, else the
    largest multiple of 64 that fits; several systems or no fitting 64-column group keep the width."""
    from redbirdpy_amd import _gpu

    monkeypatch.delenv("RB_SPMM_FLAT_MAX_BYTES", raising=False)
    n = 2_610_000
    assert _gpu.sweep_columns(n, 272, 1, False) == 64  # 2^31 / (8 n) = 102 columns
    assert _gpu.sweep_columns(n, 272, 1, True) == 272  # complex: not even 64 columns fit
    assert _gpu.sweep_columns(n, 272, 4, False) == 272
    assert _gpu.sweep_columns(115_000, 512, 1, True) == 512
    assert _gpu.sweep_columns(1_000_000, 512, 1, False) == 256  # 268 columns fit
    assert _gpu.sweep_columns(n, 64, 1, False) == 64
    monkeypatch.setenv("RB_SPMM_FLAT_MAX_BYTES", str(726 * 8 * 100))
    assert _gpu.sweep_columns(726, 272, 1, False) == 64
    assert _gpu.sweep_columns(726, 99, 1, False) == 99


def test_plans_are_kept_by_pattern_content():
    """The same pattern from other arrays (a mesh rebuilt from copies) returns the kept plan; another
    pattern builds its own."""
    import numpy as np
    import scipy.sparse as sp
    from redbirdpy_amd import _gpu

    A = sp.random(300, 300, density=0.02, format="csr", random_state=3) + sp.eye(300, format="csr")
    A.sort_indices()
    p1 = _gpu.spmm_plan(A.indptr, A.indices, "cpu")
    p2 = _gpu.spmm_plan(A.indptr.copy(), A.indices.copy(), "cpu")
    assert p1 is not None and p2 is p1
    B = A + sp.eye(300, k=5, format="csr")
    B.sort_indices()
    p3 = _gpu.spmm_plan(B.indptr, B.indices, "cpu")
    assert p3 is not p1
