"""CPU oracle: an FP64 numpy/scipy restatement of redbirdpy's FEM hot path.

TEST INFRASTRUCTURE ONLY.  This module is the checker that the HIP path is
compared against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product package
``redbirdpy_amd`` never does, and must fail loudly rather than fall back to it.

Every function restates the reference algorithm it names (file:line in
fangq/redbirdpy v0.2.0) in code written for this project.  Where the
reference calls into a third-party library that is absent here, the
published algorithm is restated instead:

* ``tsearchn``  -- iso2mesh's point location (iso2mesh >= 0.5.0, not
  installed): brute-force barycentric test, lowest containing tet wins.
* ``femsolve``  -- the reference's default path on this container is
  scipy SuperLU (``solver.py:455-476``, ``:561-588``); it is called directly.

Pinning (what checks each function, see tests/test_oracle_*.py):

* meshprep integer outputs: closed-form Kuhn-grid counts, brute force on
  1/2/6-tet meshes, exact volume sums.
* deldotdel: closed form on Kuhn tets and an independent gradient computed
  from the inverse of the [1 x y z] vertex matrix.
* femlhs: quadrature-based brute-force assembly on small meshes (exact P1
  integrals), symmetry, row-sum identities, 7-point stencil.
* femsolve/femgetdet: residual, reciprocity.
* jac: central finite differences of femgetdet under the jac derivative
  model, and the mass-matrix identity.
* getreff: independent adaptive quadrature of the Haskell integrals.
* reginvunder/reginvover: numpy lstsq on the augmented Tikhonov system;
  with priors, lstsq on [A; sqrt(lam) L] (also block-wise L^T L) and the
  closed form P A^T (A P A^T + lam I)^-1 b, P = (L^T L)^-1 (also per block).
* masksum: the adjoint of label expansion (masksum(J) v = J v[labels]).
* matreform: 'real'/'logphase' rows by central differences of detval,
  log|detval| and arg(detval); 'reim' rows give [Re z; Im z] of the complex
  Tikhonov step z = J^H (J J^H + lam I)^-1 r.
* meshinterp / remap_jacobian: linear fields reproduced exactly; the
  adjoint identity J_rec v = J (P v); column sums preserved.
* src2bc / qmeshcut / inpolygon: the reference's inpolygon known answers,
  cut points on the plane and on their edges, cross-section areas of
  horizontal and tilted planes, the depth moment of the P1 weights, the
  Fourier closed form per face, L1 norms.
* extinction: literal table entries and linear interpolation identities.
"""

import numpy as np
from scipy import sparse
from scipy.sparse.linalg import splu, spsolve
import os

C0 = 299792458000.0  # mm/s, forward.py:36
R_C0 = 1.0 / C0

_DATA = os.path.join(
    os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
    "redbirdpy_amd",
    "data",
    "hemoglobin_prahl.csv",
)


# ----------------------------------------------------------------------------
# utility.py
# ----------------------------------------------------------------------------


def forcearray(cfg, keys):
    """utility.py:1112-1117 -- list-valued fields become numpy arrays."""
    for k in keys:
        if k in cfg and isinstance(cfg[k], list):
            cfg[k] = np.array(cfg[k])
    return cfg


def deldotdel(cfg):
    """utility.py:53-98 -- V * grad(phi_i).grad(phi_j), 10 upper-tri pairs.

    Gradients come from the cofactor differences of the vertex coordinates
    divided by 6V; the cofactor index table is the reference's ``col``.
    """
    node = cfg["node"]
    evol = cfg["evol"]
    e0 = cfg["elem"][:, :4].astype(np.int64) - 1
    ne = e0.shape[0]
    xyz = node[e0, :].transpose(2, 1, 0)  # (3, 4, ne)
    tbl = ((3, 1, 2, 1), (2, 0, 3, 2), (1, 3, 0, 3), (0, 2, 1, 0))
    inv6v = 1.0 / (6.0 * evol)
    grad = np.zeros((3, 4, ne))
    for c in range(3):
        a, b = [d for d in range(3) if d != c]
        for i in range(4):
            t = tbl[i]
            grad[c, i] = (
                (xyz[a, t[0]] - xyz[a, t[1]]) * (xyz[b, t[2]] - xyz[b, t[3]])
                - (xyz[a, t[2]] - xyz[a, t[3]]) * (xyz[b, t[0]] - xyz[b, t[1]])
            ) * inv6v
    out = np.zeros((ne, 10))
    col = 0
    for i in range(4):
        for j in range(i, 4):
            out[:, col] = np.sum(grad[:, i] * grad[:, j], axis=0)
            col += 1
    out *= evol[:, None]
    return out, grad


def meshreorient(node, elem):
    """utility.py:1047-1063 -- swap local nodes 0,1 of negative-volume tets."""
    elem = np.array(elem, copy=True)
    e0 = elem[:, :4].astype(np.int64) - 1
    p = node[e0]
    v = np.einsum(
        "ij,ij->i", np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), p[:, 3] - p[:, 0]
    )
    neg = v < 0
    tmp = elem[neg, 0].copy()
    elem[neg, 0] = elem[neg, 1]
    elem[neg, 1] = tmp
    return elem


def volface(elem):
    """utility.py:1066-1087 -- boundary triangles, 1-based, outward order.

    Faces are taken in the reference's local order [0,2,1],[0,1,3],[0,3,2],
    [1,2,3], stacked face-type-major; a face is on the boundary when its
    sorted triple occurs once.  Output order is lexicographic in the sorted
    triple (np.unique), and each face keeps its original orientation.
    """
    e0 = elem[:, :4].astype(np.int64) - 1
    allf = np.vstack([e0[:, [0, 2, 1]], e0[:, [0, 1, 3]], e0[:, [0, 3, 2]], e0[:, [1, 2, 3]]])
    srt = np.sort(allf, axis=1)
    _, first, cnt = np.unique(srt, axis=0, return_index=True, return_counts=True)
    return allf[first[cnt == 1]] + 1


def elemvolume(node, elem):
    """utility.py:1090-1109 -- |det|/6 for tets, half cross-product norm for triangles."""
    e0 = elem.astype(np.int64) - 1
    if elem.shape[1] >= 4:
        a = node[e0[:, 0]]
        return np.abs(
            np.sum(np.cross(node[e0[:, 1]] - a, node[e0[:, 2]] - a) * (node[e0[:, 3]] - a), axis=1)
        ) / 6.0
    if elem.shape[1] == 3:
        a = node[e0[:, 0]]
        cr = np.cross(node[e0[:, 1]] - a, node[e0[:, 2]] - a)
        return 0.5 * np.sqrt(np.sum(cr * cr, axis=1))
    raise ValueError("unsupported element")


def nodevolume(node, elem, evol):
    """utility.py:1005-1017 -- a quarter of the volume of every incident tet."""
    e0 = elem[:, :4].astype(np.int64) - 1
    nv = np.zeros(node.shape[0])
    for j in range(4):
        np.add.at(nv, e0[:, j], evol)
    return nv * 0.25


def femnz(elem, nn):
    """utility.py:1020-1041 -- node adjacency (off-diagonal pattern).

    The reference fills Python sets, so the order of ``cols`` within a row is
    CPython's set order; here each row is sorted.  Compare as per-row sets.
    """
    e0 = elem[:, :4].astype(np.int64) - 1
    r = []
    c = []
    for i in range(4):
        for j in range(4):
            if i != j:
                r.append(e0[:, i])
                c.append(e0[:, j])
    key = np.unique(np.concatenate(r) * nn + np.concatenate(c))
    rows = key // nn
    cols = key % nn
    cnt = np.bincount(rows, minlength=nn)
    return rows, cols, cnt


def getreff(n_in, n_out=1.0):
    """utility.py:879-904 -- Haskell effective reflection, Riemann sum of step pi/2000."""
    if n_in <= n_out:
        return 0.0
    oc = np.arcsin(n_out / n_in)
    dt = np.pi / 2000
    th = np.arange(0, oc, dt)
    ct = np.cos(th)
    ctp = np.sqrt(1 - (n_in * np.sin(th)) ** 2)
    rs = 0.5 * ((n_in * ctp - n_out * ct) / (n_in * ctp + n_out * ct)) ** 2
    rs = rs + 0.5 * ((n_in * ct - n_out * ctp) / (n_in * ct + n_out * ctp)) ** 2
    thf = np.arange(0, np.pi / 2, dt)
    rf = np.ones(len(thf))
    rf[: len(rs)] = rs
    ctf = np.cos(thf)
    rphi = 2 * np.sum(np.sin(thf) * ctf * rf) * dt
    rj = 3 * np.sum(np.sin(thf) * ctf ** 2 * rf) * dt
    return (rphi + rj) / (2 - rphi + rj)


def getltr(cfg, wv=""):
    """utility.py:862-876 -- transport mean free path of the bulk medium."""
    bk = getbulk(cfg)
    if isinstance(bk, dict):
        if not wv:
            wv = list(bk.keys())[0]
        bk = bk[wv]
    return 1.0 / (bk[0] + bk[1] * (1 - bk[2]))


def getoptodes(cfg, wv=""):
    """utility.py:799-822 -- optodes moved inward by dir * l_tr."""
    ltr = getltr(cfg, wv)
    src = det = None
    if cfg.get("srcpos") is not None and np.size(cfg["srcpos"]) > 0:
        sdir = np.atleast_2d(cfg["srcdir"])
        sp = np.atleast_2d(cfg["srcpos"])
        if sdir.shape[0] == 1:
            sdir = np.tile(sdir, (sp.shape[0], 1))
        src = sp + sdir * ltr
    if cfg.get("detpos") is not None and np.size(cfg["detpos"]) > 0:
        ddir = np.atleast_2d(cfg.get("detdir", cfg["srcdir"]))
        dp = np.atleast_2d(cfg["detpos"])
        if ddir.shape[0] == 1:
            ddir = np.tile(ddir, (dp.shape[0], 1))
        det = dp + ddir * ltr
    return src, det, cfg.get("widesrc"), cfg.get("widedet")


def _counts(cfg):
    ns = np.atleast_2d(cfg["srcpos"]).shape[0] if cfg.get("srcpos") is not None and np.size(cfg["srcpos"]) else 0
    nd = np.atleast_2d(cfg["detpos"]).shape[0] if cfg.get("detpos") is not None and np.size(cfg["detpos"]) else 0
    ws = cfg["widesrc"].shape[1] if cfg.get("widesrc") is not None and cfg["widesrc"].size else 0
    wd = cfg["widedet"].shape[1] if cfg.get("widedet") is not None and cfg["widedet"].size else 0
    return ns, nd, ws, wd


def sdmap(cfg, maxdist=np.inf, **kw):
    """utility.py:675-796 -- src/det pair table; row index = d*Nsrc + s.

    Columns: [src column, det column (offset by all sources), active(, mode)].
    """
    ns, nd, ws, wd = _counts(cfg)
    if ns + ws == 0 or nd + wd == 0:
        raise ValueError("Must define at least one source and detector")
    bads = list(kw.get("excludesrc", cfg.get("excludesrc", [])))
    badd = list(kw.get("excludedet", cfg.get("excludedet", [])))
    gs = sorted(set(range(ns)) - set(bads))
    gd = sorted(set(range(nd)) - set(badd))
    off = ns + ws
    srcs = gs + list(range(ns, ns + ws))
    dets = [off + d for d in gd] + [off + d for d in range(nd, nd + wd)]
    ss, dd = np.meshgrid(srcs, dets)
    pairs = np.column_stack([ss.ravel(), dd.ravel()])

    def dist_active():
        act = np.ones(len(pairs))
        if maxdist < np.inf and ns > 0 and nd > 0:
            sp = np.atleast_2d(cfg["srcpos"])[:, :3]
            dp = np.atleast_2d(cfg["detpos"])[:, :3]
            for r, (s, d) in enumerate(pairs):
                di = int(d) - off
                if s < ns and di < nd:
                    act[r] = 1.0 if np.linalg.norm(sp[int(s)] - dp[di]) < maxdist else 0.0
        return act

    if isinstance(cfg.get("prop"), dict):
        out = {}
        for wv in cfg["prop"]:
            act = dist_active()
            out[wv] = np.column_stack([pairs, act, np.ones(len(pairs))])
        return out
    return np.column_stack([pairs, dist_active()])


# ---------------------------------------------------------------------------
# wide-field sources and detectors (utility.py:252-669)
# ---------------------------------------------------------------------------
_TET_EDGES = ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3))


def qmeshcut(elem, node, plane):
    """Plane cut of a tet mesh: iso2mesh.qmeshcut(elem, node, value, plane), called at utility.py:454-479.

    iso2mesh (>= 0.5.0) is not installed; the published algorithm is
    restated.  Signed distance d = a x + b y + c z + e for plane [a b c e];
    a node with d >= 0 counts as above.  Every mesh edge (n1 < n2) whose ends
    lie on opposite sides gives one cut point w*p1 + (1-w)*p2 with
    w = d2 / (d2 - d1).  Every crossing tet gives one polygon: a triangle
    (one node alone on its side; corners in the local order of the other
    three) or a quad (two and two: pos {a<b}, neg {c<d}, corners ac, ad, bd,
    bc, i.e. around the polygon).  Polygons are in tet order.

    Returns (cutpos (npt x 3), facedata (npoly x 4, 1-based; col 2 == col 3
    for a triangle), elemid (npoly, 1-based), nodeid (npt x 3: n1, n2
    1-based, w + 1)).  The weight column carries the +1 offset of the Python
    iso2mesh port, which the reference removes at utility.py:610-612.
    Exact zeros of d (nodes on the plane) follow the d >= 0 rule above:
    parity unpinned for such cuts.
    """
    e = np.asarray(elem)[:, :4].astype(np.int64) - 1
    node = np.asarray(node, dtype=float)
    d = node @ np.asarray(plane[:3], dtype=float) + float(plane[3])
    up = d[e] >= 0
    nup = up.sum(axis=1)
    tets = np.flatnonzero((nup > 0) & (nup < 4))
    ecut = e[tets]
    uc = up[tets]
    # global cut edges of every crossing tet (local edge k of tet t)
    pairs = []
    for a, b in _TET_EDGES:
        m = uc[:, a] != uc[:, b]
        n1 = np.minimum(ecut[:, a], ecut[:, b])
        n2 = np.maximum(ecut[:, a], ecut[:, b])
        pairs.append((m, n1, n2))
    allp = np.concatenate([np.stack([n1[m], n2[m]], 1) for m, n1, n2 in pairs]) if len(tets) else np.zeros((0, 2), int)
    uniq, inv = np.unique(allp, axis=0, return_inverse=True)
    inv = np.asarray(inv).reshape(-1)
    d1, d2 = d[uniq[:, 0]], d[uniq[:, 1]]
    w = d2 / (d2 - d1)
    cutpos = node[uniq[:, 0]] * w[:, None] + node[uniq[:, 1]] * (1 - w)[:, None]
    # point id of local edge k of each crossing tet (-1 when not cut)
    pid = np.full((len(tets), 6), -1, dtype=np.int64)
    o = 0
    for k, (m, _, _) in enumerate(pairs):
        cnt = int(m.sum())
        pid[m, k] = inv[o : o + cnt]
        o += cnt
    eidx = {ab: k for k, ab in enumerate(_TET_EDGES)}

    def edge(i, j):
        return eidx[(min(i, j), max(i, j))]

    face = np.zeros((len(tets), 4), dtype=np.int64)
    for r in range(len(tets)):
        u = uc[r]
        if nup[tets[r]] in (1, 3):
            lone = int(np.flatnonzero(u == (nup[tets[r]] == 1))[0])
            others = [q for q in range(4) if q != lone]
            c = [pid[r, edge(lone, q)] for q in others]
            face[r] = [c[0], c[1], c[2], c[2]]
        else:
            a, b = np.flatnonzero(u)
            c, dd = np.flatnonzero(~u)
            face[r] = [pid[r, edge(a, c)], pid[r, edge(a, dd)], pid[r, edge(b, dd)], pid[r, edge(b, c)]]
    nodeid = np.column_stack([uniq + 1, w + 1.0])
    return cutpos, face + 1, tets + 1, nodeid


def meshcentroid(node, face):
    """iso2mesh.meshcentroid (utility.py:482): mean of each element's vertices (1-based face)."""
    f = np.asarray(face).astype(np.int64) - 1
    return np.asarray(node, dtype=float)[f].mean(axis=1)


def rotatevec3d(pts, v):
    """iso2mesh.rotatevec3d(pts, v) (utility.py:485-486): rotate so that direction v maps onto +z.

    Rodrigues rotation about v x z by the angle between v and z (a half turn
    about x when v points along -z).  The reference only uses the rotated
    x, y coordinates for a point-in-polygon test and projections onto the
    rotated polygon's edges, which every rotation taking v to +z leaves
    unchanged.
    """
    v = np.asarray(v, dtype=float)[:3]
    v = v / np.linalg.norm(v)
    z = np.array([0.0, 0.0, 1.0])
    ax = np.cross(v, z)
    s, c = np.linalg.norm(ax), float(v @ z)
    if s < 1e-15:
        R = np.eye(3) if c > 0 else np.diag([1.0, -1.0, -1.0])
    else:
        k = ax / s
        K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
        R = np.eye(3) + s * K + (1 - c) * (K @ K)
    return np.asarray(pts, dtype=float) @ R.T


def inpolygon(x, y, px, py):
    """utility.py:639-669 -- even-odd ray casting (the reference's fallback branch)."""
    x = np.asarray(x, dtype=float)
    y = np.asarray(y, dtype=float)
    inside = np.zeros(x.shape, dtype=bool)
    for i in range(len(px) - 1):
        x1, y1, x2, y2 = px[i], py[i], px[i + 1], py[i + 1]
        hit = ((y1 <= y) & (y < y2)) | ((y2 <= y) & (y < y1))
        if np.any(hit):
            xc = (x2 - x1) * (y[hit] - y1) / (y2 - y1 + 1e-30) + x1
            inside[hit] ^= x[hit] < xc
    return inside


def _planar_source(cfg, srcpos, srcdir, p1, p2, srctype, pattern, z0):
    """utility.py:423-577 -- illuminated faces and per-pattern face weights."""
    p1 = np.asarray(p1, dtype=float)
    p2 = np.asarray(p2, dtype=float)
    srcpos = np.asarray(srcpos, dtype=float)[:3]
    srcdir = np.asarray(srcdir, dtype=float)[:3]
    poly = np.array([srcpos, srcpos + p1[:3], srcpos + p1[:3] + p2[:3], srcpos + p2[:3], srcpos])
    node = np.asarray(cfg["node"], dtype=float)
    nodeid = None
    sink = cfg.get("iscolimated", True)
    if sink:
        plane = np.concatenate([srcdir, [-(srcdir @ (srcpos + srcdir * z0))]])
        pnode, fd, _, nodeid = qmeshcut(cfg["elem"], node, plane)
        tri = fd[:, 2] == fd[:, 3]
        q = np.flatnonzero(~tri)
        pface = np.vstack([fd[tri, :3], fd[q][:, [0, 1, 2]], fd[q][:, [0, 2, 3]]]).astype(np.int64)
        parea = elemvolume(pnode, pface)
    else:
        pnode, pface, parea = node, np.asarray(cfg["face"]).astype(np.int64), np.asarray(cfg["area"])
    c0 = meshcentroid(pnode, pface)
    rot = rotatevec3d(np.vstack([c0, poly]), srcdir)
    spoly = rot[-5:, :2]
    cen = rot[:-5, :2]
    idx = np.flatnonzero(inpolygon(cen[:, 0], cen[:, 1], spoly[:, 0], spoly[:, 1]))
    if len(idx) == 0:
        raise ValueError("Source direction does not intersect with the domain")
    f0 = pface - 1
    nrm = np.cross(pnode[f0[idx, 1]] - pnode[f0[idx, 0]], pnode[f0[idx, 2]] - pnode[f0[idx, 0]])
    dd = nrm @ srcdir
    if sink:
        dd = -np.abs(dd)
    if np.all(dd >= 0):
        raise ValueError("Please reorient the surface triangles")
    good = idx[dd < 0]
    dp = cen[good] - spoly[0]
    dx, dy = spoly[1] - spoly[0], spoly[3] - spoly[0]
    lx, ly = np.linalg.norm(dx), np.linalg.norm(dy)
    bary = np.clip(np.column_stack([dp @ (dx / lx) / lx, dp @ (dy / ly) / ly]), 0, 1 - 1e-6)
    npat = 1
    bc = np.zeros((1, len(pface)))
    bc[0, good] = 1.0
    if pattern is not None and srctype == "pattern":
        pat = np.asarray(pattern, dtype=float)
        if pat.ndim == 2:
            pat = pat[:, :, None]
        npat = pat.shape[2]
        bc = np.zeros((npat, len(pface)))
        ix = np.clip((bary[:, 0] * pat.shape[1]).astype(int), 0, pat.shape[1] - 1)
        iy = np.clip((bary[:, 1] * pat.shape[0]).astype(int), 0, pat.shape[0] - 1)
        for i in range(npat):
            bc[i, good] = pat[iy, ix, i]
    elif srctype == "fourier":
        kx, ky = int(p1[3]), int(p2[3])
        phi0 = (p1[3] - kx) * 2 * np.pi
        M = 1 - (p2[3] - ky)
        npat = kx * ky
        bc = np.zeros((npat, len(pface)))
        for i in range(kx):
            for j in range(ky):
                bc[i * ky + j, good] = 0.5 * (1 + M * np.cos((i * bary[:, 0] + j * bary[:, 1]) * 2 * np.pi + phi0))
    return bc, npat, pface, parea, pnode, nodeid, sink


def _bc_weighting(cfg, bc, Reff, weight, pface, parea, nodeid, sink):
    """utility.py:580-636 -- boundary flux to nodal values, L1-normalised per pattern; (npat, nn)."""
    nn = np.asarray(cfg["node"]).shape[0]
    npat = bc.shape[0]
    wA = (parea * ((1 - Reff) / (18 * (1 + Reff))))[:, None] * bc.T
    rhs = np.zeros((nn, npat))
    f0 = pface - 1
    if nodeid is not None and sink:
        w1 = nodeid[:, 2] - 1.0
        ends = nodeid[:, :2].astype(np.int64) - 1
        for i in range(npat):
            for j in range(3):
                v = f0[:, j]
                np.add.at(rhs[:, i], ends[v, 0], wA[:, i] * w1[v])
                np.add.at(rhs[:, i], ends[v, 1], wA[:, i] * (1 - w1[v]))
    else:
        for i in range(npat):
            for j in range(3):
                np.add.at(rhs[:, i], f0[:, j], wA[:, i])
    for i in range(npat):
        ws = 1.0
        if weight is not None:
            if isinstance(weight, (int, float)):
                ws = weight
            elif len(weight) == npat:
                ws = weight[i]
        tot = np.sum(np.abs(rhs[:, i]))
        if tot > 0:
            rhs[:, i] *= ws / tot
    return rhs.T


def src2bc(cfg, isdet=False):
    """utility.py:252-420 -- wide-field sources (or detectors) as nodal boundary-flux columns.

    Adds cfg['widesrc'] (nn x npattern, patterns of every wavelength stacked),
    cfg['wfsrcmapping'] and drops the wide-field entries from cfg['srcpos']
    (or the det* counterparts).
    """
    if not isdet:
        tk, pk, dk, a1, a2, patk, wk, idk, widk, outk, mapk = (
            "srctype", "srcpos", "srcdir", "srcparam1", "srcparam2", "srcpattern", "srcweight", "srcid",
            "widesrcid", "widesrc", "wfsrcmapping")
    else:
        tk, pk, dk, a1, a2, patk, wk, idk, widk, outk, mapk = (
            "dettype", "detpos", "detdir", "detparam1", "detparam2", "detpattern", "detweight", "detid",
            "widedetid", "widedet", "wfdetmapping")
    stype = cfg.get(tk, "pencil")
    if stype in ("pencil", "isotropic") and widk not in cfg:
        return cfg
    dirs = np.atleast_2d(cfg[dk])
    pos = np.atleast_2d(cfg[pk])
    if widk in cfg:
        wid = cfg[widk] if isinstance(cfg[widk], dict) else {"": cfg[widk]}
    else:
        one = {"srctype": [stype], "srcid": [cfg.get(idk, 0)], "srcparam1": [np.atleast_1d(cfg[a1])],
               "srcparam2": [np.atleast_1d(cfg[a2])]}
        if patk in cfg:
            one["srcpattern"] = [cfg[patk]]
        if wk in cfg:
            one["srcweight"] = [cfg[wk]]
        wid = {"": one}
    prop = cfg["prop"] if isinstance(cfg.get("prop"), dict) else {"": cfg["prop"]}
    if set(wid) != set(prop) and "" in wid:
        wid = {w: wid[""] for w in prop}
    cols, mapping, ids = [], {}, []
    for wv in prop:
        par = wid[wv]
        op = np.asarray(prop[wv])
        z0 = 1.0 / (op[0] + op[1] * (1 - op[2])) if op.ndim == 1 else 1.0 / (op[1, 0] + op[1, 1] * (1 - op[1, 2]))
        mp = []
        for k in range(len(par["srcid"])):
            sid = par["srcid"][k]
            ids.append(sid)
            st = par["srctype"][k]
            pat = par["srcpattern"][k] if (st == "pattern" and "srcpattern" in par) else None
            wgt = par["srcweight"][k] if "srcweight" in par else None
            sdir = dirs[0, :3] if dirs.shape[0] == 1 else dirs[sid, :3]
            if st not in ("planar", "pattern", "fourier"):
                raise ValueError("Source type '%s' is not supported" % st)
            bc, npat, pface, parea, pnode, nodeid, sink = _planar_source(
                cfg, pos[sid, :3], sdir, par["srcparam1"][k], par["srcparam2"][k], st, pat, z0)
            reff = cfg["reff"][wv] if isinstance(cfg["reff"], dict) else cfg["reff"]
            rhs = _bc_weighting(cfg, bc, reff, wgt, pface, parea, nodeid, sink)
            mp.append([sid, len(cols), len(cols) + npat - 1])
            cols.append(rhs)
        mapping[wv] = np.array(mp) if mp else np.array([])
    wide = np.vstack(cols) if cols else np.array([])
    keep = np.ones(pos.shape[0], dtype=bool)
    keep[list(set(ids))] = False
    if len(mapping) == 1:
        mapping = mapping[list(prop)[0]]
    cfg[outk] = wide.T if wide.size else np.array([])
    cfg[mapk] = mapping
    cfg[pk] = pos[keep]
    return cfg


def meshprep(cfg):
    """utility.py:101-246 -- derived mesh fields (fallback branch, no iso2mesh)."""
    cfg = forcearray(
        cfg,
        ["node", "elem", "face", "srcpos", "srcdir", "detpos", "detdir", "prop", "seg", "widesrc", "widedet"],
    )
    if "node" not in cfg or "elem" not in cfg:
        raise ValueError("cfg must contain 'node' and 'elem'")
    node = cfg["node"]
    elem = cfg["elem"]
    if elem.shape[1] > 4 and cfg.get("seg") is None:
        cfg["seg"] = elem[:, 4].astype(int)
    if not cfg.get("isreoriented", False):
        ro = meshreorient(node, elem[:, :4])
        cfg["elem"] = np.column_stack([ro, elem[:, 4:]]) if elem.shape[1] > 4 else ro
        cfg["isreoriented"] = True
    if cfg.get("face") is None:
        cfg["face"] = volface(cfg["elem"][:, :4])
    if cfg.get("area") is None:
        cfg["area"] = elemvolume(node, cfg["face"])
    if cfg.get("evol") is None:
        cfg["evol"] = elemvolume(node, cfg["elem"][:, :4])
    if np.any(cfg["evol"] == 0):
        raise ValueError("Degenerate elements detected at indices: %s" % np.where(cfg["evol"] == 0)[0])
    if cfg.get("nvol") is None:
        cfg["nvol"] = nodevolume(node, cfg["elem"][:, :4], cfg["evol"])
    if "srcpos" not in cfg:
        raise ValueError("cfg.srcpos is required")
    if "srcdir" not in cfg:
        raise ValueError("cfg.srcdir is required")
    if isinstance(cfg.get("prop"), dict) and "param" in cfg:
        cfg["prop"] = updateprop(cfg)
    if cfg.get("reff") is None:
        bk = getbulk(cfg)
        if isinstance(bk, dict):
            cfg["reff"] = {w: getreff(p[3], 1.0) for w, p in bk.items()}
            cfg["musp0"] = {w: p[1] * (1 - p[2]) for w, p in bk.items()}
        else:
            cfg["reff"] = getreff(bk[3], 1.0)
            cfg["musp0"] = bk[1] * (1 - bk[2])
    if (cfg.get("srctype", "pencil") not in ("pencil", "isotropic") or "widesrcid" in cfg) and "widesrc" not in cfg:
        cfg["srcpos0"] = cfg["srcpos"].copy()
        cfg = src2bc(cfg, isdet=False)
    if (cfg.get("dettype", "pencil") not in ("pencil", "isotropic") or "widedetid" in cfg) and "widedet" not in cfg:
        cfg["detpos0"] = cfg["detpos"].copy()
        cfg = src2bc(cfg, isdet=True)
    if cfg.get("cols") is None:
        cfg["rows"], cfg["cols"], cfg["idxcount"] = femnz(cfg["elem"][:, :4], node.shape[0])
    if cfg.get("idxsum") is None:
        cfg["idxsum"] = np.cumsum(cfg["idxcount"])
    if cfg.get("deldotdel") is None:
        cfg["deldotdel"], _ = deldotdel(cfg)
    if "omega" not in cfg:
        cfg["omega"] = 0
    return cfg, sdmap(cfg)


def tsearchn(node, elem, pts, tol=1e-10):
    """Point location (iso2mesh.tsearchn, called at forward.py:316, :337).

    Returns (loc 1-based float with NaN when outside, bary (npts, 4)).  The
    lowest-index containing tet is returned; on a shared face or edge every
    containing tet gives the same P1 interpolant, so the RHS does not depend
    on this choice.
    """
    e0 = elem[:, :4].astype(np.int64) - 1
    p = node[e0]  # (ne,4,3)
    T = np.stack([p[:, 1] - p[:, 0], p[:, 2] - p[:, 0], p[:, 3] - p[:, 0]], axis=2)  # (ne,3,3)
    Tinv = np.linalg.inv(T)
    pts = np.atleast_2d(pts)
    loc = np.full(pts.shape[0], np.nan)
    bary = np.full((pts.shape[0], 4), np.nan)
    for k, x in enumerate(pts):
        lam = np.einsum("eij,ej->ei", Tinv, x[None, :3] - p[:, 0])
        l0 = 1.0 - lam.sum(axis=1)
        allb = np.column_stack([l0, lam])
        ok = np.all(allb >= -tol, axis=1)
        hit = np.flatnonzero(ok)
        if hit.size:
            e = hit[0]
            loc[k] = e + 1
            bary[k] = allb[e]
    return loc, bary


# ----------------------------------------------------------------------------
# property.py
# ----------------------------------------------------------------------------


def chromophore_tables():
    """property.py:439-577 -- built-in extinction tables (wavelength, value)."""
    hb = np.loadtxt(_DATA, delimiter=",", comments="#")
    conv = 2.303e-7
    tab = {
        "hbo": np.column_stack([hb[:, 0], hb[:, 1] * conv]),
        "hbr": np.column_stack([hb[:, 0], hb[:, 2] * conv]),
    }
    wwv = np.array([400, 500, 600, 650, 700, 750, 800, 850, 900, 950, 1000], dtype=float)
    wmu = np.array([0.00058, 0.00025, 0.0023, 0.0032, 0.006, 0.026, 0.02, 0.043, 0.068, 0.39, 0.36]) * 0.1
    tab["water"] = np.column_stack([wwv, wmu])
    lwv = np.arange(650, 1000, 10).astype(float)
    tab["lipids"] = np.column_stack([lwv, 0.0005 * np.ones_like(lwv)])
    awv = np.arange(650, 950, 5).astype(float)
    tab["aa3"] = np.column_stack([awv, 0.5 * np.exp(-((awv - 830) ** 2) / (2 * 50 ** 2)) + 0.4])
    return tab


def _interp_extrap(x, y, q):
    """scipy interp1d(kind='linear', fill_value='extrapolate') restated."""
    q = np.asarray(q, dtype=float)
    idx = np.clip(np.searchsorted(x, q, side="left") - 1, 0, len(x) - 2)
    # interp1d picks the interval with x[i] < q <= x[i+1] (searchsorted left)
    x0 = x[idx]
    x1 = x[idx + 1]
    return y[idx] + (q - x0) * (y[idx + 1] - y[idx]) / (x1 - x0)


def extinction(wavelengths, chromophores):
    """property.py:30-89 -- (Nwv, Nchrome) extinction, linear with extrapolation."""
    tab = chromophore_tables()
    if isinstance(wavelengths, (list, tuple)):
        wavelengths = [float(w) if isinstance(w, str) else w for w in wavelengths]
    w = np.atleast_1d(wavelengths).astype(float)
    if isinstance(chromophores, str):
        chromophores = [chromophores]
    out = np.zeros((len(w), len(chromophores)))
    for j, c in enumerate(chromophores):
        key = c.lower()
        if key not in tab:
            raise ValueError("Unknown chromophore: %s" % c)
        out[:, j] = _interp_extrap(tab[key][:, 0], tab[key][:, 1], w)
    return out, tab


def updateprop(cfg, wv=None, extin=None):
    """property.py:92-213 -- chromophores -> per-wavelength [mua, musp, g, n].

    ``extin`` (oracle of this project's extension): a dict wavelength ->
    {chromophore: value}.  When given, the table lookup is replaced and no
    default water/lipids are injected.
    """
    if "param" not in cfg or not isinstance(cfg.get("prop"), dict):
        return cfg.get("prop")
    wvs = list(cfg["prop"].keys()) if wv is None else [wv]
    par = cfg["param"]
    out = {}
    for w in wvs:
        if extin is None:
            if "water" not in par:
                par["water"] = 0.23
            if "lipids" not in par:
                par["lipids"] = 0.58
        types = [t for t in ["hbo", "hbr", "water", "lipids", "aa3"] if t in par]
        if not types:
            raise ValueError("No recognized chromophores in cfg.param")
        if extin is None:
            ex, _ = extinction(float(w), types)
            ex = ex[0]
        else:
            ex = np.array([extin[w][t] for t in types])
        first = par[types[0]]
        mua = 0.0 if np.isscalar(first) else np.zeros_like(first, dtype=float)
        for j, t in enumerate(types):
            mua = mua + ex[j] * par[t]
        musp = None
        if "scatamp" in par and "scatpow" in par:
            musp = par["scatamp"] * float(w) ** (-par["scatpow"])
        seg = cfg["prop"][w]
        nn = cfg["node"].shape[0]
        ne = cfg["elem"].shape[0]
        mlen = len(mua) if hasattr(mua, "__len__") else 1
        if mlen < min(nn, ne):
            newp = seg.copy()
            if hasattr(mua, "__len__"):
                newp[1 : mlen + 1, 0] = mua
                if musp is not None:
                    newp[1 : mlen + 1, 1] = musp
                    newp[1 : mlen + 1, 2] = 0
            else:
                newp[1, 0] = mua
                if musp is not None:
                    newp[1, 1] = musp
                    newp[1, 2] = 0
        else:
            if musp is not None:
                nr = seg[1, 3] if seg.shape[0] > 1 else 1.37
                newp = np.column_stack([mua, musp, np.zeros_like(musp), np.full_like(musp, nr)])
            else:
                newp = np.column_stack([mua, np.tile(seg[1, 1:], (len(mua), 1))])
        out[w] = newp
    return out if len(wvs) > 1 else out[wvs[0]]


def _bulk_from_prop(prop, cfg):
    """property.py:264-332."""
    dflt = np.array([0.0, 0.0, 0.0, 1.37])
    nn = cfg["node"].shape[0]
    ne = cfg["elem"].shape[0]
    if prop.shape[0] < min(nn, ne):
        if "seg" in cfg and cfg["seg"] is not None:
            seg = cfg["seg"]
            if "face" in cfg and len(seg) == nn:
                lab = seg[cfg["face"][0, 0] - 1]
            elif "face" in cfg and len(seg) == ne:
                e0 = cfg["elem"][:, :4].astype(int) - 1
                eid = np.where(np.any(e0 == cfg["face"][0, 0] - 1, axis=1))[0]
                lab = seg[eid[0]] if len(eid) else 0
            else:
                lab = seg[0]
            lab = int(lab)
            if prop.shape[0] > lab:
                return prop[lab, :]
            return prop[1, :] if prop.shape[0] > 1 else dflt
        return prop[1, :] if prop.shape[0] > 1 else dflt
    if prop.shape[0] == nn:
        if "face" in cfg:
            return prop[cfg["face"][0, 0] - 1, :]
        return prop[0, :]
    if prop.shape[0] == ne:
        if "face" in cfg:
            e0 = cfg["elem"][:, :4].astype(int) - 1
            eid = np.where(np.any(e0 == cfg["face"][0, 0] - 1, axis=1))[0]
            if len(eid):
                return prop[eid[0], :]
        return prop[0, :]
    return dflt


def getbulk(cfg):
    """property.py:216-261."""
    if "bulk" in cfg:
        b = cfg["bulk"]
        p = np.array([0.0, 0.0, 0.0, 1.37])
        if "mua" in b:
            p[0] = b["mua"]
        if "dcoeff" in b:
            p[1] = 1.0 / (3 * b["dcoeff"])
            p[2] = 0
        if "musp" in b:
            p[1] = b["musp"]
            p[2] = 0
        if "g" in b:
            p[2] = b["g"]
        if "n" in b:
            p[3] = b["n"]
        return p
    if cfg.get("prop") is None:
        return np.array([0.0, 0.0, 0.0, 1.37])
    if isinstance(cfg["prop"], dict):
        return {w: _bulk_from_prop(p, cfg) for w, p in cfg["prop"].items()}
    return _bulk_from_prop(cfg["prop"], cfg)


# ----------------------------------------------------------------------------
# forward.py
# ----------------------------------------------------------------------------


def _props_for(cfg, wavelength, mode):
    """forward.py:105-124 -- props, Reff and omega for one wavelength/mode."""
    if isinstance(cfg.get("prop"), dict) and wavelength:
        props = cfg["prop"][wavelength]
        reff = cfg["reff"][wavelength] if isinstance(cfg.get("reff"), dict) else cfg["reff"]
        om = cfg["omega"].get(wavelength, 0) if isinstance(cfg.get("omega"), dict) else cfg.get("omega", 0)
    else:
        props = cfg["prop"]
        reff = cfg.get("reff", 0.493)
        om = cfg.get("omega", 0)
    if mode == 2:
        om = 0
    return props, reff, om


def femlhs(cfg, deldotdel_mat, wavelength="", mode=1):
    """forward.py:90-256 -- FEM system matrix, COO triplets summed into CSR."""
    nn = cfg["node"].shape[0]
    ne = cfg["elem"].shape[0]
    evol = cfg["evol"]
    area = cfg["area"]
    e0 = cfg["elem"][:, :4].astype(np.int64) - 1
    f0 = cfg["face"].astype(np.int64) - 1
    props, reff, om = _props_for(cfg, wavelength, mode)
    seg = cfg.get("seg", None)
    if props.shape[0] == nn or props.shape[0] == ne:
        mua = props[:, 0]
        musp = props[:, 1] * (1 - props[:, 2]) if props.shape[1] >= 3 else props[:, 1]
        nref = props[:, 3] if props.shape[1] >= 4 else 1.37
    elif seg is not None:
        s = np.clip(seg.astype(np.int64), 0, props.shape[0] - 1)
        mua = props[s, 0]
        musp = props[s, 1] * (1 - props[s, 2]) if props.shape[1] >= 3 else props[s, 1]
        nref = props[s[0], 3] if props.shape[1] >= 4 else 1.37
    else:
        raise ValueError("Property format not recognized")
    D = 1.0 / (3.0 * (mua + musp))
    fd = om > 0
    R, C, V = [], [], []
    pairs = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    offc = [1, 2, 3, 5, 6, 8]
    diagc = [0, 4, 7, 9]
    if len(mua) == ne:
        for k, (i, j) in enumerate(pairs):
            v = deldotdel_mat[:, offc[k]] * D + 0.05 * mua * evol
            if fd:
                v = v + 1j * (0.05 * om * R_C0 * nref * evol)
            R += [e0[:, i], e0[:, j]]
            C += [e0[:, j], e0[:, i]]
            V += [v, v]
        for k in range(4):
            v = deldotdel_mat[:, diagc[k]] * D + 0.10 * mua * evol
            if fd:
                v = v + 1j * (0.10 * om * R_C0 * nref * evol)
            R.append(e0[:, k])
            C.append(e0[:, k])
            V.append(v)
    else:
        # exact P1 triple-product weights for nodal coefficients
        w_off = np.array([[2, 2, 1, 1], [2, 1, 2, 1], [2, 1, 1, 2], [1, 2, 2, 1], [1, 2, 1, 2], [1, 1, 2, 2]]).T / 120.0
        w_dia = (np.diag([2.0, 2, 2, 2]) + 1) / 60.0
        mua_e = mua[e0]
        D_e = np.mean(D[e0], axis=1)
        nref_e = nref[e0] if hasattr(nref, "__len__") and len(nref) == nn else nref
        for k, (i, j) in enumerate(pairs):
            v = deldotdel_mat[:, offc[k]] * D_e + (mua_e @ w_off[:, k]) * evol
            if fd:
                if hasattr(nref_e, "__len__"):
                    v = v + 1j * om * R_C0 * (nref_e @ w_off[:, k]) * evol
                else:
                    v = v + 1j * om * R_C0 * nref_e * 0.05 * evol
            R += [e0[:, i], e0[:, j]]
            C += [e0[:, j], e0[:, i]]
            V += [v, v]
        for k in range(4):
            v = deldotdel_mat[:, diagc[k]] * D_e + (mua_e @ w_dia[:, k]) * evol
            if fd:
                if hasattr(nref_e, "__len__"):
                    v = v + 1j * om * R_C0 * (nref_e @ w_dia[:, k]) * evol
                else:
                    v = v + 1j * om * R_C0 * nref_e * 0.10 * evol
            R.append(e0[:, k])
            C.append(e0[:, k])
            V.append(v)
    bc = (1 - reff) / (12.0 * (1 + reff))
    adiag = area * bc
    aoff = adiag * 0.5
    for i, j in [(0, 1), (0, 2), (1, 2)]:
        R += [f0[:, i], f0[:, j]]
        C += [f0[:, j], f0[:, i]]
        V += [aoff, aoff]
    for k in range(3):
        R.append(f0[:, k])
        C.append(f0[:, k])
        V.append(adiag)
    dt = complex if fd else float
    return sparse.coo_matrix(
        (np.concatenate(V), (np.concatenate(R), np.concatenate(C))), shape=(nn, nn), dtype=dt
    ).tocsr()


def femrhs(cfg, sd=None, wv="", md=1):
    """forward.py:259-362 -- RHS columns [point src | wide src | point det | wide det]."""
    src, det, wsrc, wdet = getoptodes(cfg, wv)
    ns = src.shape[0] if src is not None and src.size else 0
    nd = det.shape[0] if det is not None and det.size else 0
    ws = wsrc.shape[1] if wsrc is not None and wsrc.size else 0
    wd = wdet.shape[1] if wdet is not None and wdet.size else 0
    nn = cfg["node"].shape[0]
    tot = ns + ws + nd + wd
    if tot == 0:
        return sparse.csr_matrix((nn, 0)), np.array([]), np.array([]).reshape(0, 4), np.array([])
    rhs = sparse.lil_matrix((nn, tot))
    loc = np.full(tot, np.nan)
    bary = np.full((tot, 4), np.nan)
    elem = cfg["elem"][:, :4].astype(np.int64)
    e0 = elem - 1
    c = 0
    if ns:
        l, b = tsearchn(cfg["node"], elem, src[:, :3])
        for i in range(ns):
            if not np.isnan(l[i]):
                rhs[e0[int(l[i]) - 1, :], c + i] = b[i, :]
        loc[:ns] = l
        bary[:ns] = b
        c += ns
    if ws:
        rhs[:, c : c + ws] = wsrc
        c += ws
    if nd:
        l, b = tsearchn(cfg["node"], elem, det[:, :3])
        for i in range(nd):
            if not np.isnan(l[i]):
                rhs[e0[int(l[i]) - 1, :], c + i] = b[i, :]
        loc[ns + ws : ns + ws + nd] = l
        bary[ns + ws : ns + ws + nd] = b
        c += nd
    if wd:
        rhs[:, c : c + wd] = wdet
    opt = [x for x in (src if ns else None, det if nd else None) if x is not None]
    optode = np.vstack(opt) if opt else np.array([])
    return rhs.tocsr(), loc, bary, optode


def femsolve(Amat, rhs, method="auto", **kw):
    """solver.py:385-588, 941-948 -- default path on this container: SuperLU.

    Zero RHS columns are skipped and stay zero (solver.py:569-581).
    """
    if sparse.issparse(rhs):
        rhs = rhs.toarray()
    oned = rhs.ndim == 1
    if oned:
        rhs = rhs.reshape(-1, 1)
    n, k = rhs.shape
    cplx = np.iscomplexobj(Amat) or np.iscomplexobj(rhs)
    x = np.zeros((n, k), dtype=complex if cplx else float)
    A = Amat.tocsc()
    nz = [i for i in range(k) if np.any(rhs[:, i] != 0)]
    if nz:
        lu = splu(A)
        sol = lu.solve(np.ascontiguousarray(rhs[:, nz].astype(x.dtype)))
        x[:, nz] = sol.reshape(n, len(nz))
    return (x.ravel() if oned else x), 0


def femgetdet(phi, cfg, rhs, loc=None, bary=None):
    """forward.py:365-442 -- detval = rhs[:, det cols]^T @ phi[:, src cols]."""
    ns, nd, ws, wd = _counts(cfg)
    ts = ns + ws
    td = nd + wd
    if ts == 0 or td == 0:
        return np.array([])
    rd = rhs[:, ts : ts + td]
    rd = rd.toarray() if sparse.issparse(rd) else rd
    return rd.T @ phi[:, :ts]


def runforward(cfg, **kw):
    """forward.py:40-87 -- loop wavelengths x modes: femrhs, femlhs, solve, femgetdet."""
    rfcw = kw.get("rfcw", [1])
    if isinstance(rfcw, int):
        rfcw = [rfcw]
    if cfg.get("deldotdel") is None:
        cfg["deldotdel"], _ = deldotdel(cfg)
    wvs = list(cfg["prop"].keys()) if isinstance(cfg.get("prop"), dict) else [""]
    sd = kw.get("sd")
    if sd is None:
        sd = sdmap(cfg)
    if not isinstance(sd, dict):
        sd = {w: sd for w in wvs}
    det = {m: {"detphi": {}} for m in rfcw}
    phi = {m: {"phi": {}} for m in rfcw}
    for w in wvs:
        for m in rfcw:
            rhs, loc, bary, _ = femrhs(cfg, sd, w, m)
            A = femlhs(cfg, cfg["deldotdel"], w, m)
            x, _ = femsolve(A, rhs)
            phi[m]["phi"][w] = x
            det[m]["detphi"][w] = femgetdet(x, cfg, rhs, loc, bary)
    if len(wvs) == 1:
        for m in rfcw:
            phi[m]["phi"] = phi[m]["phi"][wvs[0]]
            det[m]["detphi"] = det[m]["detphi"][wvs[0]]
    if len(rfcw) == 1:
        phi = phi[rfcw[0]]["phi"]
        det = det[rfcw[0]]["detphi"]
    return det, phi


def jac(sd, phi, deldotdel_mat, elem, evol, iselem=False):
    """forward.py:509-569 (numpy branch) -- mua Jacobian, element then node.

    Jelem[p, e] = -(sum_i ps_i pd_i + 0.5 sum_{i!=j} ps_i pd_j) * 0.1 * V_e;
    Jnode = 0.25 * incidence @ Jelem^T, transposed back.
    """
    e0 = elem[:, :4].astype(np.int64) - 1
    ne = e0.shape[0]
    nn = phi.shape[0]
    if sd.shape[1] >= 3:
        act = sd[sd[:, 2] == 1, :2].astype(np.int64)
    else:
        act = sd[:, :2].astype(np.int64)
    Je = np.zeros((act.shape[0], ne), dtype=phi.dtype)
    v01 = 0.1 * evol
    for p, (s, d) in enumerate(act):
        ps = phi[e0, s]
        pd = phi[e0, d]
        dg = (ps * pd).sum(axis=1)
        cr = np.zeros(ne, dtype=phi.dtype)
        for i in range(4):
            for j in range(4):
                if i != j:
                    cr = cr + ps[:, i] * pd[:, j]
        Je[p] = -(dg + cr * 0.5) * v01
    P = sparse.csr_matrix((np.full(ne * 4, 0.25), (e0.ravel(), np.repeat(np.arange(ne), 4))), shape=(nn, ne))
    Jn = (P @ Je.T).T
    return Jn, Je


def jacchrome(Jmua, chromophores, extin=None):
    """forward.py:572-589 -- stack E[w, c] * Jmua[w] over wavelengths per chromophore.

    ``extin`` (project extension): (Nwv, Nchrome) matrix replacing the table.
    """
    if not isinstance(Jmua, dict):
        raise ValueError("Jmua must be a dict with wavelength keys")
    wvs = list(Jmua.keys())
    E = extinction(wvs, chromophores)[0] if extin is None else np.asarray(extin)
    out = {}
    for i, c in enumerate(chromophores):
        out[c] = np.vstack([Jmua[w] * E[j, i] for j, w in enumerate(wvs)])
    return out


# ----------------------------------------------------------------------------
# recon.py
# ----------------------------------------------------------------------------


def normalize_diag(A):
    """recon.py:784-789 -- scale to unit diagonal."""
    di = 1.0 / np.sqrt(np.abs(np.diag(A)) + 1e-16)
    return (di[:, None] * di[None, :]) * A, di


def reginvover(A, rhs, lam, LTL=None, blocks=None):
    """recon.py:363-426 -- (A^T A + lam (I | L^T L)) x = A^T rhs; an L^T L smaller than the
    Hessian is added to each diagonal block of its size (recon.py:404-412)."""
    cs = np.sum(np.abs(A), axis=0)
    keep = np.where(cs != 0)[0]
    n0 = A.shape[1]
    if len(keep) < n0:
        A = A[:, keep]
        if LTL is not None and LTL.shape[0] > len(keep):
            li = keep[keep < LTL.shape[0]]
            LTL = LTL[np.ix_(li, li)]
    rs = np.sum(np.abs(A), axis=1)
    good = rs != 0
    if good.sum() < A.shape[0]:
        A = A[good]
        rhs = rhs[good]
    g = A.T @ rhs.ravel()
    H = A.T @ A
    if LTL is None:
        H[np.diag_indices_from(H)] += lam
    elif H.shape[0] == LTL.shape[0]:
        H = H + lam * LTL
    else:
        nx = LTL.shape[0]
        for i in range(0, H.shape[0], nx):
            e = min(i + nx, H.shape[0])
            H[i:e, i:e] = H[i:e, i:e] + lam * LTL[: e - i, : e - i]
    Hn, di = normalize_diag(H)
    res = di * np.linalg.solve(Hn, di * g)
    if len(keep) < n0:
        full = np.zeros(n0, dtype=res.dtype)
        full[keep] = res
        res = full
    return res


def reginvunder(A, rhs, lam, invR=None, blocks=None):
    """recon.py:429-509 -- x = R^-1 A'^T (A' A'^T + lam I)^-1 rhs with A' = A R^-1 (R^-1 = invR),
    applied to the whole A when sizes match, else to each unknown block of invR's size
    (recon.py:458-470, :487-499); no transform without invR."""
    n0 = A.shape[1]
    cs = np.sum(np.abs(A), axis=0)
    keep = np.where(cs != 0)[0]
    if len(keep) < n0:
        A = A[:, keep]
    rs = np.sum(np.abs(A), axis=1)
    good = rs != 0
    if good.sum() < A.shape[0]:
        A = A[good]
        rhs = rhs[good]
    spans = []  # column ranges transformed by invR
    if invR is not None:
        nx = invR.shape[0]
        if nx == A.shape[1]:
            spans = [(0, nx)]
        elif blocks is not None:
            cum = np.cumsum([0] + [v[1] for v in blocks.values()])
            spans = [(cum[i], cum[i + 1]) for i in range(len(cum) - 1) if cum[i + 1] - cum[i] == nx]
        A = A.copy()
        for a, b in spans:
            A[:, a:b] = A[:, a:b] @ invR
    rhs = rhs.ravel()
    H = A @ A.T
    H[np.diag_indices_from(H)] += lam
    Hn, di = normalize_diag(H)
    y = di * np.linalg.solve(Hn, di * rhs)
    res = A.T @ y
    for a, b in spans:
        res[a:b] = invR @ res[a:b]
    if len(keep) < n0:
        full = np.zeros(n0, dtype=res.dtype)
        full[keep] = res
        res = full
    return res


def masksum(J, mask):
    """recon.py:879-893 -- sum the columns of J over each label of mask (labels ascending)."""
    labels = np.unique(mask)
    out = np.zeros((J.shape[0], len(labels)), dtype=J.dtype)
    for i, lab in enumerate(labels):
        out[:, i] = J[:, mask == lab].sum(axis=1)
    return out


def reginv(A, rhs, lam, Areg=None, blocks=None):
    """recon.py:338-360 -- over- or under-determined branch by shape."""
    Areg = Areg or {}
    if A.shape[0] >= A.shape[1]:
        return reginvover(A, rhs, lam, Areg.get("ltl"), blocks)
    return reginvunder(A, rhs, lam, Areg.get("lir"), blocks)


def matreform(A, ymeas, ymodel, form="complex"):
    """recon.py:512-567."""
    rhs = ymeas - ymodel
    if form == "complex":
        return A, rhs, 1
    if form in ("real", "reim"):
        if not np.isreal(rhs).all() and not np.isreal(A).all():
            if form == "reim":
                nA = np.block([[np.real(A), -np.imag(A)], [np.imag(A), np.real(A)]])
            else:
                nA = np.vstack([np.real(A), np.imag(A)])
            return nA, np.concatenate([np.real(rhs), np.imag(rhs)]), 2
        return np.real(A), np.real(rhs), 1
    if form == "logphase":
        t = np.conj(ymodel) / np.abs(ymodel * ymodel)
        t = t[:, None] * A if A.ndim == 2 else t * A
        if np.isreal(ymodel).all():
            return np.real(t), np.log(np.abs(ymeas)) - np.log(np.abs(ymodel)), 1
        return (
            np.vstack([np.real(t), np.imag(t)]),
            np.concatenate([np.log(np.abs(ymeas)) - np.log(np.abs(ymodel)), np.angle(ymeas) - np.angle(ymodel)]),
            2,
        )
    raise ValueError("Unknown form: %s" % form)


def matflat(A, weight=None):
    """recon.py:570-593."""
    if isinstance(A, np.ndarray):
        return A
    keys = list(A.keys())
    if weight is None:
        weight = np.ones(len(keys))
    first = A[keys[0]]
    if isinstance(first, dict):
        inner = list(first.keys())
        return np.vstack([np.hstack([A[k][w] * weight[j] for j, k in enumerate(keys)]) for w in inner])
    return np.hstack([A[k] * weight[i] for i, k in enumerate(keys)])


def flatten_detphi(detphi, sd, wavelengths, rfcw):
    """recon.py:792-815."""
    if isinstance(detphi, np.ndarray):
        return detphi.flatten()
    out = []
    for w in wavelengths:
        pw = detphi.get(w, detphi) if isinstance(detphi, dict) else detphi
        if isinstance(pw, dict):
            for m in rfcw:
                out.extend(pw.get(m, {}).get("detphi", pw).flatten())
        else:
            out.extend(np.asarray(pw).flatten())
    return np.array(out)


def meshinterp(fromval, elemid, elembary, fromelem, toval=None):
    """utility.py:973-999 -- nodal values of a source mesh at points located in it.

    elemid: 1-based element of each point (NaN: outside, left at toval or 0);
    elembary: (npts, 4) barycentric weights.
    """
    fv = np.asarray(fromval, dtype=float)
    if fv.ndim == 1:
        fv = fv[:, None]
    e0 = np.asarray(fromelem)[:, :4].astype(np.int64) - 1
    eid = np.asarray(elemid, dtype=float)
    out = np.zeros((len(eid), fv.shape[1])) if toval is None else np.array(toval, dtype=float).reshape(len(eid), -1)
    ok = ~np.isnan(eid)
    nodes = e0[eid[ok].astype(np.int64) - 1]
    out[ok] = np.einsum("pk,pkc->pc", np.asarray(elembary)[ok], fv[nodes])
    return out if out.shape[1] > 1 else out[:, 0]


def remap_jacobian(J, recon, cfg):
    """recon.py:818-876 -- forward-node Jacobian columns moved onto recon-mesh nodes.

    J_new[:, n] = sum over forward nodes i located in a recon element holding
    n of J[:, i] * mapweight[i, j] (the transpose of meshinterp), forward nodes
    taken in ascending order; NaN or out-of-range elements are skipped.
    """
    nrec = np.asarray(recon["node"]).shape[0]
    e0 = np.asarray(recon["elem"])[:, :4].astype(np.int64) - 1
    mid = np.asarray(recon["mapid"], dtype=float)
    mw = np.asarray(recon["mapweight"], dtype=float)
    out = np.zeros((J.shape[0], nrec), dtype=J.dtype)
    for i in range(J.shape[1]):
        if np.isnan(mid[i]):
            continue
        e = int(mid[i]) - 1
        if e < 0 or e >= e0.shape[0] or np.any(e0[e] < 0) or np.any(e0[e] >= nrec):
            continue
        for j in range(4):
            out[:, e0[e, j]] += J[:, i] * mw[i, j]
    return out


def _dual(recon):
    return "mapid" in recon and "mapweight" in recon and "elem" in recon


def syncprop(cfg, recon):
    """recon.py:669-781 -- recon unknowns onto the forward mesh (label copy, same-mesh copy, or dual-mesh
    interpolation with meshinterp over recon['mapid'] / ['mapweight'])."""
    if "param" in recon:
        keys = list(recon["param"].keys())
        first = recon["param"][keys[0]]
        plen = len(first) if hasattr(first, "__len__") else 1
        if plen < 50:
            cfg["param"] = {k: (v.copy() if hasattr(v, "copy") else v) for k, v in recon["param"].items()}
        else:
            cfg.setdefault("param", {})
            for k in keys:
                cfg["param"][k] = (meshinterp(recon["param"][k], recon["mapid"], recon["mapweight"], recon["elem"])
                                   if _dual(recon) else recon["param"][k].copy())
    elif "prop" in recon:
        if not isinstance(recon["prop"], dict):
            if recon["prop"].shape[0] >= 50 and _dual(recon):
                cfg["prop"] = meshinterp(recon["prop"], recon["mapid"], recon["mapweight"], recon["elem"])
            else:
                cfg["prop"] = recon["prop"].copy()
        else:
            first = next(iter(recon["prop"].values()))
            if first.shape[0] >= 50 and _dual(recon):
                cfg["prop"] = {k: meshinterp(v, recon["mapid"], recon["mapweight"], recon["elem"])
                               for k, v in recon["prop"].items()}
            else:
                cfg["prop"] = {k: v.copy() for k, v in recon["prop"].items()}
    return cfg, recon


def runrecon(cfg, recon, detphi0, sd=None, extin=None, **kw):
    """recon.py:39-335 -- Gauss-Newton loop (single or dual mesh; an explicit ``lmat`` / ``ltl``
    prior; label compression of the Jacobian by a 1-D recon['seg'] of fewer than 50 labels).

    ``extin`` (project extension): dict wavelength -> {chromophore: value},
    used by updateprop and jacchrome in place of the built-in table.
    """
    maxiter = kw.get("maxiter", 5)
    lam = kw.get("lambda_", recon.get("lambda", 0.05))
    reform = kw.get("reform", "real")
    rfcw = kw.get("rfcw", [1])
    tol = kw.get("tol", 0)
    if isinstance(rfcw, int):
        rfcw = [rfcw]
    if sd is None:
        sd = sdmap(cfg)
    if "prop" in recon and isinstance(recon["prop"], np.ndarray) and recon["prop"].ndim == 1:
        recon["prop"] = recon["prop"].reshape(1, -1)
    resid = np.zeros(maxiter)
    updates = []
    Aregu = {}  # recon.py:127-133 (an explicit lmat, else ltl; the prior() builder is product code)
    if "lmat" in kw:
        Aregu["lmat"] = kw["lmat"]
    elif "ltl" in kw:
        Aregu["ltl"] = kw["ltl"]
    for it in range(maxiter):
        if "param" in recon or "prop" in recon:
            cfg, recon = syncprop(cfg, recon)
        if "param" in cfg and isinstance(cfg.get("prop"), dict):
            cfg["prop"] = updateprop(cfg, extin=extin)
        detphi, phi = runforward(cfg, sd=sd, rfcw=rfcw)
        wvs = list(cfg["prop"].keys()) if isinstance(cfg.get("prop"), dict) else [""]
        Jm = {}
        for w in wvs:
            sdw = sd.get(w, sd) if isinstance(sd, dict) else sd
            pw = phi.get(w, phi) if isinstance(phi, dict) else phi
            Jn, _ = jac(sdw, pw, cfg["deldotdel"], cfg["elem"], cfg["evol"])
            Jm["mua" if w == "" else w] = Jn
        if isinstance(cfg.get("prop"), dict) and "param" in cfg:
            chs = [k for k in cfg["param"].keys() if k in ["hbo", "hbr", "water", "lipids", "aa3"]]
            if chs:
                E = None
                if extin is not None:
                    E = np.array([[extin[w][c] for c in chs] for w in wvs])
                Jm = jacchrome(Jm, chs, extin=E)
        y0 = flatten_detphi(detphi0, sd, wvs, rfcw)
        y = flatten_detphi(detphi, sd, wvs, rfcw)
        blocks = {k: v.shape for k, v in Jm.items()}
        J = matflat(Jm)
        if reform != "complex":
            J, mis, _ = matreform(J, y0, y, reform)
        else:
            mis = y0 - y
        if "node" in recon and "elem" in recon and "mapid" in recon:  # recon.py:200-205
            J = remap_jacobian(J, recon, cfg)
            nrec = np.asarray(recon["node"]).shape[0]
            blocks = {k: (v[0], nrec) for k, v in blocks.items()}
        if "seg" in recon and np.ndim(recon["seg"]) == 1:  # recon.py:210-218
            seg = np.asarray(recon["seg"])
            nlab = len(np.unique(seg))
            if len(seg) == J.shape[1] and nlab < 50:
                J = masksum(J, seg)
                blocks = {k: (v[0], nlab) for k, v in blocks.items()}
        resid[it] = np.sum(np.abs(mis))
        if it == 0 and Aregu and "lmat" in Aregu and "ltl" not in Aregu:  # recon.py:225-233
            if J.shape[0] >= J.shape[1]:
                Aregu["ltl"] = Aregu["lmat"].T @ Aregu["lmat"]
            else:
                from scipy.linalg import qr

                _, R = qr(Aregu["lmat"])
                Aregu["lir"] = np.linalg.inv(np.triu(R))
        s = 1.0 / np.sqrt(np.sum(J ** 2))
        J = J * s
        dmu = reginv(J, mis, lam, Aregu, blocks) * s
        upd = {}
        idx = 0
        for key in blocks:
            n = blocks[key][1]
            dx = dmu[idx : idx + n]
            upd[key] = dx
            idx += n
            if key in ("mua", "dcoeff"):
                pi = 0 if key == "mua" else 1
                if "prop" in recon and isinstance(recon["prop"], np.ndarray):
                    pr = recon["prop"]
                    if pr.shape[0] < len(dx) and pr.shape[0] < 50:
                        for li in range(min(pr.shape[0], len(dx))):
                            if key == "dcoeff":
                                recon["prop"][li, pi] = 1.0 / (3 * (1.0 / (3 * pr[li, pi]) + dx[li]))
                            else:
                                recon["prop"][li, pi] += dx[li]
                    else:
                        if key == "dcoeff":
                            recon["prop"][:, pi] = 1.0 / (3 * (1.0 / (3 * pr[:, pi]) + dx))
                        else:
                            recon["prop"][:, pi] += dx
            elif key in ("hbo", "hbr", "water", "lipids", "scatamp", "scatpow"):
                if "param" in recon and key in recon["param"]:
                    tgt = recon
                elif "param" in cfg and key in cfg["param"]:
                    tgt = cfg
                else:
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
        updates.append(upd)
        if it > 0 and abs(resid[it] - resid[it - 1]) / resid[0] < tol:
            resid = resid[: it + 1]
            break
    recon["lambda"] = lam
    return recon, resid, cfg, updates, Jm, detphi, phi
