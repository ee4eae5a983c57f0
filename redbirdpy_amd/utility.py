"""Mesh and optode utilities (drop-in for redbirdpy.utility on the hot path).

elem/face stay 1-based in cfg, as in redbirdpy (utility.py:1-6).  The CSR
connectivity (``rows``/``cols``/``idxcount``) is produced by the GPU pattern
kernel; the remaining geometry is vectorised host code, as in the reference.
"""

__all__ = [
    "meshprep",
    "meshprep_batch",
    "deldotdel",
    "sdmap",
    "getoptodes",
    "getdistance",
    "getltr",
    "getreff",
    "elem2node",
    "forcearray",
    "src2bc",
    "meshinterp",
    "addnoise",
    "HAS_ISO2MESH",
]

import warnings

import ctypes

import numpy as np
import torch

# iso2mesh is not used: plane cuts, point location and boundary faces are
# computed here and on the GPU (the flag keeps the reference's name)
HAS_ISO2MESH = False

_ARRAY_KEYS = ["node", "elem", "face", "srcpos", "srcdir", "detpos", "detdir", "prop", "seg", "widesrc", "widedet"]


def forcearray(cfg, keys):
    """List-valued entries become numpy arrays (utility.py:1112-1117)."""
    for key in keys:
        if key in cfg and isinstance(cfg[key], list):
            cfg[key] = np.array(cfg[key])
    return cfg


def _tet_points(node, elem):
    e0 = np.asarray(elem)[:, :4].astype(np.int64) - 1
    return node[e0]  # (ne, 4, 3)


def _signed_volume6(node, elem):
    p = _tet_points(node, elem)
    a = p[:, 1] - p[:, 0]
    b = p[:, 2] - p[:, 0]
    c = p[:, 3] - p[:, 0]
    return np.einsum("ij,ij->i", np.cross(a, b), c)


def reorient(node, elem):
    """Swap local nodes 0 and 1 of tets with negative signed volume (utility.py:1047-1063)."""
    out = np.array(elem, copy=True)
    neg = _signed_volume6(node, elem) < 0
    out[neg, 0], out[neg, 1] = elem[neg, 1], elem[neg, 0]
    return out


def boundary_faces(elem):
    """Faces that belong to a single tet, outward-ordered, 1-based (utility.py:1066-1087).

    Local faces (0,2,1), (0,1,3), (0,3,2), (1,2,3) are stacked face-type-major;
    the result is ordered lexicographically by the sorted vertex triple.
    """
    e0 = np.asarray(elem)[:, :4].astype(np.int64) - 1
    loc = np.array([[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]])
    tri = np.concatenate([e0[:, loc[k]] for k in range(4)], axis=0)
    key = np.sort(tri, axis=1)
    order = np.lexsort((key[:, 2], key[:, 1], key[:, 0]))
    ks = key[order]
    newgrp = np.ones(len(ks), dtype=bool)
    newgrp[1:] = np.any(ks[1:] != ks[:-1], axis=1)
    start = np.flatnonzero(newgrp)
    cnt = np.diff(np.append(start, len(ks)))
    single = start[cnt == 1]
    return tri[order[single]] + 1


def elemvolume(node, elem):
    """|det|/6 for tets, half cross-product norm for triangles (utility.py:1090-1109)."""
    e = np.asarray(elem)
    if e.shape[1] >= 4:
        return np.abs(_signed_volume6(node, e)) / 6.0
    if e.shape[1] == 3:
        e0 = e.astype(np.int64) - 1
        cr = np.cross(node[e0[:, 1]] - node[e0[:, 0]], node[e0[:, 2]] - node[e0[:, 0]])
        return 0.5 * np.sqrt(np.einsum("ij,ij->i", cr, cr))
    raise ValueError("Unsupported element type with %d nodes" % e.shape[1])


def nodevolume(nn, elem, evol):
    """A quarter of each incident tet's volume (utility.py:1005-1017)."""
    e0 = np.asarray(elem)[:, :4].astype(np.int64) - 1
    nv = np.zeros(nn)
    for j in range(4):
        nv += np.bincount(e0[:, j], weights=evol, minlength=nn)
    return nv * 0.25


def deldotdel(cfg):
    """V * grad(phi_i) . grad(phi_j) for the 10 upper-triangle pairs (utility.py:53-98).

    Gradients are the cofactors of the vertex coordinates divided by 6V.
    Returns (ne x 10, delphi 3 x 4 x ne).
    """
    node = cfg["node"]
    evol = np.asarray(cfg["evol"], dtype=float)
    p = _tet_points(node, cfg["elem"]).transpose(2, 1, 0)  # (3, 4, ne)
    ne = p.shape[2]
    cof = ((3, 1, 2, 1), (2, 0, 3, 2), (1, 3, 0, 3), (0, 2, 1, 0))
    g = np.empty((3, 4, ne))
    scale = 1.0 / (evol * 6.0)
    for c in range(3):
        u, v = [x for x in range(3) if x != c]
        for i, (a, b, cc, d) in enumerate(cof):
            g[c, i] = ((p[u, a] - p[u, b]) * (p[v, cc] - p[v, d]) - (p[u, cc] - p[u, d]) * (p[v, a] - p[v, b])) * scale
    out = np.empty((ne, 10))
    k = 0
    for i in range(4):
        for j in range(i, 4):
            out[:, k] = np.einsum("cn,cn->n", g[:, i], g[:, j])
            k += 1
    out *= evol[:, None]
    return out, g


def getreff(n_in, n_out=1.0):
    """Effective reflection coefficient, Haskell et al. 1994 (utility.py:879-904).

    Riemann sums of step pi/2000 over the Fresnel reflectance, as in the reference.
    """
    if n_in <= n_out:
        return 0.0
    step = np.pi / 2000
    crit = np.arcsin(n_out / n_in)
    th = np.arange(0, crit, step)
    ci = np.cos(th)
    ct = np.sqrt(1 - (n_in * np.sin(th)) ** 2)
    rperp = ((n_in * ct - n_out * ci) / (n_in * ct + n_out * ci)) ** 2
    rpar = ((n_in * ci - n_out * ct) / (n_in * ci + n_out * ct)) ** 2
    full = np.arange(0, np.pi / 2, step)
    R = np.ones(len(full))
    R[: len(th)] = 0.5 * rperp + 0.5 * rpar
    s = np.sin(full)
    c = np.cos(full)
    r_phi = 2 * np.sum(s * c * R) * step
    r_j = 3 * np.sum(s * c * c * R) * step
    return (r_phi + r_j) / (2 - r_phi + r_j)


def getltr(cfg, wv=""):
    """Transport mean free path 1/(mua + mus(1-g)) of the bulk (utility.py:862-876)."""
    from .property import getbulk

    bk = getbulk(cfg)
    if isinstance(bk, dict):
        bk = bk[wv if wv else list(bk.keys())[0]]
    return 1.0 / (bk[0] + bk[1] * (1 - bk[2]))


def _tile_dir(d, n):
    d = np.atleast_2d(np.asarray(d, dtype=float))
    return np.tile(d, (n, 1)) if d.shape[0] == 1 else d


def getoptodes(cfg, wv=""):
    """Point optodes moved inward by dir * l_tr; wide-field patterns as given (utility.py:799-822)."""
    ltr = getltr(cfg, wv)
    src = det = None
    if cfg.get("srcpos") is not None and np.size(cfg["srcpos"]) > 0:
        sp = np.atleast_2d(cfg["srcpos"]).astype(float)
        src = sp + _tile_dir(cfg["srcdir"], sp.shape[0]) * ltr
    if cfg.get("detpos") is not None and np.size(cfg["detpos"]) > 0:
        dp = np.atleast_2d(cfg["detpos"]).astype(float)
        det = dp + _tile_dir(cfg.get("detdir", cfg["srcdir"]), dp.shape[0]) * ltr
    return src, det, cfg.get("widesrc", None), cfg.get("widedet", None)


def optode_counts(cfg):
    """(point src, point det, wide src, wide det) counts as femrhs/femgetdet count them."""

    def npos(key):
        v = cfg.get(key)
        if v is None:
            return 0
        v = np.atleast_2d(v)
        return v.shape[0] if v.size > 0 else 0

    def nwide(key):
        v = cfg.get(key)
        return v.shape[1] if v is not None and np.size(v) > 0 else 0

    return npos("srcpos"), npos("detpos"), nwide("widesrc"), nwide("widedet")


def getdistance(srcpos, detpos, badsrc=None, baddet=None, widesrc=None, widedet=None):
    """(Ndet x Nsrc) source-detector distances, inf for excluded ones (utility.py:825-859)."""
    badsrc = badsrc or []
    baddet = baddet or []
    srcpos = np.atleast_2d(srcpos)
    detpos = np.atleast_2d(detpos)
    ns, nd = srcpos.shape[0], detpos.shape[0]
    ws = widesrc.shape[0] if widesrc is not None else 0
    wd = widedet.shape[0] if widedet is not None else 0
    dist = np.full((nd + wd, ns + ws), np.inf)
    gs = sorted(set(range(ns)) - set(badsrc))
    gd = sorted(set(range(nd)) - set(baddet))
    if gs and gd:
        diff = srcpos[gs, None, :3].transpose(1, 0, 2) - detpos[gd, None, :3]
        dist[np.ix_(gd, gs)] = np.sqrt(np.sum(diff ** 2, axis=2))
    return dist


def sdmap(cfg, maxdist=np.inf, **kwargs):
    """Source-detector pair table, row = d*Nsrc + s (utility.py:675-796).

    Columns: [src column, det column (after all sources), active(, mode)];
    multi-wavelength cfgs get a dict of 4-column tables.
    """
    ns, nd, ws, wd = optode_counts(cfg)
    if ns + ws == 0 or nd + wd == 0:
        raise ValueError("Must define at least one source and detector")
    bads = list(kwargs.get("excludesrc", cfg.get("excludesrc", [])))
    badd = list(kwargs.get("excludedet", cfg.get("excludedet", [])))
    goods = sorted(set(range(ns)) - set(bads))
    goodd = sorted(set(range(nd)) - set(badd))
    off = ns + ws
    srcs = goods + list(range(ns, ns + ws))
    dets = [off + d for d in goodd] + [off + d for d in range(nd, nd + wd)]
    ss, dd = np.meshgrid(srcs, dets)
    pairs = np.column_stack([ss.ravel(), dd.ravel()])
    act = np.ones(len(pairs))
    if maxdist < np.inf and ns > 0 and nd > 0:
        dist = getdistance(cfg["srcpos"], cfg["detpos"], bads, baddet=badd)
        for r, (s, d) in enumerate(pairs):
            di = int(d) - off
            if s < ns and di < nd:
                act[r] = 1.0 if dist[di, int(s)] < maxdist else 0.0
    if isinstance(cfg.get("prop"), dict):
        return {wv: np.column_stack([pairs, act, np.ones(len(pairs))]) for wv in cfg["prop"]}
    return np.column_stack([pairs, act])


def elem2node(elem, elemval, nodelen=None):
    """Average element values onto nodes with weight 1/4 per incident tet (utility.py:910-929)."""
    if isinstance(elem, dict):
        nodelen = elem["node"].shape[0]
        elem = elem["elem"]
    e0 = np.asarray(elem)[:, :4].astype(np.int64) - 1
    v = np.asarray(elemval, dtype=float)
    v2 = v[:, None] if v.ndim == 1 else v
    out = np.zeros((nodelen, v2.shape[1]))
    for j in range(4):
        np.add.at(out, e0[:, j], v2)
    return (out * 0.25).squeeze()


def _connectivity(cfg):
    """rows/cols/idxcount of the off-diagonal node adjacency from the GPU CSR pattern."""
    from . import _gpu

    mesh = _gpu.mesh_for(cfg)
    indptr, indices, _ = mesh.pattern_host()
    rows = np.repeat(np.arange(mesh.nn), np.diff(indptr))
    off = indices != rows
    return rows[off], indices[off], np.bincount(rows[off], minlength=mesh.nn)


# ---------------------------------------------------------------------------
# wide-field sources and detectors (utility.py:252-669)
# ---------------------------------------------------------------------------
def _rotation_to_z(v):
    """3x3 rotation taking direction v onto +z (about v x z; a half turn about x for v along -z)."""
    v = np.asarray(v, dtype=float)[:3]
    v = v / np.linalg.norm(v)
    ax = np.array([v[1], -v[0], 0.0])  # v x z
    s, c = np.linalg.norm(ax), v[2]
    if s < 1e-15:
        return np.eye(3) if c > 0 else np.diag([1.0, -1.0, -1.0])
    k = ax / s
    K = np.array([[0.0, -k[2], k[1]], [k[2], 0.0, -k[0]], [-k[1], k[0], 0.0]])
    return np.eye(3) + s * K + (1 - c) * (K @ K)


def _wide_columns(cfg, srcpos, srcdir, p1, p2, stype, pattern, z0, reff, weight):
    """(npat, nn) normalised columns of one planar / pattern / Fourier source (utility.py:423-636).

    The geometry (sink-plane cut or boundary faces, rectangle test, pattern
    lookup, flux split over the cut edges) runs on the GPU (rb_widefield);
    the per-pattern L1 normalisation follows (utility.py:622-634).
    """
    from . import _gpu, _abi

    p1 = np.asarray(p1, dtype=float)
    p2 = np.asarray(p2, dtype=float)
    srcpos = np.asarray(srcpos, dtype=float)[:3]
    srcdir = np.asarray(srcdir, dtype=float)[:3]
    corners = np.array([srcpos, srcpos + p1[:3], srcpos + p1[:3] + p2[:3], srcpos + p2[:3], srcpos])
    R = _rotation_to_z(srcdir)
    poly = corners @ R[:2].T
    ex, ey = poly[1] - poly[0], poly[3] - poly[0]
    lx, ly = np.linalg.norm(ex), np.linalg.norm(ey)
    sink = cfg.get("iscolimated", True)
    plane = np.concatenate([srcdir, [-(srcdir @ (srcpos + srcdir * z0))]]) if sink else np.zeros(4)
    kind, npat, kx, ky, M, phi0, img = _abi.RB_WF_UNIFORM, 1, 0, 0, 0.0, 0.0, None
    if stype == "pattern" and pattern is not None:
        img = np.asarray(pattern, dtype=float)
        if img.ndim == 2:
            img = img[:, :, None]
        kind, npat = _abi.RB_WF_IMAGE, img.shape[2]
    elif stype == "fourier":
        kx, ky = int(p1[3]), int(p2[3])
        phi0 = (p1[3] - kx) * 2 * np.pi
        M = 1 - (p2[3] - ky)
        kind, npat = _abi.RB_WF_FOURIER, kx * ky
    nn = np.asarray(cfg["node"]).shape[0]
    if npat == 0:
        return np.zeros((0, nn))
    frame = np.concatenate([srcdir, R[0], R[1], poly[:, 0], poly[:, 1], poly[0], ex / lx, ey / ly, [lx, ly],
                            plane, [(1 - reff) / (18 * (1 + reff)), M, phi0]])
    mesh = _gpu.mesh_for(cfg)
    rhs, counts = _gpu.widefield(mesh, _abi.RB_WF_SINK if sink else _abi.RB_WF_FACES, frame, kind, npat, kx, ky,
                                 img, area=cfg.get("area"))
    if counts[0] == 0:
        raise ValueError("Source direction does not intersect with the domain")
    if counts[1] == 0:
        raise ValueError("Please reorient the surface triangles")
    ws = np.ones(npat)
    if weight is not None:
        if isinstance(weight, (int, float)):
            ws[:] = weight
        elif len(weight) == npat:
            ws = np.asarray(weight, dtype=float)
    import torch

    tot = rhs.abs().sum(0)
    one = torch.ones_like(tot)
    wst = torch.as_tensor(ws, dtype=rhs.dtype, device=rhs.device)
    rhs = rhs * torch.where(tot > 0, wst / torch.where(tot > 0, tot, one), one)
    return mesh.user_rows(rhs).cpu().numpy().T


def src2bc(cfg, isdet=False):
    """Wide-field sources (or detectors) as nodal boundary-flux columns (utility.py:252-420).

    Adds cfg['widesrc'] (nn x npattern; the patterns of every wavelength,
    stacked), cfg['wfsrcmapping'] ([srcid, first, last] column per source,
    per wavelength) and removes the wide-field entries from cfg['srcpos'];
    isdet=True does the same for the det* fields.  Types: planar, pattern
    (cfg['srcpattern'], ny x nx [x npattern]) and fourier (kx, ky in the 4th
    entries of srcparam1/2, integer part; phase and depth in the fraction).
    A collimated source (cfg['iscolimated'], default True) is sunk to 1/mus'
    below the surface along its direction.
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
    blocks, mapping, used = [], {}, []
    ncol = 0
    for wv in prop:
        par = wid[wv]
        op = np.asarray(prop[wv])
        # sink depth 1/mus' of the first label row (of the single row for 1-D props)
        z0 = 1.0 / (op[0] + op[1] * (1 - op[2])) if op.ndim == 1 else 1.0 / (op[1, 0] + op[1, 1] * (1 - op[1, 2]))
        rows = []
        for k in range(len(par["srcid"])):
            sid = par["srcid"][k]
            used.append(sid)
            st = par["srctype"][k]
            if st not in ("planar", "pattern", "fourier"):
                raise ValueError("Source type '%s' is not supported" % st)
            pat = par["srcpattern"][k] if (st == "pattern" and "srcpattern" in par) else None
            wgt = par["srcweight"][k] if "srcweight" in par else None
            sdir = dirs[0, :3] if dirs.shape[0] == 1 else dirs[sid, :3]
            reff = cfg["reff"][wv] if isinstance(cfg["reff"], dict) else cfg["reff"]
            cols = _wide_columns(cfg, pos[sid, :3], sdir, par["srcparam1"][k], par["srcparam2"][k], st, pat, z0,
                                 reff, wgt)
            rows.append([sid, ncol, ncol + cols.shape[0] - 1])
            ncol += cols.shape[0]
            blocks.append(cols)
        mapping[wv] = np.array(rows) if rows else np.array([])
    wide = np.vstack(blocks) if blocks else np.array([])
    keep = np.ones(pos.shape[0], dtype=bool)
    keep[list(set(used))] = False
    if len(mapping) == 1:
        mapping = mapping[list(prop)[0]]
    cfg[outk] = wide.T if wide.size else np.array([])
    cfg[mapk] = mapping
    cfg[pk] = pos[keep]
    return cfg


def addnoise(data, snrshot, snrthermal=np.inf, randseed=123456789):
    """Shot and thermal noise at the given SNRs in dB (utility.py:932-970).

    Shot noise scales with sqrt(|data|) (sigma 10^(-snr/20)), thermal noise
    with the largest amplitude; complex data also get phase noise from the
    imaginary parts of the SNRs.  The draws come from numpy's global
    generator seeded with randseed, in the reference's order, so a seed
    gives the reference's noise.
    """
    np.random.seed(randseed)
    data = np.asarray(data)
    if np.isinf(snrshot) and np.isinf(snrthermal):
        warnings.warn("No noise added")
        return data.copy()
    s_shot = 10 ** (-np.real(snrshot) / 20)
    s_therm = np.max(np.abs(data)) * 10 ** (-np.real(snrthermal) / 20)
    if np.isreal(data).all():
        out = data + np.sqrt(np.abs(data)) * np.random.randn(*data.shape) * s_shot
        return out + np.random.randn(*data.shape) * s_therm
    ph_shot = 10 ** (-np.imag(snrshot) / 20)
    ph_therm = 10 ** (-np.imag(snrthermal) / 20)
    a1 = np.random.randn(*data.shape) * s_shot
    p1 = np.random.randn(*data.shape) * ph_shot * 2 * np.pi
    a2 = np.random.randn(*data.shape) * s_therm
    p2 = np.random.randn(*data.shape) * ph_therm * 2 * np.pi
    return data + np.sqrt(np.abs(data)) * (a1 * np.exp(1j * p1)) + a2 * np.exp(1j * p2)


def meshinterp(fromval, elemid, elembary, fromelem, toval=None):
    """Nodal values of a source mesh at located points (utility.py:973-999).

    elemid: 1-based source element per point (NaN: outside, keeps toval or
    0); elembary: (npts, 4) barycentric weights; fromval: (nn_src[, ncol]).
    """
    fv = np.asarray(fromval, dtype=float)
    fv = fv[:, None] if fv.ndim == 1 else fv
    e0 = np.asarray(fromelem)[:, :4].astype(np.int64) - 1
    eid = np.asarray(elemid, dtype=float)
    if toval is None:
        out = np.zeros((len(eid), fv.shape[1]))
    else:
        out = np.array(toval, dtype=float).reshape(len(eid), -1)
    inside = np.flatnonzero(~np.isnan(eid))
    corners = e0[eid[inside].astype(np.int64) - 1]  # (npts, 4) source nodes
    w = np.asarray(elembary, dtype=float)[inside]
    acc = np.zeros((len(inside), fv.shape[1]))
    for j in range(4):
        acc += w[:, j : j + 1] * fv[corners[:, j]]
    out[inside] = acc
    return out if out.shape[1] > 1 else out[:, 0]


_MESH_FIELDS = ("isreoriented", "face", "area", "evol", "nvol", "cols", "idxsum", "deldotdel")


def meshprep_batch(cfgs):
    """meshprep for many meshes (utility.py:101-246), their geometry in one GPU launch sequence.

    The meshes are concatenated with node offsets (faces never cross meshes)
    and reoriented, measured and given their deldotdel (rb_tet_prep), node
    incidence, boundary faces in the reference's order (rb_boundary_faces),
    face areas, node volumes and CSR pattern on the device; the results are
    split back per mesh.  The rest of meshprep (properties, Reff, wide-field
    sources, the sd map) runs per mesh as in meshprep.  A cfg that already
    carries mesh fields goes through meshprep itself.  Returns [(cfg, sd)].
    """
    from . import _gpu
    from ._abi import call, ptr, stream_ptr

    cfgs = [forcearray(c, _ARRAY_KEYS) for c in cfgs]
    fresh = [i for i, c in enumerate(cfgs) if not any(c.get(k) is not None and c.get(k) is not False
                                                       for k in _MESH_FIELDS)]
    out = [None] * len(cfgs)
    for i, c in enumerate(cfgs):
        if "node" not in c or "elem" not in c:
            raise ValueError("cfg must contain 'node' and 'elem'")
        if i not in fresh:
            out[i] = meshprep(c)
    if not fresh:
        return out
    nodes, elems, noff, eoff = [], [], [0], [0]
    for i in fresh:
        c = cfgs[i]
        if c["elem"].shape[1] > 4 and c.get("seg") is None:
            c["seg"] = c["elem"][:, 4].astype(int)
        nodes.append(np.asarray(c["node"], dtype=np.float64))
        elems.append(np.asarray(c["elem"])[:, :4].astype(np.int64) - 1 + noff[-1])
        noff.append(noff[-1] + nodes[-1].shape[0])
        eoff.append(eoff[-1] + elems[-1].shape[0])
    nn, ne = noff[-1], eoff[-1]
    dev = _gpu.device()
    s = stream_ptr()
    nd = _gpu.h2d(np.vstack(nodes), dev)
    el_in = _gpu.h2d(np.vstack(elems).astype(np.int32), dev)
    el = torch.empty_like(el_in)
    evol = torch.empty(ne, dtype=torch.float64, device=dev)
    dd = torch.empty((ne, 10), dtype=torch.float64, device=dev)
    call("rb_tet_prep", nn, ne, ptr(nd), ptr(el_in), ptr(el), ptr(evol), ptr(dd), s)
    tinc_ptr = torch.empty(nn + 1, dtype=torch.int32, device=dev)
    tinc = torch.empty(4 * ne, dtype=torch.int32, device=dev)
    call("rb_incidence", ptr(el), ne, 4, nn, ptr(tinc_ptr), ptr(tinc), s)
    face_ptr = torch.empty(nn + 1, dtype=torch.int32, device=dev)
    nf = ctypes.c_int64(0)
    call("rb_boundary_faces", nn, ptr(el), ptr(tinc_ptr), ptr(tinc), ptr(face_ptr), None, ctypes.byref(nf), s)
    faces = torch.empty((max(nf.value, 1), 3), dtype=torch.int32, device=dev)
    call("rb_boundary_faces", nn, ptr(el), ptr(tinc_ptr), ptr(tinc), ptr(face_ptr), ptr(faces), ctypes.byref(nf), s)
    area = torch.empty(max(nf.value, 1), dtype=torch.float64, device=dev)
    call("rb_face_area", nf.value, ptr(nd), ptr(faces), ptr(area), s)
    nvol = torch.empty(nn, dtype=torch.float64, device=dev)
    call("rb_node_volume", nn, ptr(evol), ptr(tinc_ptr), ptr(tinc), ptr(nvol), s)
    row_ptr = torch.empty(nn + 1, dtype=torch.int32, device=dev)
    nnz = ctypes.c_int64(0)
    call("rb_csr_pattern", ptr(el), ne, nn, ptr(tinc_ptr), ptr(tinc), ptr(row_ptr), None, 0, ctypes.byref(nnz), None,
         s)
    col_idx = torch.empty(max(nnz.value, 1), dtype=torch.int32, device=dev)
    call("rb_csr_pattern", ptr(el), ne, nn, ptr(tinc_ptr), ptr(tinc), ptr(row_ptr), ptr(col_idx), nnz.value,
         ctypes.byref(nnz), None, s)
    el_h, evol_h, dd_h = el.cpu().numpy(), evol.cpu().numpy(), dd.cpu().numpy()
    fp_h, fc_h, ar_h = face_ptr.cpu().numpy(), faces[: nf.value].cpu().numpy(), area[: nf.value].cpu().numpy()
    nv_h, rp_h, ci_h = nvol.cpu().numpy(), row_ptr.cpu().numpy(), col_idx[: nnz.value].cpu().numpy()
    for j, i in enumerate(fresh):
        c = cfgs[i]
        n0, n1, e0, e1 = noff[j], noff[j + 1], eoff[j], eoff[j + 1]
        ro = (el_h[e0:e1] - n0 + 1).astype(c["elem"].dtype if np.issubdtype(c["elem"].dtype, np.integer) else np.int64)
        c["elem"] = np.column_stack([ro, c["elem"][:, 4:]]) if c["elem"].shape[1] > 4 else ro
        c["isreoriented"] = True
        ev = evol_h[e0:e1].copy()
        if np.any(ev == 0):
            raise ValueError("Degenerate elements detected at indices: %s" % np.flatnonzero(ev == 0))
        c["evol"] = ev
        f0, f1 = fp_h[n0], fp_h[n1]
        c["face"] = fc_h[f0:f1].astype(np.int64) - n0 + 1
        c["area"] = ar_h[f0:f1].copy()
        c["nvol"] = nv_h[n0:n1].copy()
        rp = rp_h[n0 : n1 + 1].astype(np.int64) - rp_h[n0]
        ci = ci_h[rp_h[n0] : rp_h[n1]].astype(np.int64) - n0
        rows = np.repeat(np.arange(n1 - n0), np.diff(rp))
        off = ci != rows
        c["rows"], c["cols"] = rows[off], ci[off]
        c["idxcount"] = np.bincount(rows[off], minlength=n1 - n0)
        c["deldotdel"] = dd_h[e0:e1].copy()
        out[i] = meshprep(c)  # properties, Reff, wide-field sources, idxsum, omega and the sd map
    return out


def meshprep(cfg):
    """Derived mesh fields and the sd map (utility.py:101-246).

    Adds reoriented elem, face, area, evol, nvol, reff, musp0, rows/cols/
    idxcount/idxsum, deldotdel and omega to cfg (in place) and returns
    (cfg, sd).
    """
    from .property import updateprop, getbulk

    cfg = forcearray(cfg, _ARRAY_KEYS)
    if "node" not in cfg or "elem" not in cfg:
        raise ValueError("cfg must contain 'node' and 'elem'")
    node = cfg["node"]
    elem = cfg["elem"]
    if elem.shape[1] > 4 and cfg.get("seg") is None:
        cfg["seg"] = elem[:, 4].astype(int)
    if not cfg.get("isreoriented", False):
        ro = reorient(node, elem[:, :4])
        cfg["elem"] = np.column_stack([ro, elem[:, 4:]]) if elem.shape[1] > 4 else ro
        cfg["isreoriented"] = True
    if cfg.get("face") is None:
        cfg["face"] = boundary_faces(cfg["elem"][:, :4])
    if cfg.get("area") is None:
        cfg["area"] = elemvolume(node, cfg["face"])
    if cfg.get("evol") is None:
        cfg["evol"] = elemvolume(node, cfg["elem"][:, :4])
    if np.any(cfg["evol"] == 0):
        raise ValueError("Degenerate elements detected at indices: %s" % np.flatnonzero(cfg["evol"] == 0))
    if cfg.get("nvol") is None:
        cfg["nvol"] = nodevolume(node.shape[0], cfg["elem"], cfg["evol"])
    if "srcpos" not in cfg:
        raise ValueError("cfg.srcpos is required")
    if "srcdir" not in cfg:
        raise ValueError("cfg.srcdir is required")
    if isinstance(cfg.get("prop"), dict) and "param" in cfg:
        cfg["prop"] = updateprop(cfg)
    if cfg.get("reff") is None:
        bk = getbulk(cfg)
        if isinstance(bk, dict):
            cfg["reff"] = {wv: getreff(p[3], 1.0) for wv, p in bk.items()}
            cfg["musp0"] = {wv: p[1] * (1 - p[2]) for wv, p in bk.items()}
        else:
            cfg["reff"] = getreff(bk[3], 1.0)
            cfg["musp0"] = bk[1] * (1 - bk[2])
    # wide-field sources / detectors become boundary-flux columns (utility.py:210-224)
    if (cfg.get("srctype", "pencil") not in ("pencil", "isotropic") or "widesrcid" in cfg) and "widesrc" not in cfg:
        cfg["srcpos0"] = cfg["srcpos"].copy()
        cfg = src2bc(cfg, isdet=False)
    if (cfg.get("dettype", "pencil") not in ("pencil", "isotropic") or "widedetid" in cfg) and "widedet" not in cfg:
        cfg["detpos0"] = cfg["detpos"].copy()
        cfg = src2bc(cfg, isdet=True)
    if cfg.get("cols") is None:
        cfg["rows"], cfg["cols"], cfg["idxcount"] = _connectivity(cfg)
    if cfg.get("idxsum") is None:
        cfg["idxsum"] = np.cumsum(cfg["idxcount"])
    if cfg.get("deldotdel") is None:
        cfg["deldotdel"], _ = deldotdel(cfg)
    if "omega" not in cfg:
        cfg["omega"] = 0
    return cfg, sdmap(cfg)
