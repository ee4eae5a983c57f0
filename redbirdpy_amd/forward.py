"""FEM forward model on the GPU (drop-in for redbirdpy.forward).

``runforward`` assembles every (wavelength, mode) system in one assembly
launch, builds all right-hand sides in one dense device block
[point src | wide src | point det | wide det] per system, and solves every
column of every system in one multi-RHS Krylov solve; ``femgetdet`` and
``jac`` read the device-resident fields.  Public functions take and return
numpy/scipy objects like the reference; ``forward_device`` keeps results on
the GPU for the reconstruction loop and for autograd.
"""

__all__ = ["runforward", "femlhs", "femrhs", "femgetdet", "jac", "jacchrome", "C0", "forward_batch", "BatchForward"]

import numpy as np
import torch
from scipy import sparse

from . import _gpu
from .utility import sdmap, getoptodes, deldotdel, optode_counts
from .property import extinction

C0 = 299792458000.0  # speed of light in mm/s (forward.py:36)
R_C0 = 1.0 / C0

last_info = None


def femlhs(cfg, deldotdel_mat, wavelength="", mode=1):
    """FEM system matrix as scipy CSR, assembled on the GPU (forward.py:90-256)."""
    mesh = _gpu.mesh_for(cfg)
    sysm = _gpu.assemble(mesh, cfg, deldotdel_mat, [(wavelength, mode)])
    indptr, indices, _ = mesh.pattern_host()
    data = mesh.user_values(sysm.vals64[0]).cpu().numpy()
    return sparse.csr_matrix((data, indices, indptr), shape=(mesh.nn, mesh.nn))


def _locate(mesh, pts):
    """GPU point location; returns (loc 1-based float with NaN, bary with NaN)."""
    if pts is None or np.size(pts) == 0:
        return np.zeros(0), np.zeros((0, 4))
    l0, b = _gpu.tsearch(mesh, np.atleast_2d(pts)[:, :3])
    loc = np.where(l0 >= 0, l0 + 1, np.nan).astype(float)
    b = np.where((l0 >= 0)[:, None], b, np.nan)
    return loc, b


class _Optodes:
    """Point-optode locations and wide-field patterns of one wavelength."""

    def __init__(self, cfg, mesh, wv):
        src, det, wsrc, wdet = getoptodes(cfg, wv)
        self.ns, self.nd, self.ws, self.wd = optode_counts(cfg)
        self.src = np.atleast_2d(src) if self.ns else None
        self.det = np.atleast_2d(det) if self.nd else None
        self.wsrc = np.asarray(wsrc) if self.ws else None
        self.wdet = np.asarray(wdet) if self.wd else None
        self.sloc, self.sbary = _locate(mesh, self.src)
        self.dloc, self.dbary = _locate(mesh, self.det)
        self.k = self.ns + self.ws + self.nd + self.wd

    def rhs_csr(self, cfg, nn):
        """RHS as scipy CSR (nn x k) in the reference's column order."""
        e0 = cfg["elem"][:, :4].astype(np.int64) - 1
        rows, cols, vals = [], [], []

        def points(loc, bary, c0):
            for i in range(len(loc)):
                if not np.isnan(loc[i]):
                    rows.append(e0[int(loc[i]) - 1])
                    cols.append(np.full(4, c0 + i))
                    vals.append(bary[i])

        def wide(w, c0):
            r, c = np.nonzero(w)
            rows.append(r)
            cols.append(c + c0)
            vals.append(w[r, c])

        points(self.sloc, self.sbary, 0)
        if self.ws:
            wide(self.wsrc, self.ns)
        points(self.dloc, self.dbary, self.ns + self.ws)
        if self.wd:
            wide(self.wdet, self.ns + self.ws + self.nd)
        if rows:
            A = sparse.coo_matrix(
                (np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(nn, self.k)
            ).tocsr()
            A.eliminate_zeros()
            return A
        return sparse.csr_matrix((nn, self.k))

    def fill_block(self, mesh, B, c0, is_complex):
        """Write this wavelength's columns into the device block B at column c0."""
        if self.ns:
            _gpu.rhs_block(mesh, B, c0, _loc0(self.sloc), self.sbary, is_complex)
        if self.ws:
            B[:, c0 + self.ns : c0 + self.ns + self.ws] = torch.as_tensor(mesh.mesh_nodes(self.wsrc),
                                                                          device=B.device).to(B.dtype)
        if self.nd:
            _gpu.rhs_block(mesh, B, c0 + self.ns + self.ws, _loc0(self.dloc), self.dbary, is_complex)
        if self.wd:
            c = c0 + self.ns + self.ws + self.nd
            B[:, c : c + self.wd] = torch.as_tensor(mesh.mesh_nodes(self.wdet), device=B.device).to(B.dtype)

    def det_columns(self, cfg, mesh):
        """Detector columns as device CSC (ptr, row, val) for rb_getdet, rows in device numbering."""
        ptrs, rows, vals = [0], [], []
        found = [i for i in range(self.nd) if not np.isnan(self.dloc[i])]
        tets = {}
        if found:  # device numbering of the located tets' nodes only
            t = np.array([int(self.dloc[i]) - 1 for i in found], dtype=np.int64)
            tets = dict(zip(found, mesh.iperm[np.asarray(cfg["elem"])[t, :4].astype(np.int64) - 1]))
        for i in range(self.nd):
            if i in tets:
                rows.append(tets[i])
                vals.append(self.dbary[i])
                ptrs.append(ptrs[-1] + 4)
            else:
                ptrs.append(ptrs[-1])
        for j in range(self.wd):
            r = np.flatnonzero(self.wdet[:, j])
            rows.append(mesh.iperm[r])
            vals.append(self.wdet[r, j])
            ptrs.append(ptrs[-1] + len(r))
        dev = mesh.dev
        row = np.concatenate(rows) if rows else np.zeros(0, np.int64)
        val = np.concatenate(vals) if vals else np.zeros(0)
        return (
            _gpu.h2d(np.array(ptrs), dev, np.int32),
            _gpu.h2d(row, dev, np.int32),
            _gpu.h2d(val, dev, np.float64),
        )


def _loc0(loc):
    return np.where(np.isnan(loc), -1, np.nan_to_num(loc, nan=0) - 1).astype(np.int64)


def femrhs(cfg, sd=None, wv="", md=1):
    """RHS columns [point src | wide src | point det | wide det] (forward.py:259-362).

    Returns (rhs csr, loc 1-based with NaN for wide-field, bary, optode).
    """
    mesh = _gpu.mesh_for(cfg)
    op = _Optodes(cfg, mesh, wv)
    nn = cfg["node"].shape[0]
    if op.k == 0:
        return sparse.csr_matrix((nn, 0)), np.array([]), np.array([]).reshape(0, 4), np.array([])
    loc = np.full(op.k, np.nan)
    bary = np.full((op.k, 4), np.nan)
    loc[: op.ns] = op.sloc
    bary[: op.ns] = op.sbary
    d0 = op.ns + op.ws
    loc[d0 : d0 + op.nd] = op.dloc
    bary[d0 : d0 + op.nd] = op.dbary
    opts = [x for x in (op.src, op.det) if x is not None]
    optode = np.vstack(opts) if opts else np.array([])
    return op.rhs_csr(cfg, nn), loc, bary, optode


def femgetdet(phi, cfg, rhs, loc=None, bary=None):
    """detval = rhs[:, det cols]^T phi[:, src cols], (ndet x nsrc) (forward.py:365-442)."""
    ns, nd, ws, wd = optode_counts(cfg)
    ts, td = ns + ws, nd + wd
    if ts == 0 or td == 0:
        return np.array([])
    R = sparse.csc_matrix(rhs)[:, ts : ts + td] if sparse.issparse(rhs) else sparse.csc_matrix(np.asarray(rhs)[:, ts : ts + td])
    R.sort_indices()
    dev = _gpu.device()
    phi_t = phi if torch.is_tensor(phi) else torch.as_tensor(np.ascontiguousarray(phi), device=dev)
    phi_t = phi_t.to(dev)
    is_complex = phi_t.is_complex() or np.iscomplexobj(R.data)
    phi_t = phi_t.to(torch.complex128 if is_complex else torch.float64).contiguous()
    if np.iscomplexobj(R.data):
        raise NotImplementedError("complex detector weights are not supported")
    out = _gpu.getdet(
        torch.as_tensor(R.indptr.astype(np.int32), device=dev),
        torch.as_tensor(R.indices.astype(np.int32), device=dev),
        torch.as_tensor(R.data.astype(np.float64), device=dev),
        phi_t, 0, ts, is_complex)
    return out.cpu().numpy()


class ForwardResult:
    """Device-resident output of forward_device (field rows in the mesh's device numbering)."""

    def __init__(self, systems, k, X, optodes, detvals, info, is_complex, mesh=None):
        self.mesh = mesh
        self.systems = systems  # [(wv, mode)]
        self.k = k  # columns per system
        self.X = X  # (nn, nsys*k) device
        self.optodes = optodes  # wv -> _Optodes
        self.detvals = detvals  # [(ndet x nsrc) device tensors] per system
        self.info = info
        self.is_complex = is_complex

    def phi(self, i):
        return self.X[:, i * self.k : (i + 1) * self.k]


def locate_optodes(cfg):
    """Located optodes per wavelength (reusable by forward_device while optodes and bulk optics stay fixed)."""
    wvs = list(cfg["prop"].keys()) if isinstance(cfg.get("prop"), dict) else [""]
    mesh = _gpu.mesh_for(cfg)
    return {wv: _Optodes(cfg, mesh, wv) for wv in wvs}


def _element_props(cfg):
    """(ne x 4) element rows [mua, mus', 0, n] equivalent to cfg's label / element properties."""
    props, _, _ = _gpu.props_for(cfg)
    kind, plen, arrays = _gpu.expand_props(cfg, props)
    ne = cfg["elem"].shape[0]
    if arrays[0] == "label":
        tab, seg = arrays[1], np.asarray(arrays[2]).astype(np.int64)
        return np.column_stack([tab[seg, 0], tab[seg, 1], np.zeros(ne), tab[seg, 2]])
    mua, musp, nref = arrays[1], arrays[2], np.asarray(arrays[3], dtype=float)
    if len(mua) != ne:
        raise NotImplementedError("forward_batch takes label or element properties (node properties: one by one)")
    nref = nref if len(nref) == ne else np.full(ne, nref.ravel()[0])
    return np.column_stack([mua, musp, np.zeros(ne), nref])


class BatchResult:
    """Per-mesh outputs of forward_batch: detval (ndet x nsrc), phi (nn x k, [src | det] columns),
    and the node Jacobian when asked for."""

    def __init__(self, detval, phi, jac=None):
        self.detval = detval
        self.phi = phi
        self.jac = jac


class BatchForward:
    """Forward problems of several meshes in one multi-RHS solve (one launch sequence for all).

    cfgs: meshprep'ed single-wavelength cfgs with label or element
    properties, the same omega and the same Reff (the Robin coefficient is
    one per assembly).  The meshes are translated apart (FEM matrices do not
    depend on translation; the node order then keeps each mesh contiguous),
    assembled as one block-diagonal system, and every mesh puts its sources
    in the shared columns [0, TS) and its detectors in [TS, TS + TD) of its
    own rows.  The solve runs to tol / sqrt(number of meshes) relative to a
    column's norm over the batch after every mesh's part of every column is
    scaled to unit norm, which bounds each mesh's own relative residual by
    tol whatever the meshes' RHS norms.  Construction keeps the batch's mesh,
    located optodes and right-hand sides on the device; solve() assembles and
    solves (each call the same work) and returns a BatchResult per cfg.
    """

    def __init__(self, cfgs):
        if not cfgs:
            raise ValueError("BatchForward needs at least one cfg")
        for c in cfgs:
            if isinstance(c.get("prop"), dict):
                raise NotImplementedError("forward_batch takes single-wavelength cfgs")
        om = float(cfgs[0].get("omega", 0))
        reff = float(cfgs[0]["reff"])
        if any(float(c.get("omega", 0)) != om or float(c["reff"]) != reff for c in cfgs):
            raise NotImplementedError("forward_batch needs one omega and one Reff for the whole batch")
        nodes, elems, faces, evol, area, dd, props, shifts, noff = [], [], [], [], [], [], [], [], []
        x0, nn = 0.0, 0
        for c in cfgs:
            nd = np.asarray(c["node"], dtype=float)
            shift = np.array([x0 - nd[:, 0].min(), 0.0, 0.0])
            x0 += (nd[:, 0].max() - nd[:, 0].min()) + 10.0
            nodes.append(nd + shift)
            elems.append(np.asarray(c["elem"])[:, :4].astype(np.int64) + nn)
            faces.append(np.asarray(c["face"])[:, :3].astype(np.int64) + nn)
            evol.append(np.asarray(c["evol"], dtype=float))
            area.append(np.asarray(c["area"], dtype=float))
            if c.get("deldotdel") is None:
                c["deldotdel"], _ = deldotdel(c)
            dd.append(np.asarray(c["deldotdel"], dtype=float))
            props.append(_element_props(c))
            shifts.append(shift)
            noff.append(nn)
            nn += nd.shape[0]
        big = {"node": np.vstack(nodes), "elem": np.vstack(elems), "face": np.vstack(faces),
               "evol": np.concatenate(evol), "area": np.concatenate(area), "prop": np.vstack(props),
               "omega": om, "reff": reff, "isreoriented": True}
        big["deldotdel"] = np.vstack(dd)
        mesh = _gpu.mesh_for(big)
        ops = []
        for m, c in enumerate(cfgs):
            src, det, wsrc, wdet = getoptodes(c)
            op = _Optodes.__new__(_Optodes)
            op.ns, op.nd, op.ws, op.wd = optode_counts(c)
            op.src = np.atleast_2d(src) + shifts[m] if op.ns else None
            op.det = np.atleast_2d(det) + shifts[m] if op.nd else None
            op.sloc, op.sbary = _locate(mesh, op.src)
            op.dloc, op.dbary = _locate(mesh, op.det)
            nm = np.asarray(c["node"]).shape[0]

            def embed(w):  # this mesh's wide columns in the batch's node rows
                if w is None:
                    return None
                e = np.zeros((nn, w.shape[1]))
                e[noff[m] : noff[m] + nm] = w
                return e

            op.wsrc = embed(np.asarray(wsrc)) if op.ws else None
            op.wdet = embed(np.asarray(wdet)) if op.wd else None
            op.k = op.ns + op.ws + op.nd + op.wd
            ops.append(op)
        TS = max(o.ns + o.ws for o in ops)
        TD = max(o.nd + o.wd for o in ops)
        C = TS + TD
        if C > 512:
            raise ValueError("forward_batch: at most 512 source + detector columns per mesh")
        self.is_complex = om > 0
        dt = torch.complex128 if self.is_complex else torch.float64
        B = _gpu.zeros((mesh.nn, C), dt, mesh.dev)
        for m, op in enumerate(ops):
            rows = torch.as_tensor(mesh.iperm[noff[m] : noff[m] + np.asarray(cfgs[m]["node"]).shape[0]],
                                   device=mesh.dev)
            if op.ns:
                _gpu.rhs_block(mesh, B, 0, _loc0(op.sloc), op.sbary, self.is_complex)
            if op.ws:
                B[rows, op.ns : op.ns + op.ws] = torch.as_tensor(op.wsrc[noff[m] : noff[m] + len(rows)],
                                                                 device=mesh.dev, dtype=dt)
            if op.nd:
                _gpu.rhs_block(mesh, B, TS, _loc0(op.dloc), op.dbary, self.is_complex)
            if op.wd:
                B[rows, TS + op.nd : TS + op.nd + op.wd] = torch.as_tensor(op.wdet[noff[m] : noff[m] + len(rows)],
                                                                           device=mesh.dev, dtype=dt)
        # every mesh's part of every column scaled to unit norm (the system is block-diagonal, so
        # the solution block scales alike): a column's residual over the batch, <= tol/sqrt(M)
        # times its norm sqrt(#meshes using it), then bounds each mesh's own relative residual by tol
        self.row_blocks = [torch.as_tensor(mesh.iperm[noff[m] : noff[m] + np.asarray(c["node"]).shape[0]],
                                           device=mesh.dev) for m, c in enumerate(cfgs)]
        self.scales = []
        for rows in self.row_blocks:
            nrm = torch.linalg.norm(B[rows], dim=0)
            sc = torch.where(nrm > 0, nrm, torch.ones_like(nrm))
            B[rows] = B[rows] / sc
            self.scales.append(sc)
        self.detcols = [op.det_columns(big, mesh) if (op.ns + op.ws) and (op.nd + op.wd) else None for op in ops]
        self.cfgs, self.big, self.mesh, self.ops, self.noff = cfgs, big, mesh, ops, noff
        self.TS, self.TD, self.C, self.B = TS, TD, C, B
        self.columns = sum(op.k for op in ops)  # each mesh's own source and detector columns
        self.info = None

    def tune(self, precision="fp64", tol=1e-12, maxiter=None, reps=1):
        """Tune the launch choices of this batch's solve and cache them (tune.tune_system): timed on
        its first column group, whose shape every full-width group of the solve shares."""
        from . import tune as _tune

        mesh = self.mesh
        sysm = _gpu.assemble(mesh, self.big, self.big["deldotdel"], [("", 1)], want32=(precision == "mixed"))
        if maxiter is None:
            maxiter = max(4000, 4 * int(np.sqrt(mesh.nn)) + 1000)
        grp = _gpu.sweep_columns(mesh.nn, self.C, 1, sysm.is_complex)
        return _tune.tune_system(sysm, self.B[:, :grp].contiguous(), grp, precision,
                                 tol / np.sqrt(len(self.cfgs)), maxiter, reps=reps)

    def solve(self, precision="fp64", tol=1e-12, maxiter=None, fields=True, jacobian=False):
        mesh, big, TS, C = self.mesh, self.big, self.TS, self.C
        sysm = _gpu.assemble(mesh, big, big["deldotdel"], [("", 1)], want32=(precision == "mixed"))
        if maxiter is None:
            maxiter = max(4000, 4 * int(np.sqrt(mesh.nn)) + 1000)
        X, info = _gpu.solve_system(sysm, self.B, C, precision, tol / np.sqrt(len(self.cfgs)), maxiter)
        for rows, sc in zip(self.row_blocks, self.scales):
            X[rows] = X[rows] * sc
        _gpu.warn_unconverged(info, "forward_batch")
        self.info = info
        self.X = X
        Xu = mesh.user_rows(X) if fields or jacobian else None
        out = []
        for m, op in enumerate(self.ops):
            ts, td = op.ns + op.ws, op.nd + op.wd
            det_val = None
            if self.detcols[m] is not None:
                p, r, v = self.detcols[m]
                det_val = _gpu.getdet(p, r, v, X, 0, ts, sysm.is_complex).cpu().numpy()
            phi = J = None
            if Xu is not None:
                nm = np.asarray(self.cfgs[m]["node"]).shape[0]
                blk = Xu[self.noff[m] : self.noff[m] + nm]
                phi = torch.cat([blk[:, :ts], blk[:, TS : TS + td]], dim=1).cpu().numpy()
                if jacobian and ts and td:
                    c = self.cfgs[m]
                    J, _ = jac(sdmap(c), phi, c["deldotdel"], c["elem"], c["evol"], return_elem=False)
            out.append(BatchResult(det_val, phi, J))
        return out


def forward_batch(cfgs, precision="fp64", tol=1e-12, maxiter=None, jacobian=False):
    """Forward problems of several meshes in one multi-RHS solve (BatchForward(cfgs).solve(...)).

    cfgs: meshprep'ed single-wavelength cfgs with label or element
    properties, the same omega and the same Reff (the Robin coefficient is
    one per assembly).  The meshes are translated apart (FEM matrices do not
    depend on translation; the node order then keeps each mesh contiguous),
    assembled as one block-diagonal system, and every mesh puts its sources
    in the shared columns [0, TS) and its detectors in [TS, TS + TD) of its
    own rows.  The solve runs to tol / sqrt(number of meshes) relative to a
    column's norm over the batch after every mesh's part of every column is
    scaled to unit norm, which bounds each mesh's own relative residual by
    tol whatever the meshes' RHS norms.  Returns a BatchResult per cfg.
    """
    if not cfgs:
        return []
    bf = BatchForward(cfgs)
    out = bf.solve(precision, tol, maxiter, fields=True, jacobian=jacobian)
    forward_batch.last_info = bf.info
    return out


def check_inputs(cfg, wvs):
    """ValueError for NaN/Inf optical properties or optode positions (before any launch)."""
    for wv in wvs:
        props = _gpu.props_for(cfg, wv, 1)[0]  # refuses complex-valued properties
        if not np.all(np.isfinite(props)):
            raise ValueError("optical properties%s hold NaN or Inf" % ((" of wavelength %s" % wv) if wv else ""))
    for key in ("srcpos", "detpos", "srcdir", "detdir", "widesrc", "widedet"):
        v = cfg.get(key)
        if v is not None and np.size(v) and not np.all(np.isfinite(np.asarray(v, dtype=float))):
            raise ValueError("cfg['%s'] holds NaN or Inf" % key)


def forward_system(cfg, rfcw=(1,), precision="fp64", optodes=None, keep_optodes=False):
    """The device systems of a forward run: (mesh, System, B, k, systems, optodes, wavelengths), B the
    right-hand sides of every (wavelength, mode) system, k columns each (optodes as in forward_device)."""
    if cfg.get("deldotdel") is None:
        cfg["deldotdel"], _ = deldotdel(cfg)
    wvs = list(cfg["prop"].keys()) if isinstance(cfg.get("prop"), dict) else [""]
    check_inputs(cfg, wvs)
    systems = [(wv, md) for wv in wvs for md in rfcw]
    mesh = _gpu.mesh_for(cfg)
    sysm = _gpu.assemble(mesh, cfg, cfg["deldotdel"], systems, want32=(precision == "mixed"))
    # point optodes move inward by the bulk transport length, which changes
    # with the optical properties: a previous set is reused only where the
    # moved positions are unchanged
    old = optodes or {}
    optodes = {}
    for wv in wvs:
        if keep_optodes and wv in old:
            optodes[wv] = old[wv]
            continue
        src, det, _, _ = getoptodes(cfg, wv)
        op = old.get(wv)
        same = op is not None and all(
            (a is None and b is None) or (a is not None and b is not None and np.array_equal(np.atleast_2d(a), b))
            for a, b in ((src, op.src), (det, op.det)))
        optodes[wv] = op if same else _Optodes(cfg, mesh, wv)
    k = optodes[wvs[0]].k
    nsys = len(systems)
    dt = torch.complex128 if sysm.is_complex else torch.float64
    B = _gpu.zeros((mesh.nn, nsys * k), dt, mesh.dev)
    for i, (wv, md) in enumerate(systems):
        optodes[wv].fill_block(mesh, B, i * k, sysm.is_complex)
    return mesh, sysm, B, k, systems, optodes, wvs


def forward_device(cfg, rfcw=(1,), precision="fp64", tol=1e-12, maxiter=None, X0=None, optodes=None,
                   keep_optodes=False, where="runforward"):
    """Assemble, build RHS and solve all (wavelength, mode) systems on the GPU.

    optodes: located optodes of an earlier call, reused where their moved
    positions are unchanged (or as given, with keep_optodes=True: the
    optodes then stay where they were when the optical properties change).
    A solve that stops short of ``tol`` (iteration limit, breakdown) raises
    a SolverWarning naming ``where``; its flag is in the result's info.
    """
    mesh, sysm, B, k, systems, optodes, wvs = forward_system(cfg, rfcw, precision, optodes, keep_optodes)
    nsys = len(systems)
    if maxiter is None:
        maxiter = max(4000, 4 * int(np.sqrt(mesh.nn)) + 1000)
    if nsys * k <= 512:
        X, info = _gpu.solve_system(sysm, B, k, precision, tol, maxiter, X0)
    else:
        X = torch.empty_like(B)
        info = None
        for i in range(nsys):
            sub = _gpu.System(mesh, sysm.vals64[i : i + 1], None if sysm.vals32 is None else sysm.vals32[i : i + 1],
                              1, sysm.is_complex)
            x0 = None if X0 is None else X0[:, i * k : (i + 1) * k]
            Xi, inf = _gpu.solve_system(sub, B[:, i * k : (i + 1) * k].contiguous(), k, precision, tol, maxiter, x0)
            X[:, i * k : (i + 1) * k] = Xi
            info = inf if info is None or (inf.flag, inf.relres) > (info.flag, info.relres) else info
    _gpu.warn_unconverged(info, where)
    detvals = []
    for i, (wv, md) in enumerate(systems):
        op = optodes[wv]
        ts = op.ns + op.ws
        if ts == 0 or op.nd + op.wd == 0:
            detvals.append(None)
            continue
        if not hasattr(op, "_detcols"):
            op._detcols = op.det_columns(cfg, mesh)
        p, r, v = op._detcols
        detvals.append(_gpu.getdet(p, r, v, X[:, i * k : (i + 1) * k], 0, ts, sysm.is_complex))
    return ForwardResult(systems, k, X, optodes, detvals, info, sysm.is_complex, mesh)


def runforward(cfg, **kwargs):
    """Forward simulation at all sources and wavelengths (forward.py:40-87).

    Returns (detval, phi) with the reference's nesting: dicts by wavelength
    when cfg['prop'] is a dict, by mode when rfcw has several entries.
    Extra keywords: precision ('fp64' | 'mixed'), tol, maxiter.
    """
    global last_info
    rfcw = kwargs.get("rfcw", [1])
    if isinstance(rfcw, int):
        rfcw = [rfcw]
    res = forward_device(cfg, rfcw, kwargs.get("precision", "fp64"), kwargs.get("tol", 1e-12),
                         kwargs.get("maxiter", None))
    last_info = res.info
    wvs = list(cfg["prop"].keys()) if isinstance(cfg.get("prop"), dict) else [""]
    det = {md: {"detphi": {}} for md in rfcw}
    phi = {md: {"phi": {}} for md in rfcw}
    Xh = res.mesh.user_rows(res.X).cpu().numpy()
    for i, (wv, md) in enumerate(res.systems):
        phi[md]["phi"][wv] = np.ascontiguousarray(Xh[:, i * res.k : (i + 1) * res.k])
        dv = res.detvals[i]
        det[md]["detphi"][wv] = dv.cpu().numpy() if dv is not None else np.array([])
    if len(wvs) == 1:
        for md in rfcw:
            phi[md]["phi"] = phi[md]["phi"][wvs[0]]
            det[md]["detphi"] = det[md]["detphi"][wvs[0]]
    if len(rfcw) == 1:
        phi = phi[rfcw[0]]["phi"]
        det = det[rfcw[0]]["detphi"]
    return det, phi


def _active_pairs(sd):
    sd = np.asarray(sd)
    if sd.shape[1] >= 3:
        sd = sd[sd[:, 2] == 1]
    return sd[:, 0].astype(np.int64), sd[:, 1].astype(np.int64)


def jac(sd, phi, deldotdel_mat, elem, evol, iselem=False, return_elem=True):
    """mua Jacobian (forward.py:509-569): (Jnode nsd x nn, Jelem nsd x ne).

    Rows are the active sd pairs in order.  ``return_elem=False`` skips the
    element Jacobian (the reference always builds it) and returns None there.
    """
    s, d = _active_pairs(sd)
    dev = _gpu.device()
    phi_t = phi if torch.is_tensor(phi) else torch.as_tensor(np.ascontiguousarray(phi), device=dev)
    phi_t = phi_t.to(dev)
    if not phi_t.is_complex():
        phi_t = phi_t.to(torch.float64)
    mesh = _gpu.mesh_from_elem(elem, phi_t.shape[0], evol)
    out = torch.complex128 if phi_t.is_complex() else torch.float64
    JT, JE = _gpu.jac_node(mesh, mesh.mesh_rows(phi_t).contiguous(), s, d, out, want_elem=return_elem)
    Jn = mesh.user_rows(JT).cpu().numpy().T
    Je = JE.cpu().numpy().T if JE is not None else None
    return Jn, Je


def jacchrome(Jmua, chromophores, extin=None):
    """Chromophore Jacobians: stack over wavelengths of E[w, c] * Jmua[w] (forward.py:572-589).

    ``extin`` may replace the table with an (Nwv x Nchrome) matrix.  This
    is a scaling and stacking of arrays already computed; the reconstruction
    path never materialises it (it uses the Kronecker form of the Gram).
    """
    if not isinstance(Jmua, dict):
        raise ValueError("Jmua must be a dict with wavelength keys")
    wvs = list(Jmua.keys())
    E = extinction(wvs, chromophores)[0] if extin is None else np.asarray(extin, dtype=float)
    out = {}
    for i, ch in enumerate(chromophores):
        out[ch] = np.vstack([Jmua[wv] * E[j, i] for j, wv in enumerate(wvs)])
    return out
