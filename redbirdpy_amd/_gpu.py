"""Device-resident engine: mesh structures and the hot-path calls on torch tensors.

Everything here runs through librbhip (hand-written HIP kernels); torch only
allocates device memory and provides the stream.  The public redbirdpy-
compatible functions in ``forward``/``solver``/``recon`` convert numpy/scipy
inputs at their boundary and call into this module.
"""

import ctypes
import os
import numpy as np
import torch

from . import _abi
from ._abi import ptr, call

C0 = 299792458000.0  # mm/s (redbirdpy forward.py:36)


def device():
    _abi.require_gpu()
    return torch.device("cuda", torch.cuda.current_device())


_NP2TORCH = {np.dtype(np.float64): torch.float64, np.dtype(np.float32): torch.float32,
             np.dtype(np.int32): torch.int32, np.dtype(np.int64): torch.int64,
             np.dtype(np.complex128): torch.complex128, np.dtype(np.complex64): torch.complex64}


def h2d(a, dev, dtype=None):
    """Host array -> device tensor through a pinned staging copy enqueued on the
    current stream without waiting for it (a pageable copy first drains the
    stream, which leaves the GPU idle while the host prepares the next launch).
    The staging block is torch's cached pinned memory, held until the copy ends."""
    a = np.asarray(a) if dtype is None else np.asarray(a, dtype=dtype)
    tdt = _NP2TORCH.get(a.dtype)
    if tdt is None or a.size == 0:
        return torch.as_tensor(np.ascontiguousarray(a), device=dev)
    stage = torch.empty(a.shape, dtype=tdt, pin_memory=True)
    stage.numpy()[...] = a
    return stage.to(dev, non_blocking=True)


def _i32(a, dev):
    return h2d(a, dev, np.int32)


def _f64(a, dev):
    return h2d(a, dev, np.float64)


def _spread_bits(v):
    """Insert two zero bits between the low 21 bits of v (3-D Morton interleave)."""
    v = v & np.uint64(0x1FFFFF)
    v = (v | (v << np.uint64(32))) & np.uint64(0x1F00000000FFFF)
    v = (v | (v << np.uint64(16))) & np.uint64(0x1F0000FF0000FF)
    v = (v | (v << np.uint64(8))) & np.uint64(0x100F00F00F00F00F)
    v = (v | (v << np.uint64(4))) & np.uint64(0x10C30C30C30C30C3)
    v = (v | (v << np.uint64(2))) & np.uint64(0x1249249249249249)
    return v


def morton_order(node):
    """Node order along a 3-D Morton curve of the coordinates (new -> old).

    Consecutive nodes of the new order form compact bricks, so a block of
    consecutive CSR rows has a small union of neighbour rows (the SpMM
    stages that union in LDS).  The order is a stable sort: deterministic.
    """
    node = np.asarray(node, dtype=np.float64)
    lo = node.min(axis=0)
    span = float(np.max(node.max(axis=0) - lo))
    if span <= 0:
        return np.arange(node.shape[0])
    q = np.floor((node - lo) / span * (2 ** 21 - 1)).astype(np.uint64)
    code = _spread_bits(q[:, 0]) | (_spread_bits(q[:, 1]) << np.uint64(1)) | (_spread_bits(q[:, 2]) << np.uint64(2))
    return np.argsort(code, kind="stable")


class GpuMesh:
    """Device copies of a tet mesh plus its CSR pattern and assembly slot map.

    elem/face are taken 1-based (redbirdpy convention, __init__.py:13-15).
    On the device the nodes are renumbered along a Morton curve (``perm``:
    new -> old, ``iperm``: old -> new) for SpMM locality; every device array
    indexed by node (fields, RHS, Jacobian rows, node properties) uses the
    new numbering, and ``user_rows``/``mesh_rows`` convert at the API.
    """

    def __init__(self, node, elem, face=None, evol=None, area=None, dev=None, reorder=True):
        dev = dev or device()
        self.dev = dev
        node = np.asarray(node, dtype=np.float64)
        e0u = np.asarray(elem)[:, :4].astype(np.int64) - 1
        self.nn = int(node.shape[0])
        self.ne = int(e0u.shape[0])
        self.perm = morton_order(node) if reorder else np.arange(self.nn)
        self.iperm = np.empty_like(self.perm)
        self.iperm[self.perm] = np.arange(self.nn)
        self.reordered = bool(reorder)
        self.perm_t = torch.as_tensor(self.perm, device=dev)
        self.iperm_t = torch.as_tensor(self.iperm, device=dev)
        e0 = self.iperm[e0u]
        self.node = _f64(node[self.perm], dev)
        self.el = _i32(e0, dev)
        self.evol = _f64(evol, dev) if evol is not None else None
        if face is not None and np.size(face):
            f0 = self.iperm[np.asarray(face)[:, :3].astype(np.int64) - 1]
            self.nf = int(f0.shape[0])
            self.fc = _i32(f0, dev)
            self.area = _f64(area, dev) if area is not None else None
        else:
            self.nf = 0
            self.fc = None
            self.area = None
        s = _abi.stream_ptr()
        nn, ne, nf = self.nn, self.ne, self.nf
        self.tinc_ptr = torch.empty(nn + 1, dtype=torch.int32, device=dev)
        self.tinc = torch.empty(max(4 * ne, 1), dtype=torch.int32, device=dev)
        call("rb_incidence", ptr(self.el), ne, 4, nn, ptr(self.tinc_ptr), ptr(self.tinc), s)
        if nf:
            self.finc_ptr = torch.empty(nn + 1, dtype=torch.int32, device=dev)
            self.finc = torch.empty(3 * nf, dtype=torch.int32, device=dev)
            call("rb_incidence", ptr(self.fc), nf, 3, nn, ptr(self.finc_ptr), ptr(self.finc), s)
        else:
            self.finc_ptr = None
            self.finc = None
        self.row_ptr = torch.empty(nn + 1, dtype=torch.int32, device=dev)
        nnz = ctypes.c_int64(0)
        call("rb_csr_pattern", ptr(self.el), ne, nn, ptr(self.tinc_ptr), ptr(self.tinc), ptr(self.row_ptr),
             None, 0, ctypes.byref(nnz), None, s)
        self.nnz = int(nnz.value)
        self.col_idx = torch.empty(max(self.nnz, 1), dtype=torch.int32, device=dev)
        self.diag = torch.empty(nn, dtype=torch.int32, device=dev)
        call("rb_csr_pattern", ptr(self.el), ne, nn, ptr(self.tinc_ptr), ptr(self.tinc), ptr(self.row_ptr),
             ptr(self.col_idx), self.nnz, ctypes.byref(nnz), ptr(self.diag), s)
        self.slot_ptr = torch.empty(self.nnz + 1, dtype=torch.int32, device=dev)
        self.codes = torch.empty(max(16 * ne + 9 * nf, 1), dtype=torch.int32, device=dev)
        call("rb_slot_map", ptr(self.el), ne, ptr(self.fc), nf, nn, ptr(self.tinc_ptr), ptr(self.tinc),
             ptr(self.finc_ptr), ptr(self.finc), ptr(self.row_ptr), ptr(self.col_idx), self.nnz,
             ptr(self.slot_ptr), ptr(self.codes), s)
        self.maxnp = int((self.row_ptr[1:] - self.row_ptr[:-1]).max().item()) if nn else 0
        self._dd_src = None
        self.dd = None

    def set_geometry(self, deldotdel, evol, area):
        """Upload element geometry (cached by object identity)."""
        if self._dd_src is not deldotdel:
            self.dd = _f64(np.asarray(deldotdel), self.dev)
            self._dd_src = deldotdel
        if evol is not None and (self.evol is None or getattr(self, "_evol_src", None) is not evol):
            self.evol = _f64(evol, self.dev)
            self._evol_src = evol
        if area is not None and self.nf and (self.area is None or getattr(self, "_area_src", None) is not area):
            self.area = _f64(area, self.dev)
            self._area_src = area

    def asm_workspace(self, nsys):
        """Property tables of rb_assemble_ws for nsys systems (kept for reuse; stream-ordered, so one
        assembly at a time per stream)."""
        nbytes = ctypes.c_size_t(0)
        call("rb_assemble_workspace", self.nn, self.ne, nsys, ctypes.byref(nbytes))
        w = getattr(self, "_asm_work", None)
        if w is None or w.numel() < nbytes.value:
            w = torch.empty(int(nbytes.value), dtype=torch.uint8, device=self.dev)
            self._asm_work = w
        return w

    def asm_bins(self):
        """Workgroup bins of the slot map for rb_assemble_ws (built once)."""
        b = getattr(self, "_asm_bins", None)
        if b is None:
            cnt = ctypes.c_int64(0)
            call("rb_assemble_bin_count", self.ne, self.nf, ctypes.byref(cnt))
            b = torch.empty(int(cnt.value), dtype=torch.int32, device=self.dev)
            call("rb_assemble_bins", ptr(self.slot_ptr), self.nnz, self.ne, self.nf, ptr(b), _abi.stream_ptr())
            self._asm_bins = b
        return b

    def spmm_plan(self, sets=1):
        """The LDS-staged SpMM plan of this pattern (built once) with room for `sets` record value
        sets; None when it does not apply."""
        if getattr(self, "_plan", False) is False:
            self._plan = spmm_plan(self.row_ptr.cpu().numpy(), self.col_idx[: self.nnz].cpu().numpy(), self.dev)
        if os.environ.get("RB_LDS_SPMM", "1") == "0" or self._plan is None:
            return None
        return self._plan.ensure_sets(sets)

    def pattern_host(self):
        """(indptr, indices, slot order) of the CSR pattern in the user's node numbering.

        data_user = data_device[order] maps device slot values to that pattern.
        """
        if getattr(self, "_user_pattern", None) is None:
            rp = self.row_ptr.cpu().numpy().astype(np.int64)
            ci = self.col_idx[: self.nnz].cpu().numpy().astype(np.int64)
            rows = self.perm[np.repeat(np.arange(self.nn), np.diff(rp))]
            cols = self.perm[ci]
            order = np.lexsort((cols, rows))
            indptr = np.zeros(self.nn + 1, dtype=np.int64)
            np.cumsum(np.bincount(rows, minlength=self.nn), out=indptr[1:])
            self._user_pattern = (indptr, cols[order], order)
            self._user_order_t = torch.as_tensor(order, device=self.dev)
        return self._user_pattern

    def user_values(self, vals):
        """Device slot values -> the user-numbered CSR data order."""
        self.pattern_host()
        return vals[..., self._user_order_t]

    def user_rows(self, X):
        """Rows of a node-indexed device array in the user's numbering."""
        return X if not self.reordered else X[self.iperm_t]

    def mesh_rows(self, X):
        """User-numbered node rows -> the device numbering."""
        return X if not self.reordered else X[self.perm_t]

    def mesh_nodes(self, a):
        """Host array indexed by user node -> device numbering."""
        return a if not self.reordered else np.asarray(a)[self.perm]


class _Plan:
    """An rb_spmm_plan (ctypes struct of device pointers) with the tensors that own its memory."""

    def __init__(self, tensors, nblk, nrec, limits):
        self.tensors = tensors
        self.limits = limits
        self.nrec = nrec
        self.struct = _abi.SpmmPlan(nblk, nrec, limits[0], limits[1],
                                    *[ctypes.c_void_p(t.data_ptr()) for t in tensors], 1)

    def ensure_sets(self, sets):
        """Room for `sets` record value sets (one per real float32 system pair)."""
        if self.struct.recv_sets < sets:
            recv = torch.empty(sets * (self.nrec + _abi.RB_LDS_RECPAD), dtype=torch.float64,
                               device=self.tensors[-1].device)
            self.tensors[-1] = recv
            self.struct.recv = ctypes.c_void_p(recv.data_ptr())
            self.struct.recv_sets = sets
        return self


_PLAN_CACHE = []  # [(key, plan)]: plans by pattern content, most recent first


def _pattern_key(rp, ci, dev, limits):
    try:
        import xxhash

        h = xxhash.xxh3_64()
    except ImportError:  # pragma: no cover - xxhash ships with this environment
        import hashlib

        h = hashlib.blake2b(digest_size=16)
    h.update(rp.tobytes())
    h.update(ci.tobytes())
    return (str(dev), tuple(limits), rp.size, ci.size, h.hexdigest())


def spmm_plan(row_ptr, col_idx, dev, limits=None):
    """LDS-staged SpMM plan of a CSR pattern (host arrays, device row order): blocks of consecutive
    rows with their neighbour-row unions and padded entry records (rb_spmm_plan_count/fill on the
    host), uploaded to dev.  None when some row alone exceeds the block limits (RB_ERR_SIZE).  Plans
    are kept by pattern content (a small LRU): a mesh rebuilt from the same arrays reuses its plan."""
    rp = np.ascontiguousarray(np.asarray(row_ptr).astype(np.int32))
    ci = np.ascontiguousarray(np.asarray(col_idx).astype(np.int32))
    key = _pattern_key(rp, ci, dev, limits or _abi.RB_LDS_LIMITS)
    for i, (k, p) in enumerate(_PLAN_CACHE):
        if k == key:
            _PLAN_CACHE.insert(0, _PLAN_CACHE.pop(i))
            return p
    p = _build_spmm_plan(rp, ci, dev, limits)
    _PLAN_CACHE.insert(0, (key, p))
    del _PLAN_CACHE[4:]
    return p


def _build_spmm_plan(rp, ci, dev, limits=None):
    n = rp.size - 1
    nblk, nun, nrec = ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_int64(0)
    L = _abi.lib()
    hp = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    limits = limits or _abi.RB_LDS_LIMITS
    rc = L.rb_spmm_plan_count(n, hp(rp), hp(ci), limits[0], limits[1], ctypes.byref(nblk), ctypes.byref(nun),
                              ctypes.byref(nrec))
    if rc == -3:  # RB_ERR_SIZE: the gather SpMM stays
        return None
    _abi.check(rc, "rb_spmm_plan_count")
    nb, nu, nr = nblk.value, nun.value, nrec.value
    brow = np.zeros(nb + 1, np.int32)
    bu = np.zeros(nb + 1, np.int32)
    ul = np.zeros(nu + _abi.RB_LDS_UPAD, np.int32)
    ri = np.zeros(2 * (n + 1) + _abi.RB_LDS_RPAD, np.int32)
    rm = np.zeros(nr + _abi.RB_LDS_RECPAD, np.int32)
    lo = np.zeros(nr + _abi.RB_LDS_RECPAD, np.int32)
    _abi.check(L.rb_spmm_plan_fill(n, hp(rp), hp(ci), limits[0], limits[1], hp(brow), hp(bu), hp(ul), hp(ri),
                                   hp(rm), hp(lo)), "rb_spmm_plan_fill")
    t = [torch.as_tensor(a, device=dev) for a in (brow, bu, ul, ri, rm, lo)]
    t.append(torch.empty(nr + _abi.RB_LDS_RECPAD, dtype=torch.float64, device=dev))  # 8-byte record values
    return _Plan(t, nb, nr, limits)


_MESH_CACHE = []  # [(key objects, GpuMesh)], small LRU


def mesh_for(cfg):
    """GpuMesh for cfg (node/elem/face), cached on the identity of those arrays."""
    key = (cfg["node"], cfg["elem"], cfg.get("face"))
    for i, (k, m) in enumerate(_MESH_CACHE):
        if all(a is b for a, b in zip(k, key)):
            _MESH_CACHE.insert(0, _MESH_CACHE.pop(i))
            return m
    m = GpuMesh(cfg["node"], cfg["elem"], cfg.get("face"), cfg.get("evol"), cfg.get("area"))
    _MESH_CACHE.insert(0, (key, m))
    del _MESH_CACHE[4:]
    return m


def mesh_from_elem(elem, nn, evol):
    """GpuMesh without faces, for jac() which only receives elem/evol."""
    key = (None, elem, None)
    for i, (k, m) in enumerate(_MESH_CACHE):
        if k[1] is elem and m.nn == nn:
            _MESH_CACHE.insert(0, _MESH_CACHE.pop(i))
            if m.evol is None or getattr(m, "_evol_src", None) is not evol:
                m.evol = _f64(evol, m.dev)
                m._evol_src = evol
            return m
    node = np.zeros((nn, 3))
    m = GpuMesh(node, elem, None, evol, None, reorder=False)
    m._evol_src = evol
    _MESH_CACHE.insert(0, (key, m))
    del _MESH_CACHE[4:]
    return m


# ---------------------------------------------------------------------------
# properties (redbirdpy forward.py:105-142)
# ---------------------------------------------------------------------------
def props_for(cfg, wavelength="", mode=1):
    """Return (props, reff, omega) as femlhs sees them for one wavelength/mode."""
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
    props = np.asarray(props)
    if np.iscomplexobj(props):  # e.g. concentrations after a reform='complex' update (recon.py:305-313)
        raise NotImplementedError("complex-valued optical properties%s: the GPU assembly takes real mua, mus, g "
                                  "and n" % ((" of wavelength %s" % wavelength) if wavelength else ""))
    return np.asarray(props, dtype=np.float64), float(reff), float(om)


def expand_props(cfg, props):
    """Property layout as femlhs chooses it (forward.py:128-142).

    Returns (kind, plen, arrays) where arrays is ("rows", mua, musp, nref)
    for per-node/per-element tables, or ("label", table, seg) for a label
    table indexed by cfg['seg'] (n taken from the first element's label).
    """
    nn = cfg["node"].shape[0]
    ne = cfg["elem"].shape[0]
    seg = cfg.get("seg", None)
    if props.ndim == 1:
        props = props.reshape(1, -1)
    nc = props.shape[1]
    if props.shape[0] == nn or props.shape[0] == ne:
        mua = props[:, 0]
        musp = props[:, 1] * (1 - props[:, 2]) if nc >= 3 else props[:, 1]
        nref = props[:, 3] if nc >= 4 else np.array([1.37])
        kind = _abi.RB_PROP_ELEM if len(mua) == ne else _abi.RB_PROP_NODE
        return kind, len(mua), ("rows", np.asarray(mua, float), np.asarray(musp, float), np.asarray(nref, float))
    if seg is not None:
        seg = np.asarray(seg)
        if seg.shape[0] != ne:
            raise ValueError("label properties need one seg label per element")
        tab = np.zeros((props.shape[0], 3))
        tab[:, 0] = props[:, 0]
        tab[:, 1] = props[:, 1] * (1 - props[:, 2]) if nc >= 3 else props[:, 1]
        s0 = int(np.clip(seg[0], 0, props.shape[0] - 1))
        tab[:, 2] = props[s0, 3] if nc >= 4 else 1.37
        return _abi.RB_PROP_ELEM, ne, ("label", tab, seg)
    raise ValueError("Property format not recognized")


def _seg_device(mesh, seg, nlabel):
    """Clipped int64 labels on the device, cached on the seg array identity and sum."""
    key = (id(seg), seg.shape[0], int(np.sum(seg)), nlabel)
    c = getattr(mesh, "_seg_cache", None)
    if c is not None and c[0] == key and c[2] is seg:
        return c[1]
    sd = torch.as_tensor(np.clip(seg.astype(np.int64), 0, nlabel - 1), device=mesh.dev)
    mesh._seg_cache = (key, sd, seg)
    return sd


class System:
    """One assembled FEM system set: nsys matrices sharing the mesh pattern."""

    def __init__(self, mesh, vals64, vals32, nsys, is_complex):
        self.mesh = mesh
        self.vals64 = vals64
        self.vals32 = vals32
        self.nsys = nsys
        self.is_complex = is_complex


class AsmInputs:
    """Device inputs of one assembly (rb_assemble / rb_assemble_ws arguments after the mesh's)."""

    def __init__(self, kind, plen, nlen, mua, musp, nref, om, bc, is_complex, nsys):
        self.kind, self.plen, self.nlen = kind, plen, nlen
        self.mua, self.musp, self.nref, self.om, self.bc = mua, musp, nref, om, bc
        self.is_complex, self.nsys = is_complex, nsys


def assemble_inputs(mesh, cfg, deldotdel, systems):
    """Property arrays of (wavelength, mode) systems on the device, in the layout rb_assemble takes."""
    dev = mesh.dev
    mesh.set_geometry(deldotdel, cfg["evol"], cfg.get("area"))
    cols = {"mua": [], "musp": [], "nref": []}
    oms, bcs = [], []
    kind = plen = None
    nlen = 1
    for wv, md in systems:
        props, reff, om = props_for(cfg, wv, md)
        k, pl, arr = expand_props(cfg, props)
        if kind is None:
            kind, plen = k, pl
        elif kind != k or plen != pl:
            raise ValueError("all systems of one assembly must share the property layout")
        if arr[0] == "label":
            tab = torch.as_tensor(arr[1], device=dev)
            seg = _seg_device(mesh, arr[2], arr[1].shape[0])
            rows = tab[seg]  # (ne, 3) gather on the device
            cols["mua"].append(rows[:, 0])
            cols["musp"].append(rows[:, 1])
            cols["nref"].append(tab[:1, 2])  # n of the first element's label, as femlhs uses
        else:
            mua, musp, nref = _f64(arr[1], dev), _f64(arr[2], dev), _f64(arr[3], dev)
            if k == _abi.RB_PROP_NODE:  # node properties follow the device numbering (gathered on the device)
                mua, musp = mesh.mesh_rows(mua), mesh.mesh_rows(musp)
                if nref.numel() == mesh.nn:
                    nref = mesh.mesh_rows(nref)
            cols["mua"].append(mua)
            cols["musp"].append(musp)
            cols["nref"].append(nref)
            nlen = max(nlen, nref.numel())
        oms.append(om)
        bcs.append((1 - reff) / (12.0 * (1 + reff)))
    nref_rows = [x if x.numel() == nlen else x.expand(nlen) for x in cols["nref"]]
    return AsmInputs(kind, plen, nlen, torch.stack(cols["mua"]).contiguous(), torch.stack(cols["musp"]).contiguous(),
                     torch.stack(nref_rows).contiguous(), _f64(np.array(oms), dev), _f64(np.array(bcs), dev),
                     any(o > 0 for o in oms), len(systems))


def assemble_call(mesh, a, want32=False, entry="rb_assemble_ws"):
    """Run one assembly of AsmInputs a through entry (rb_assemble_ws, or rb_assemble): (v64, v32)."""
    shape = (a.nsys, mesh.nnz)
    dev = mesh.dev
    v64 = torch.empty(shape, dtype=torch.complex128 if a.is_complex else torch.float64, device=dev)
    v32 = torch.empty(shape, dtype=torch.complex64 if a.is_complex else torch.float32, device=dev) if want32 else None
    args = (ptr(mesh.slot_ptr), ptr(mesh.codes), mesh.nnz, ptr(mesh.el), mesh.ne, ptr(mesh.fc), mesh.nf, mesh.nn,
            ptr(mesh.dd), ptr(mesh.evol), ptr(mesh.area), a.nsys, a.kind, ptr(a.mua), ptr(a.musp), ptr(a.nref),
            a.plen, a.nlen, ptr(a.om), ptr(a.bc), int(a.is_complex), ptr(v64), ptr(v32))
    if entry == "rb_assemble_ws":
        work = mesh.asm_workspace(a.nsys)
        call("rb_assemble_ws", *args, ptr(mesh.asm_bins()), ptr(work), work.numel(), _abi.stream_ptr())
    else:
        call(entry, *args, _abi.stream_ptr())
    return v64, v32


def assemble(mesh, cfg, deldotdel, systems, want32=False):
    """Assemble femlhs matrices for a list of (wavelength, mode) systems in one launch sequence.

    Returns a System with values (nsys x nnz) float64/complex128 (+ 32-bit copy).
    """
    a = assemble_inputs(mesh, cfg, deldotdel, systems)
    v64, v32 = assemble_call(mesh, a, want32)
    return System(mesh, v64, v32, a.nsys, a.is_complex)


def volume_weights(mesh):
    """W[n, i] = sum of V_e over tets holding n and i (kind RB_PROP_VOLUME)."""
    w = torch.empty(mesh.nnz, dtype=torch.float64, device=mesh.dev)
    call("rb_assemble", ptr(mesh.slot_ptr), ptr(mesh.codes), mesh.nnz, ptr(mesh.el), mesh.ne, None, 0, mesh.nn,
         ptr(mesh.dd), ptr(mesh.evol), None, 1, _abi.RB_PROP_VOLUME, None, None, None, 0, 1, None, None, 0,
         ptr(w), None, _abi.stream_ptr())
    return w


# ---------------------------------------------------------------------------
# solve
# ---------------------------------------------------------------------------
_PREC = {"fp64": _abi.RB_PREC_FP64, "mixed": _abi.RB_PREC_MIXED}


class SolveInfo:
    def __init__(self, iters, relres, flag, groups=None):
        self.iters = iters
        self.relres = relres
        self.flag = flag
        self.groups = groups  # [(columns, iterations)] of the column groups a wide solve ran as

    def __repr__(self):
        return "SolveInfo(iters=%d, relres=%.3e, flag=%d)" % (self.iters, self.relres, self.flag)


def krylov_solve(row_ptr, col_idx, diag, n, nnz, vals64, vals32, B, k, nsys, is_complex, precision="fp64",
                 tol=1e-12, maxiter=2000, X0=None, check_every=16, plan=None):
    """Solve A_s X = B for nsys systems x k columns; B is (n, nsys*k) on device."""
    dev = B.device
    prec = _PREC[precision]
    C = nsys * k
    if C > 512:
        raise ValueError("at most 512 columns per solve call")
    cdt = torch.complex128 if is_complex else torch.float64
    B = B.to(cdt).contiguous()
    if X0 is not None:
        X = X0.to(cdt).clone().contiguous()
        use_x0 = 1
    else:
        X = torch.empty_like(B)
        use_x0 = 0
    nbytes = ctypes.c_size_t(0)
    call("rb_krylov_workspace", n, nsys, k, int(is_complex), prec, ctypes.byref(nbytes))
    work = torch.empty(int(nbytes.value), dtype=torch.uint8, device=dev)
    it = ctypes.c_int(0)
    rel = ctypes.c_double(0)
    fl = ctypes.c_int(0)
    if prec == _abi.RB_PREC_MIXED and vals32 is None:
        vals32 = vals64.to(torch.complex64 if is_complex else torch.float32)
    if plan is not None:
        plan = plan.struct
        call("rb_krylov_solve_planned", n, nnz, ptr(row_ptr), ptr(col_idx), ptr(diag), nsys, k, int(is_complex),
             ptr(vals64), ptr(vals32), ptr(B), ptr(X), prec, float(tol), int(maxiter), use_x0, int(check_every),
             ptr(work), nbytes.value, ctypes.byref(plan), ctypes.byref(it), ctypes.byref(rel), ctypes.byref(fl),
             _abi.stream_ptr())
    else:
        call("rb_krylov_solve", n, nnz, ptr(row_ptr), ptr(col_idx), ptr(diag), nsys, k, int(is_complex),
             ptr(vals64), ptr(vals32), ptr(B), ptr(X), prec, float(tol), int(maxiter), use_x0, int(check_every),
             ptr(work), nbytes.value, ctypes.byref(it), ctypes.byref(rel), ctypes.byref(fl), _abi.stream_ptr())
    return X, SolveInfo(it.value, rel.value, fl.value)


def bicgstab_solve(row_ptr, col_idx, diag, n, nnz, vals64, B, is_complex, tol=1e-12, maxiter=2000, X0=None,
                   check_every=16):
    """BiCGSTAB (rb_bicgstab_solve) for one system: B (n, ncol) on device, float64 / complex128."""
    dev = B.device
    cdt = torch.complex128 if is_complex else torch.float64
    B = B.to(cdt).contiguous()
    ncol = B.shape[1]
    if ncol > 512:
        raise ValueError("at most 512 columns per solve call")
    X = X0.to(cdt).clone().contiguous() if X0 is not None else torch.empty_like(B)
    nbytes = ctypes.c_size_t(0)
    call("rb_bicgstab_workspace", n, ncol, int(is_complex), ctypes.byref(nbytes))
    work = torch.empty(int(nbytes.value), dtype=torch.uint8, device=dev)
    it = ctypes.c_int(0)
    rel = ctypes.c_double(0)
    fl = ctypes.c_int(0)
    vals = vals64.reshape(-1)[:nnz].to(cdt).contiguous()
    call("rb_bicgstab_solve", n, nnz, ptr(row_ptr), ptr(col_idx), ptr(diag), ncol, int(is_complex), ptr(vals),
         ptr(B), ptr(X), float(tol), int(maxiter), int(X0 is not None), int(check_every), ptr(work), nbytes.value,
         ctypes.byref(it), ctypes.byref(rel), ctypes.byref(fl), _abi.stream_ptr())
    return X, SolveInfo(it.value, rel.value, fl.value)


class SolverWarning(RuntimeWarning):
    """A forward solve stopped before every column reached its tolerance."""


def warn_unconverged(info, where):
    """Warn (SolverWarning) when a solve's flag is nonzero: 1 = iteration limit, 4 = breakdown."""
    if info is not None and info.flag:
        import warnings

        why = {1: "reached the iteration limit", 4: "broke down"}.get(info.flag, "failed (flag %d)" % info.flag)
        warnings.warn("%s: the Krylov solve %s after %d iterations; largest relative residual %.3e, so some "
                      "fields and detector values are not at the requested tolerance"
                      % (where, why, info.iters, info.relres), SolverWarning, stacklevel=3)


SWEEP_MAX_BYTES = 1 << 31


def lds_plan_sets(nsys, k, is_complex, precision):
    """Record value sets an LDS-staged SpMM plan needs for this solve shape (0: it does not apply):
    for 64-column solves of one system with 8-byte or 4-byte inner values (complex64 or float32 mixed,
    float64 real) two, A and A D^-1 (the polynomial preconditioner's product); for mixed solves of an
    even number of real 64-column systems one per pair, then one per pair for the pairs' A D^-1."""
    if k != 64:
        return 0
    if nsys == 1 and (precision == "mixed" or not is_complex):
        return 2
    if precision == "mixed" and not is_complex and nsys >= 2 and nsys % 2 == 0:
        return nsys
    return 0


def sweep_columns(n, k, nsys, is_complex):
    """Columns per solve call so that a one-system solve of k columns keeps its float64 vectors
    (n rows) under the column sweep's 2 GiB buffer descriptor: k itself when it fits (or when
    no multiple of 64 columns fits, or several systems share the call), else the largest
    multiple of 64 that fits.  RB_SPMM_FLAT_MAX_BYTES lowers the limit as in the library."""
    lim = SWEEP_MAX_BYTES
    env = os.environ.get("RB_SPMM_FLAT_MAX_BYTES")
    if env:
        lim = min(lim, int(env))
    row = n * (16 if is_complex else 8)
    if nsys != 1 or k <= 64 or row * k < lim:
        return k
    g = (lim - 1) // row // 64 * 64
    return g if g >= 64 else k


def solve_system(sysm, B, k, precision="fp64", tol=1e-12, maxiter=2000, X0=None, tuned=None):
    """CG/COCG over every system; after a breakdown (flag 4) every system is finished by
    BiCGSTAB from its CG iterate (its columns that had converged stop at once).

    tuned: launch choices {"lds", "grid", "check_every"} (tune.py); None takes the autotuner's
    cached entry for this shape, if any (RB_AUTOTUNE=1 measures one on a shape's first solve)."""
    m = sysm.mesh
    grp = sweep_columns(m.nn, k, sysm.nsys, sysm.is_complex)
    if grp < k:
        # the column sweep addresses a solve's vectors through 32-bit buffer offsets (float64 rows
        # under 2 GiB); a wider solve of one system runs as column groups that fit it
        parts, infos = [], []
        for c0 in range(0, k, grp):
            c1 = min(k, c0 + grp)
            Xi, inf = solve_system(sysm, B[:, c0:c1].contiguous(), c1 - c0, precision, tol, maxiter,
                                   None if X0 is None else X0[:, c0:c1].contiguous())
            parts.append(Xi)
            infos.append(inf)
        return torch.cat(parts, dim=1), SolveInfo(max(i.iters for i in infos), max(i.relres for i in infos),
                                                  max(i.flag for i in infos),
                                                  [(p.shape[1], i.iters) for p, i in zip(parts, infos)])
    from . import tune

    if tuned is None:
        tuned = tune.lookup(sysm, k, precision)
        if tuned is None and os.environ.get("RB_AUTOTUNE") == "1":
            tuned = tune.tune_system(sysm, B, k, precision, tol if tol > 0 else 1e-10, maxiter)
    # the LDS-staged SpMM serves 64-column solves of one system with 8-byte or 4-byte inner values
    # (complex64 or float32 mixed, float64 plain) and mixed solves of real 64-column systems in pairs
    lds_sets = lds_plan_sets(sysm.nsys, k, sysm.is_complex, precision)
    lds = lds_sets > 0 and (tuned is None or bool(tuned.get("lds", True)))
    check_every = int(tuned.get("check_every", 16)) if tuned else 16
    with tune.applied(tuned):
        X, info = krylov_solve(m.row_ptr, m.col_idx, m.diag, m.nn, m.nnz, sysm.vals64, sysm.vals32, B, k,
                               sysm.nsys, sysm.is_complex, precision, tol, maxiter, X0,
                               check_every=check_every, plan=m.spmm_plan(lds_sets) if lds else None)
    if info.flag == 4:
        if sysm.nsys == 1:
            return bicgstab_solve(m.row_ptr, m.col_idx, m.diag, m.nn, m.nnz, sysm.vals64, B, sysm.is_complex, tol,
                                  maxiter, X)
        parts = []
        for i in range(sysm.nsys):
            sl = slice(i * k, (i + 1) * k)
            Xi, inf = bicgstab_solve(m.row_ptr, m.col_idx, m.diag, m.nn, m.nnz, sysm.vals64[i], B[:, sl].contiguous(),
                                     sysm.is_complex, tol, maxiter, X[:, sl].contiguous())
            X[:, sl] = Xi
            parts.append(inf)
        info = SolveInfo(info.iters + max(p.iters for p in parts), max(p.relres for p in parts),
                         max(p.flag for p in parts))
    return X, info


# ---------------------------------------------------------------------------
# optodes and right-hand sides (redbirdpy forward.py:259-362)
# ---------------------------------------------------------------------------
def widefield(mesh, mode, frame, kind, npat, kx=0, ky=0, img=None, area=None):
    """Un-normalised wide-field columns (nn x npat, device numbering) and the
    (inside, inside and facing) triangle counts of one source (rb_widefield)."""
    dev = mesh.dev
    rhs = torch.empty((mesh.nn, npat), dtype=torch.float64, device=dev)
    counts = torch.zeros(2, dtype=torch.int64, device=dev)
    fr = (ctypes.c_double * _abi.RB_WF_FRAME)(*[float(v) for v in frame])
    imgt, pd0, pd1 = None, 0, 0
    if img is not None:
        imgt = _f64(np.ascontiguousarray(img, dtype=np.float64), dev)
        pd0, pd1 = int(img.shape[0]), int(img.shape[1])
    ar = mesh.area
    if mode == _abi.RB_WF_FACES and ar is None and area is not None:
        ar = _f64(area, dev)
    call("rb_widefield", mesh.nn, ptr(mesh.node), ptr(mesh.el), mesh.ne, ptr(mesh.tinc_ptr), ptr(mesh.tinc),
         ptr(mesh.fc), mesh.nf, ptr(ar), ptr(mesh.finc_ptr), ptr(mesh.finc), int(mode), ctypes.addressof(fr),
         int(kind), int(npat), int(kx), int(ky), ptr(imgt), pd0, pd1, ptr(rhs), ptr(counts), _abi.stream_ptr())
    return rhs, counts.cpu().numpy()


def tsearch(mesh, pts, tol=1e-10):
    """GPU point location; returns (loc0 int64 with -1 outside, bary (npts,4))."""
    pts = np.atleast_2d(np.asarray(pts, dtype=np.float64))[:, :3]
    npts = pts.shape[0]
    if npts == 0:
        return np.zeros(0, np.int64), np.zeros((0, 4))
    dev = mesh.dev
    p = _f64(pts, dev)
    loc = torch.empty(npts, dtype=torch.int32, device=dev)
    bary = torch.empty((npts, 4), dtype=torch.float64, device=dev)
    call("rb_tsearch", ptr(mesh.node), ptr(mesh.el), mesh.ne, ptr(p), npts, float(tol), ptr(loc), ptr(bary),
         _abi.stream_ptr())
    loc = loc.cpu().numpy().astype(np.int64)
    loc[loc == np.iinfo(np.int32).max] = -1
    return loc, bary.cpu().numpy()


def rhs_block(mesh, B, col0, loc0, bary, is_complex):
    """Write point-optode columns col0.. into the dense device block B (n x ld)."""
    ok = loc0 >= 0
    if not np.any(ok):
        return
    dev = mesh.dev
    cols = _i32(np.arange(len(loc0))[ok] + col0, dev)
    l = _i32(loc0[ok], dev)
    b = _f64(bary[ok], dev)
    ld = B.shape[1]
    call("rb_rhs_points", ptr(mesh.el), ptr(l), ptr(b), int(ok.sum()), ptr(cols), ld, int(is_complex), ptr(B),
         _abi.stream_ptr())


def zeros(shape, dtype, dev):
    t = torch.empty(shape, dtype=dtype, device=dev)
    call("rb_zero", ptr(t), t.numel() * t.element_size(), _abi.stream_ptr())
    return t


def to_host(t):
    """Device tensor -> numpy array through page-locked memory (a full-rate DMA; torch's host
    allocator caches the pinned block, so repeated copies of one size do not allocate again)."""
    h = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
    h.copy_(t)
    return h.numpy()


# ---------------------------------------------------------------------------
# detector readout and Jacobian
# ---------------------------------------------------------------------------
def getdet(det_ptr, det_row, det_val, phi, src0, nsrc, is_complex):
    """out[d, s] = sum_j w_j phi[row_j, src0 + s] for sparse detector columns."""
    ndet = det_ptr.numel() - 1
    dev = phi.device
    out = torch.empty((ndet, nsrc), dtype=torch.complex128 if is_complex else torch.float64, device=dev)
    if phi.stride(1) != 1:
        phi = phi.contiguous()
    call("rb_getdet", ndet, ptr(det_ptr), ptr(det_row), ptr(det_val), ptr(phi), phi.stride(0), src0, nsrc,
         int(is_complex), ptr(out), _abi.stream_ptr())
    return out


def pair_grid(src_cols, det_cols):
    """Reduce an ordered pair list to (unique srcs, unique dets, pmap, dup_rows).

    pmap[d_idx * Ns + s_idx] = output row of that pair (-1 when absent);
    dup_rows lists (row, first_row) for repeated pairs.
    """
    src_cols = np.asarray(src_cols, dtype=np.int64)
    det_cols = np.asarray(det_cols, dtype=np.int64)
    S, si = np.unique(src_cols, return_inverse=True)
    D, di = np.unique(det_cols, return_inverse=True)
    Ns, Nd = len(S), len(D)
    pmap = np.full(Ns * Nd, -1, dtype=np.int64)
    cell = di * Ns + si
    dups = []
    for r, c in enumerate(cell):
        if pmap[c] < 0:
            pmap[c] = r
        else:
            dups.append((r, int(pmap[c])))
    return S, D, pmap, dups


def patch_matrices(mesh):
    """Per-node patch matrices Q_n of the node Jacobian (rb_jac_patch_q), built once per mesh
    geometry and kept on the mesh (sum of np^2 doubles, ~0.2 GB at 115k nodes)."""
    key = (id(mesh.evol), mesh.evol.data_ptr())
    c = getattr(mesh, "_patch_q", None)
    if c is not None and c[0] == key:
        return c[1], c[2]
    npr = (mesh.row_ptr[1:] - mesh.row_ptr[:-1]).to(torch.int64)
    qptr = torch.zeros(mesh.nn + 1, dtype=torch.int64, device=mesh.dev)
    torch.cumsum(npr * npr, 0, out=qptr[1:])
    Qc = torch.empty(int(qptr[-1].item()), dtype=torch.float64, device=mesh.dev)
    call("rb_jac_patch_q", mesh.nn, ptr(mesh.el), ptr(mesh.evol), ptr(mesh.tinc_ptr), ptr(mesh.tinc),
         ptr(mesh.row_ptr), ptr(mesh.col_idx), mesh.maxnp, ptr(qptr), ptr(Qc), _abi.stream_ptr())
    mesh._patch_q = (key, qptr, Qc)
    return qptr, Qc


def jac_node(mesh, phi, src_cols, det_cols, out_dtype=torch.float64, want_elem=False, out=None):
    """JT (nn x npairs) node-major Jacobian (and optional element Jacobian, ne x npairs).

    ``out`` may be a column slice of a wider node-major block (row stride
    out.stride(0), unit column stride) that receives the Jacobian in place.
    """
    npairs = len(src_cols)
    dev = phi.device
    is_complex = phi.is_complex()
    S, D, pmap, dups = pair_grid(src_cols, det_cols)
    full = (len(S) * len(D) == npairs) and np.array_equal(pmap, np.arange(npairs))
    code = {torch.float64: _abi.RB_F64, torch.float32: _abi.RB_F32, torch.complex128: _abi.RB_C128,
            torch.complex64: _abi.RB_C64}[out_dtype]
    if out is not None:
        if out.shape != (mesh.nn, npairs) or out.dtype != out_dtype or out.stride(1) != 1:
            raise ValueError("jac_node: out must be an (nn, npairs) view with unit column stride")
        JT = out
    else:
        JT = torch.empty((mesh.nn, npairs), dtype=out_dtype, device=dev)
    JE = torch.empty((mesh.ne, npairs), dtype=out_dtype, device=dev) if want_elem else None
    if npairs == 0:
        return JT, JE
    s_d = _i32(S, dev)
    d_d = _i32(D, dev)
    pm = None if full else _i32(pmap, dev)
    phi = phi.contiguous()
    qptr, Qc = patch_matrices(mesh)
    call("rb_jac_grid", mesh.nn, mesh.ne, ptr(mesh.el), ptr(mesh.evol), ptr(mesh.tinc_ptr), ptr(mesh.tinc),
         ptr(mesh.row_ptr), ptr(mesh.col_idx), mesh.maxnp, ptr(phi), phi.shape[1], int(is_complex), len(S), len(D),
         ptr(s_d), ptr(d_d), ptr(pm), ptr(JT), JT.stride(0), code, ptr(JE), ptr(qptr), ptr(Qc), _abi.stream_ptr())
    for r, f in dups:
        JT[:, r] = JT[:, f]
        if JE is not None:
            JE[:, r] = JE[:, f]
    return JT, JE


def jac_contract(mesh, phi, S, D, w):
    """g[n] = sum_{d,s} w[d, s] Jnode[n, (s, d)] for full source x detector grids."""
    dev = phi.device
    is_complex = phi.is_complex()
    Ns, Nd = len(S), len(D)
    dt = phi.dtype
    w = w.to(dt).resolve_conj().resolve_neg().contiguous()
    psi = torch.empty((mesh.nn, Ns), dtype=dt, device=dev)
    h = torch.empty(mesh.ne, dtype=dt, device=dev)
    g = torch.empty(mesh.nn, dtype=dt, device=dev)
    s_d = _i32(S, dev)
    d_d = _i32(D, dev)
    phi = phi.contiguous()
    call("rb_jac_contract_grid", mesh.nn, mesh.ne, ptr(mesh.el), ptr(mesh.evol), ptr(mesh.tinc_ptr),
         ptr(mesh.tinc), ptr(phi), phi.shape[1], int(is_complex), Ns, Nd, ptr(s_d), ptr(d_d), ptr(w), ptr(psi),
         ptr(h), ptr(g), _abi.stream_ptr())
    return g


GRAM_MODES = {"fp64": 0, "bf16": 1, "bf16x3": 3}


def gram(JT, d=None, mode="fp64", want_colsq=False):
    """G = diag(d) JT^T JT diag(d) in float64 (m x m); mode fp64 | bf16 | bf16x3 (MFMA).

    want_colsq (MFMA modes): also return the exact float64 sum_k JT[k, c]^2,
    taken from the same read of JT.
    """
    K, m = JT.shape
    G = torch.empty((m, m), dtype=torch.float64, device=JT.device)
    code = _abi.RB_F64 if JT.dtype == torch.float64 else _abi.RB_F32
    md = GRAM_MODES[mode] if isinstance(mode, str) else int(mode)
    if md == 0 and d is None and JT.dtype == torch.float64:
        # plain float64 SYRK: the library DGEMM (MFMA, ~32 TFLOP/s at m=4096,
        # K=115k, 61 ms) outruns rb_gram's FMA kernel (87 ms)
        return torch.matmul(JT.t(), JT)
    if md == 0:
        call("rb_gram", K, m, ptr(JT), JT.stride(0), code, ptr(d), 0, ptr(G), _abi.stream_ptr())
        return G
    nb = ctypes.c_size_t(0)
    call("rb_gram_workspace", K, m, md, ctypes.byref(nb))
    work = torch.empty(int(nb.value), dtype=torch.uint8, device=JT.device)
    colsq = torch.empty(m, dtype=torch.float64, device=JT.device) if want_colsq else None
    call("rb_gram_mfma", K, m, ptr(JT), JT.stride(0), code, ptr(d), md, ptr(G), ptr(colsq), ptr(work), nb.value,
         _abi.stream_ptr())
    return (G, colsq) if want_colsq else G


def jt_rows(JT, nw, npairs, v):
    """t[n, w] = sum_p JT[n, w*np + p] v[w*np + p] in float64 (J_w^T v_w for every wavelength)."""
    K, M = JT.shape
    t = torch.empty((K, nw), dtype=torch.float64, device=JT.device)
    code = _abi.RB_F64 if JT.dtype == torch.float64 else _abi.RB_F32
    v = v.to(torch.float64).contiguous()
    call("rb_jt_rows", K, M, ptr(JT), JT.stride(0), code, nw, npairs, ptr(v), ptr(t), _abi.stream_ptr())
    return t


def jt_hv_ok(JT, nw, npairs):
    """Whether rb_jt_hv takes this shape (float64, np a multiple of 64, nw <= 16, M <= 8192)."""
    K, M = JT.shape
    return (JT.dtype == torch.float64 and JT.stride(1) == 1 and npairs % 64 == 0 and 0 < nw <= 16
            and nw * npairs == M and M <= 8192)


def jt_hv(JT, W, npairs, v, s2, lam):
    """s2 * J W (J^T v) + lam * v in one read of JT (float64 accumulation)."""
    K, M = JT.shape
    nw = W.shape[0]
    W = W.to(torch.float64).contiguous()
    v = v.to(torch.float64).contiguous()
    out = torch.empty(M, dtype=torch.float64, device=JT.device)
    nb = ctypes.c_size_t(0)
    call("rb_jt_hv_workspace", K, M, ctypes.byref(nb))
    work = torch.empty(int(nb.value), dtype=torch.uint8, device=JT.device)
    call("rb_jt_hv", K, M, ptr(JT), JT.stride(0), nw, npairs, ptr(W), ptr(v), float(s2), float(lam), ptr(out),
         ptr(work), nb.value, _abi.stream_ptr())
    return out


def cholesky(H):
    """Lower Cholesky factor of a float64 SPD device matrix (blocked 64 x 64, rb_potrf).

    n is padded to a multiple of 64 with an identity block.  Raises
    torch.linalg.LinAlgError at a non-positive pivot, like torch.linalg.cholesky.
    """
    n = H.shape[0]
    npad = -(-n // 64) * 64
    A = torch.zeros((npad, npad), dtype=torch.float64, device=H.device)
    A[:n, :n] = H
    if npad > n:
        A.diagonal()[n:] = 1.0
    nb = ctypes.c_size_t(0)
    call("rb_potrf_workspace", npad, ctypes.byref(nb))
    work = torch.empty(int(nb.value), dtype=torch.uint8, device=H.device)
    info = torch.empty(1, dtype=torch.int32, device=H.device)
    call("rb_potrf", npad, ptr(A), npad, ptr(work), nb.value, ptr(info), _abi.stream_ptr())
    bad = int(info.item())
    if bad:
        raise torch.linalg.LinAlgError(f"cholesky: the leading minor of order {bad} is not positive-definite")
    L = A[:n, :n].tril_()
    return L if npad == n else L.contiguous()


def remap_operator(mapid, mapweight, relem, nrec, perm_rows=None, dev=None):
    """Device CSR (ptr, idx, w) over recon nodes of the interpolation's transpose (recon.py:848-874 loop order).

    Forward node i (ascending) located in recon element mapid[i] (1-based,
    NaN outside) gives weight mapweight[i, j] to recon node relem[e, j].
    perm_rows maps user forward-node numbers to device rows (iperm).
    """
    dev = dev or device()
    mid = np.asarray(mapid, dtype=float)
    mw = np.asarray(mapweight, dtype=float)
    e0 = np.asarray(relem)[:, :4].astype(np.int64) - 1
    fwd = np.flatnonzero(~np.isnan(mid))
    e = mid[fwd].astype(np.int64) - 1
    ok = (e >= 0) & (e < e0.shape[0])
    fwd, e = fwd[ok], e[ok]
    nodes = e0[e]
    ok = np.all((nodes >= 0) & (nodes < nrec), axis=1)
    fwd, nodes = fwd[ok], nodes[ok]
    rows = nodes.ravel()
    cols = np.repeat(fwd, 4)
    vals = mw[fwd].ravel()
    order = np.lexsort((np.tile(np.arange(4), len(fwd)), cols, rows))  # per recon node: forward nodes ascending
    rows, cols, vals = rows[order], cols[order], vals[order]
    ptr_ = np.zeros(nrec + 1, dtype=np.int64)
    np.cumsum(np.bincount(rows, minlength=nrec), out=ptr_[1:])
    if perm_rows is not None:
        cols = perm_rows[cols]
    return _i32(ptr_, dev), _i32(cols, dev), _f64(vals, dev)


def jt_remap(op, JT, nrec):
    """(nrec x M) recon-mesh Jacobian rows from the node-major forward Jacobian JT (rb_jt_remap)."""
    p, i, w = op
    M = JT.shape[1]
    Y = torch.empty((nrec, M), dtype=torch.float64, device=JT.device)
    call("rb_jt_remap", nrec, ptr(p), ptr(i), ptr(w), M, ptr(JT), JT.stride(0), ptr(Y), Y.stride(0),
         _abi.stream_ptr())
    return Y


def jt_cols(JT, nw, npairs, z=None):
    """u[j] = sum_n JT[n, j] z[n, j // np] (J_w z_w), or sum_n JT[n, j]^2 when z is None."""
    K, M = JT.shape
    u = torch.empty(M, dtype=torch.float64, device=JT.device)
    code = _abi.RB_F64 if JT.dtype == torch.float64 else _abi.RB_F32
    nb = ctypes.c_size_t(0)
    call("rb_jt_cols_workspace", K, M, ctypes.byref(nb))
    work = torch.empty(int(nb.value), dtype=torch.uint8, device=JT.device)
    zz = None if z is None else z.to(torch.float64).contiguous()
    call("rb_jt_cols", K, M, ptr(JT), JT.stride(0), code, nw, npairs, ptr(zz), 0 if z is not None else 1, ptr(u),
         ptr(work), nb.value, _abi.stream_ptr())
    return u
