"""Optical properties (drop-in for redbirdpy.property on the hot path).

``extinction`` reads the built-in absorption tables (hemoglobin from S.
Prahl's compilation in data/hemoglobin_prahl.csv; water, lipids and aa3 as
defined in redbirdpy property.py:507-575) and interpolates linearly with
linear extrapolation, like scipy's interp1d(fill_value='extrapolate').
``updateprop`` and ``jacchrome``/``runrecon`` accept an optional ``extin``
override (a dict wavelength -> {chromophore: value} or an Nwv x Nchrome
matrix) in place of the table; with it, no default water/lipids are added.
"""

__all__ = ["extinction", "updateprop", "getbulk", "get_chromophore_table", "musp2sasp", "setmesh"]

import os
import weakref

import numpy as np

_HB_CSV = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "hemoglobin_prahl.csv")
_TABLES = None
CHROMOPHORES = ["hbo", "hbr", "water", "lipids", "aa3"]


def _tables():
    global _TABLES
    if _TABLES is None:
        hb = np.loadtxt(_HB_CSV, delimiter=",", comments="#")
        to_mm_uM = 2.303e-7  # 1/(cm M) -> 1/(mm uM), property.py:538-540
        t = {
            "hbo": np.column_stack([hb[:, 0], hb[:, 1] * to_mm_uM]),
            "hbr": np.column_stack([hb[:, 0], hb[:, 2] * to_mm_uM]),
        }
        wv = np.array([400, 500, 600, 650, 700, 750, 800, 850, 900, 950, 1000], dtype=float)
        mu_cm = np.array([0.00058, 0.00025, 0.0023, 0.0032, 0.006, 0.026, 0.02, 0.043, 0.068, 0.39, 0.36])
        t["water"] = np.column_stack([wv, mu_cm * 0.1])
        lw = np.arange(650, 1000, 10).astype(float)
        t["lipids"] = np.column_stack([lw, np.full(lw.shape, 0.0005)])
        aw = np.arange(650, 950, 5).astype(float)
        t["aa3"] = np.column_stack([aw, 0.5 * np.exp(-((aw - 830.0) ** 2) / (2 * 50.0 ** 2)) + 0.4])
        _TABLES = t
    return _TABLES


def get_chromophore_table(name):
    t = _tables()
    key = name.lower()
    if key not in t:
        raise ValueError("Unknown chromophore: %s. Available: %s" % (name, list(t)))
    return t[key].copy()


def _lerp_extrapolate(x, y, q):
    """Piecewise-linear interpolation that extends the end segments."""
    i = np.searchsorted(x, q, side="left")
    i = np.clip(i, 1, len(x) - 1)
    x0, x1 = x[i - 1], x[i]
    y0, y1 = y[i - 1], y[i]
    return y0 + (q - x0) * ((y1 - y0) / (x1 - x0))


def extinction(wavelengths, chromophores, **interp_opts):
    """(Nwv x Nchrome) extinction coefficients and the tables (property.py:30-89)."""
    t = _tables()
    if isinstance(wavelengths, (list, tuple)):
        wavelengths = [float(w) if isinstance(w, str) else w for w in wavelengths]
    w = np.atleast_1d(wavelengths).astype(float)
    if isinstance(chromophores, str):
        chromophores = [chromophores]
    out = np.zeros((len(w), len(chromophores)))
    for j, c in enumerate(chromophores):
        key = c.lower()
        if key not in t:
            raise ValueError("Unknown chromophore: %s. Available: %s" % (c, list(t)))
        out[:, j] = _lerp_extrapolate(t[key][:, 0], t[key][:, 1], w)
    return out, t


def _extin_row(extin, wv, types):
    if isinstance(extin, dict):
        row = extin[wv]
        return np.array([row[t] for t in types], dtype=float)
    raise ValueError("extin for updateprop must be a dict wavelength -> {chromophore: value}")


def updateprop(cfg, wv=None, extin=None):
    """Chromophore concentrations -> per-wavelength [mua, musp, g, n] (property.py:92-213)."""
    if "param" not in cfg or not isinstance(cfg.get("prop"), dict):
        return cfg.get("prop")
    wvs = list(cfg["prop"].keys()) if wv is None else [wv]
    par = cfg["param"]
    out = {}
    for w in wvs:
        if extin is None:
            par.setdefault("water", 0.23)
            par.setdefault("lipids", 0.58)
        types = [t for t in CHROMOPHORES if t in par]
        if not types:
            raise ValueError(
                "No recognized chromophores in cfg.param. Expected one or more of: hbo, hbr, water, lipids, aa3"
            )
        ex = extinction(float(w), types)[0][0] if extin is None else _extin_row(extin, w, types)
        first = par[types[0]]
        mua = 0.0 if np.isscalar(first) else np.zeros_like(first, dtype=float)
        for j, t in enumerate(types):
            mua = mua + ex[j] * par[t]
        musp = None
        if "scatamp" in par and "scatpow" in par:
            musp = par["scatamp"] * float(w) ** (-par["scatpow"])
        tmpl = cfg["prop"][w]
        nn = cfg["node"].shape[0] if "node" in cfg else np.size(mua)
        ne = cfg["elem"].shape[0] if "elem" in cfg else np.size(mua)
        mlen = len(mua) if hasattr(mua, "__len__") else 1
        if mlen < min(nn, ne):
            newp = tmpl.copy()
            rows = slice(1, mlen + 1) if hasattr(mua, "__len__") else 1
            newp[rows, 0] = mua
            if musp is not None:
                newp[rows, 1] = musp
                newp[rows, 2] = 0
        elif musp is not None:
            nref = tmpl[1, 3] if tmpl.shape[0] > 1 else 1.37
            if hasattr(mua, "__len__"):
                # column_stack in the reference: complex concentrations give complex rows
                newp = np.empty((len(mua), 4), dtype=np.result_type(np.asarray(mua), np.asarray(musp), float))
                newp[:, 0] = mua
                newp[:, 1] = musp
                newp[:, 2] = 0.0
                newp[:, 3] = nref
            else:
                newp = np.array([[mua, musp, 0, nref]])
        elif hasattr(mua, "__len__"):
            # the template row repeated (a contiguous copy), then the mua column
            dt = np.result_type(np.asarray(mua), tmpl)
            newp = np.repeat(tmpl[1:2].astype(dt, copy=False), len(mua), axis=0)
            newp[:, 0] = mua
        else:
            newp = np.array([[mua] + list(tmpl[1, 1:])])
        out[w] = newp
    return out if len(wvs) > 1 else out[wvs[0]]


_DEFAULT_BULK = (0.0, 0.0, 0.0, 1.37)


_FIRST_ELEM_CACHE = {}  # (id(elem), id(face)) -> (weakref elem, weakref face, first element)


def _first_face_elem(cfg):
    """First element holding the first face's first node (a scan of every
    element, cached per (elem, face) array pair: getoptodes asks on every
    forward call)."""
    elem, face = cfg["elem"], cfg["face"]
    key = (id(elem), id(face))
    hit = _FIRST_ELEM_CACHE.get(key)
    if hit is not None and hit[0]() is elem and hit[1]() is face:
        return hit[2]
    rows = np.flatnonzero(np.any(np.asarray(elem)[:, :4] == face[0, 0], axis=1))
    first = int(rows[0]) if len(rows) else None
    try:
        if len(_FIRST_ELEM_CACHE) > 16:
            _FIRST_ELEM_CACHE.clear()
        _FIRST_ELEM_CACHE[key] = (weakref.ref(elem), weakref.ref(face), first)
    except TypeError:  # not weak-referenceable (e.g. a list): no caching
        pass
    return first


def _bulk_of(prop, cfg):
    """Bulk property row from a label/node/element table (property.py:264-332)."""
    dflt = np.array(_DEFAULT_BULK)
    nn = cfg["node"].shape[0] if "node" in cfg else prop.shape[0]
    ne = cfg["elem"].shape[0] if "elem" in cfg else prop.shape[0]
    if prop.shape[0] < min(nn, ne):
        if "seg" in cfg and cfg["seg"] is not None:
            seg = cfg["seg"]
            if "face" in cfg and len(seg) == nn:
                label = seg[cfg["face"][0, 0] - 1]
            elif "face" in cfg and len(seg) == ne:
                e = _first_face_elem(cfg)
                label = seg[e] if e is not None else 0
            else:
                label = seg[0]
            label = int(label)
            if prop.shape[0] > label:
                return prop[label, :]
        return prop[1, :] if prop.shape[0] > 1 else dflt
    if prop.shape[0] == nn:
        return prop[cfg["face"][0, 0] - 1, :] if "face" in cfg else prop[0, :]
    if prop.shape[0] == ne:
        if "face" in cfg:
            e = _first_face_elem(cfg)
            if e is not None:
                return prop[e, :]
        return prop[0, :]
    return dflt


def getbulk(cfg):
    """Bulk [mua, mus, g, n]: cfg.bulk > surface property > default (property.py:216-261)."""
    if "bulk" in cfg:
        b = cfg["bulk"]
        p = np.array(_DEFAULT_BULK)
        if "mua" in b:
            p[0] = b["mua"]
        if "dcoeff" in b:
            p[1], p[2] = 1.0 / (3 * b["dcoeff"]), 0
        if "musp" in b:
            p[1], p[2] = b["musp"], 0
        if "g" in b:
            p[2] = b["g"]
        if "n" in b:
            p[3] = b["n"]
        return p
    if cfg.get("prop") is None:
        return np.array(_DEFAULT_BULK)
    prop = cfg["prop"]
    if isinstance(prop, dict):
        return {wv: _bulk_of(np.atleast_2d(p), cfg) for wv, p in prop.items()}
    return _bulk_of(np.atleast_2d(prop), cfg)


def musp2sasp(musp, wavelength):
    """Scattering amplitude and power from mus' at two wavelengths (property.py:335-366).

    Model mus'(lambda) = sa (lambda / 500 nm)^(-sp): sp from the ratio of the
    two values, sa the mean of the two back-solved amplitudes.
    """
    musp = np.asarray(musp, dtype=float)
    wavelength = np.asarray(wavelength, dtype=float)
    if len(musp) < 2 or len(wavelength) < 2:
        raise ValueError("Need at least 2 wavelengths to fit scattering parameters")
    x = wavelength[:2] / 500.0
    sp = np.log(musp[0] / musp[1]) / np.log(x[1] / x[0])
    sa = 0.5 * (musp[0] * x[0] ** sp + musp[1] * x[1] ** sp)
    return sa, sp


_MESH_DERIVED = ("face", "area", "evol", "deldotdel", "isreoriented", "nvol", "cols", "rows", "idxsum", "idxcount",
                 "musp0", "reff")


def setmesh(cfg0, node, elem, prop=None, propidx=None):
    """A copy of cfg0 on a new mesh, its mesh-derived fields recomputed by meshprep (property.py:369-432).

    elem is 1-based; a fifth column (or propidx) gives the labels.
    """
    from .utility import meshprep

    elem = np.asarray(elem)
    cfg = {k: v for k, v in cfg0.items() if k not in _MESH_DERIVED}
    cfg["node"] = node
    cfg["elem"] = elem[:, :4] if elem.shape[1] > 4 else elem
    if prop is not None:
        cfg["prop"] = prop
    if propidx is not None:
        cfg["seg"] = propidx
    elif elem.shape[1] > 4:
        cfg["seg"] = elem[:, 4].astype(int)
    cfg, _ = meshprep(cfg)
    return cfg
