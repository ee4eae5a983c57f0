"""Synthetic workloads on a 60 x 60 x 30 mm box (the benchmark configurations).

* ``cw_box``: homogeneous box at 2 mm spacing (31x31x16 = 15,376 nodes,
  81,000 tets), mua 0.01/mm, mus' 1.0/mm, n 1.37, one source at (30,30,0),
  detectors at (30,40,0) and (40,30,0), CW.
* ``fd_box``: the box at 1 mm spacing (61x61x31 = 115,351 nodes, 648,000
  tets), background [0.008, 1.0, 0, 1.37], a sphere of radius 8 mm at
  (30,30,15) with [0.02, 1.2, 0, 1.37] (tets whose centroid lies inside),
  omega = 2 pi 100 MHz, 32 sources and 32 detectors interleaved on an 8x8
  lattice of z=0 points x, y in linspace(10, 50, 8): even lattice rows are
  sources, odd rows detectors.
* ``jac_box``: the 1 mm box and optodes in CW with homogeneous
  [0.01, 1.0, 0, 1.37] (1024 pairs x 115,351 nodes).
* ``widefield_box``: ``cw_box`` with its source turned into a wide-field
  one: a 20 x 20 mm rectangle from (30,30,0) (planar, pattern or fourier
  with kx, ky in the 4th parameter entries), collimated along +z.
* ``chrome_box``: the 1 mm box at 4 wavelengths with 3 chromophores whose
  4x3 extinction matrix is uniform in [5e-4, 5e-3] /mm/uM (seeded), node
  concentrations 20/10/0.5 doubled inside the sphere.

Every builder takes a ``h`` (spacing) so tests can use coarser versions.
"""

import numpy as np

from .mesh import kuhn_box, sphere_labels

BOX = (60.0, 60.0, 30.0)
SPHERE_C = (30.0, 30.0, 15.0)
SPHERE_R = 8.0
OMEGA_100MHZ = 2 * np.pi * 100e6


def lattice_optodes(n=8, lo=10.0, hi=50.0):
    """Interleaved sources/detectors: lattice rows with even y index are sources."""
    xs = np.linspace(lo, hi, n)
    src, det = [], []
    for j, y in enumerate(xs):
        row = [[x, y, 0.0] for x in xs]
        (src if j % 2 == 0 else det).extend(row)
    return np.array(src), np.array(det)


def _base(h, box=BOX):
    node, elem = kuhn_box([0, 0, 0], box, h)
    return node, elem


def cw_box(h=2.0, box=BOX):
    node, elem = _base(h, box)
    return {
        "node": node,
        "elem": elem,
        "prop": np.array([[0, 0, 1, 1], [0.01, 1.0, 0, 1.37]]),
        "seg": np.ones(elem.shape[0], dtype=int),
        "srcpos": np.array([[box[0] / 2, box[1] / 2, 0.0]]),
        "srcdir": np.array([[0.0, 0.0, 1.0]]),
        "detpos": np.array([[box[0] / 2, box[1] / 2 + 10, 0.0], [box[0] / 2 + 10, box[1] / 2, 0.0]]),
        "detdir": np.array([[0.0, 0.0, 1.0]]),
        "omega": 0,
    }


def widefield_box(h=2.0, srctype="planar", kx=0, ky=0, pattern=None, box=BOX):
    cfg = cw_box(h, box)
    cfg["srctype"] = srctype
    cfg["srcparam1"] = np.array([20.0, 0.0, 0.0, float(kx)])
    cfg["srcparam2"] = np.array([0.0, 20.0, 0.0, float(ky)])
    if pattern is not None:
        cfg["srcpattern"] = np.asarray(pattern, dtype=float)
    return cfg


def batch_boxes(count=64, seed=5, sides=(20.0, 80.0), spacing=(0.8, 2.5), nsrc=(8, 128), nn_range=(2000, 200000),
                kf=4, omega=0.0):
    """Uneven batch: Kuhn boxes with seeded side lengths and spacing (resampled until the node
    count lies in nn_range), a seeded spherical inclusion (mua in [0.005, 0.03], mus' in
    [0.5, 1.5], radius 3-10 mm, at most 0.3 of the shortest side) on a [0.01, 1.0] background, nsrc point sources on a z=0
    grid plus kf*kf Fourier patterns over the top face, and as many detectors."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        box = tuple(rng.uniform(*sides, 3))
        h = rng.uniform(*spacing)
        dims = [int(round(b / h)) + 1 for b in box]
        nn = dims[0] * dims[1] * dims[2]
        if not (nn_range[0] <= nn <= nn_range[1]):
            continue
        box = tuple((d - 1) * h for d in dims)
        node, elem = kuhn_box([0, 0, 0], box, h)
        r = min(rng.uniform(3.0, 10.0), 0.3 * min(box))  # the sphere stays inside the box
        cen = np.array([rng.uniform(r, b - r) for b in box])
        inc = sphere_labels(node, elem, cen, r)
        prop = np.array([[0, 0, 1, 1], [0.01, 1.0, 0, 1.37],
                         [rng.uniform(0.005, 0.03), rng.uniform(0.5, 1.5), 0, 1.37]])
        ns = int(rng.integers(nsrc[0], nsrc[1] + 1))
        g = int(np.ceil(np.sqrt(ns)))
        xs = np.linspace(0.15 * box[0], 0.85 * box[0], g)
        ys = np.linspace(0.15 * box[1], 0.85 * box[1], g)
        grid = np.array([[x, y, 0.0] for y in ys for x in xs])[:ns]
        cfg = {"node": node, "elem": elem, "prop": prop, "seg": inc,
               "srcpos": grid, "srcdir": np.array([[0.0, 0.0, 1.0]]),
               "detpos": grid + np.array([0.5 * (xs[1] - xs[0]) if g > 1 else 1.0, 0.0, 0.0]),
               "detdir": np.array([[0.0, 0.0, 1.0]]), "omega": omega}
        if kf:
            cfg["widesrcid"] = {"": {"srctype": ["fourier"], "srcid": [ns],
                                     "srcparam1": [np.array([0.7 * box[0], 0, 0, kf])],
                                     "srcparam2": [np.array([0, 0.7 * box[1], 0, kf])]}}
            cfg["srcpos"] = np.vstack([grid, [[0.15 * box[0], 0.15 * box[1], 0.0]]])
        out.append(cfg)
    return out


def fd_box(h=1.0, box=BOX, n_lattice=8, omega=OMEGA_100MHZ):
    node, elem = _base(h, box)
    lab = sphere_labels(node, elem, SPHERE_C, SPHERE_R, inside=2, outside=1)
    lo, hi = box[0] / 6, box[0] * 5 / 6
    src, det = lattice_optodes(n_lattice, lo, hi)
    return {
        "node": node,
        "elem": elem,
        "prop": np.array([[0, 0, 1, 1], [0.008, 1.0, 0, 1.37], [0.02, 1.2, 0, 1.37]]),
        "seg": lab,
        "srcpos": src,
        "srcdir": np.array([[0.0, 0.0, 1.0]]),
        "detpos": det,
        "detdir": np.array([[0.0, 0.0, 1.0]]),
        "omega": omega,
    }


def jac_box(h=1.0, box=BOX, n_lattice=8):
    cfg = fd_box(h, box, n_lattice, omega=0)
    cfg["prop"] = np.array([[0, 0, 1, 1], [0.01, 1.0, 0, 1.37]])
    cfg["seg"] = np.ones(cfg["elem"].shape[0], dtype=int)
    return cfg


WAVELENGTHS = ["690", "750", "800", "850"]
CHROMES = ["hbo", "hbr", "water"]


def chrome_extinction(seed=4):
    E = np.random.default_rng(seed).uniform(5e-4, 5e-3, (4, 3))
    return {wv: {c: E[i, j] for j, c in enumerate(CHROMES)} for i, wv in enumerate(WAVELENGTHS)}, E


CHROME_BASE = {"hbo": 20.0, "hbr": 10.0, "water": 0.5}


def chrome_box(h=1.0, box=BOX, n_lattice=8):
    """Multispectral setup: returns (cfg, background params, true params, extin dict, E).

    cfg carries label properties per wavelength (background optics) and no
    'param', so meshprep does not run updateprop on it; the caller puts the
    concentrations in after meshprep (or in recon['param']).
    """
    cfg = jac_box(h, box, n_lattice)
    node = cfg["node"]
    nn = node.shape[0]
    inside = np.sum((node - np.array(SPHERE_C)) ** 2, axis=1) <= SPHERE_R ** 2
    truth = {c: np.where(inside, 2 * v, v).astype(float) for c, v in CHROME_BASE.items()}
    background = {c: np.full(nn, v) for c, v in CHROME_BASE.items()}
    cfg["prop"] = {wv: np.array([[0, 0, 1, 1], [0.01, 1.0, 0, 1.37]]) for wv in WAVELENGTHS}
    extin, E = chrome_extinction()
    return cfg, background, truth, extin, E
