"""Structured tetrahedral meshes for synthetic workloads.

The reference builds its test meshes with ``iso2mesh.meshabox``
(test/test_forward.py:33), which is not available here.  This module
generates the equivalent box geometry as a Kuhn (Freudenthal) subdivision:
every grid cube is split into six tetrahedra, one per permutation of the
axes, along the monotone path corner -> +e_a -> +e_a+e_b -> opposite corner.

Numbering is fixed so that meshes are reproducible bit for bit:

* node id = i + nx*(j + ny*k)   (x fastest), stored 0-based in ``node``;
* cube id = i + (nx-1)*(j + (ny-1)*k), six tets per cube in lexicographic
  permutation order (xyz, xzy, yxz, yzx, zxy, zyx);
* ``elem`` is returned 1-based, like iso2mesh, and *not* reoriented: the
  odd permutations have negative signed volume, so ``meshprep`` has real
  reorientation work to do.
"""

import numpy as np

__all__ = ["kuhn_box", "sphere_labels", "kuhn_sizes"]

_PERMS = np.array(
    [[0, 1, 2], [0, 2, 1], [1, 0, 2], [1, 2, 0], [2, 0, 1], [2, 1, 0]], dtype=np.int64
)


def kuhn_sizes(p0, p1, h):
    """Return (nx, ny, nz) grid node counts for a box [p0, p1] at spacing h."""
    p0 = np.asarray(p0, dtype=float)
    p1 = np.asarray(p1, dtype=float)
    n = np.rint((p1 - p0) / float(h)).astype(np.int64) + 1
    if np.any(n < 2):
        raise ValueError("box must span at least one cell in each direction")
    return int(n[0]), int(n[1]), int(n[2])


def kuhn_box(p0, p1, h):
    """Mesh the box [p0, p1] with the 6-tet Kuhn subdivision at spacing h.

    Returns ``(node, elem)``: node is (nn, 3) float64, elem is (ne, 4) int64
    with 1-based node indices.
    """
    p0 = np.asarray(p0, dtype=float)
    p1 = np.asarray(p1, dtype=float)
    nx, ny, nz = kuhn_sizes(p0, p1, h)
    hx = (p1 - p0) / np.array([nx - 1, ny - 1, nz - 1], dtype=float)

    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    node = np.column_stack(
        [p0[0] + i.ravel() * hx[0], p0[1] + j.ravel() * hx[1], p0[2] + k.ravel() * hx[2]]
    )

    ck, cj, ci = np.meshgrid(
        np.arange(nz - 1), np.arange(ny - 1), np.arange(nx - 1), indexing="ij"
    )
    corner = np.column_stack([ci.ravel(), cj.ravel(), ck.ravel()])  # (ncube, 3)
    stride = np.array([1, nx, nx * ny], dtype=np.int64)
    base = corner @ stride  # (ncube,)

    ncube = base.shape[0]
    elem = np.empty((ncube, 6, 4), dtype=np.int64)
    for t, perm in enumerate(_PERMS):
        step0 = stride[perm[0]]
        step1 = stride[perm[1]]
        elem[:, t, 0] = base
        elem[:, t, 1] = base + step0
        elem[:, t, 2] = base + step0 + step1
        elem[:, t, 3] = base + stride.sum()
    return node, elem.reshape(ncube * 6, 4) + 1


def sphere_labels(node, elem, center, radius, inside=2, outside=1):
    """Label tets whose centroid lies within ``radius`` of ``center``.

    ``elem`` is 1-based.  Returns an int64 array of length ne.
    """
    elem_0 = np.asarray(elem)[:, :4].astype(np.int64) - 1
    cent = node[elem_0].mean(axis=1)
    d2 = np.sum((cent - np.asarray(center, dtype=float)) ** 2, axis=1)
    return np.where(d2 <= float(radius) ** 2, inside, outside).astype(np.int64)
