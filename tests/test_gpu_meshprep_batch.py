"""meshprep_batch: the geometry of many meshes in one GPU launch sequence (rb_tet_prep,
rb_boundary_faces, rb_face_area, rb_node_volume, rb_csr_pattern) against meshprep of each mesh
alone (host reorientation, np.unique face extraction, volumes, deldotdel, GPU CSR): integer fields
bit-exact, floating-point fields to 1e-13, the same sd maps and wide-field columns."""

import copy

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _meshes():
    from conftest import box_cfg, grid_optodes
    from redbirdpy_amd import workloads
    from redbirdpy_amd.mesh import kuhn_box

    out = [box_cfg((12, 10, 6), 2.0), box_cfg((20, 20, 10), 2.5, omega=2 * np.pi * 1e8)]
    c = box_cfg((16, 12, 8), 2.0)  # half the tets given with negative orientation
    c["elem"] = c["elem"].copy()
    c["elem"][::2, [0, 1]] = c["elem"][::2, [1, 0]]
    out.append(c)
    node, elem = kuhn_box([0, 0, 0], (16, 16, 8), 2.0)  # centroid-split tets: irregular rows
    cen = node[elem[:, :4] - 1].mean(axis=1)
    split = np.flatnonzero(np.sum((cen - [8.0, 8.0, 4.0]) ** 2, axis=1) < 12.0)
    keep = np.setdiff1d(np.arange(len(elem)), split)
    ids = node.shape[0] + 1 + np.arange(len(split))
    quads = []
    for (a, b, cc, d), x in zip(elem[split, :4], ids):
        quads += [(a, b, cc, x), (a, b, x, d), (a, x, cc, d), (x, b, cc, d)]
    c = box_cfg((16, 16, 8), 2.0, src=grid_optodes((16, 16, 8), 2), det=grid_optodes((16, 16, 8), 3, margin=3.0))
    c["node"] = np.vstack([node, cen[split]])
    c["elem"] = np.vstack([elem[keep, :4], np.array(quads)]).astype(int)
    c["seg"] = np.ones(len(c["elem"]), dtype=int)
    out.append(c)
    out.append(workloads.widefield_box(2.0, "fourier", kx=2, ky=2))  # wide-field source columns
    cfg0, bg, truth, extin, E = workloads.chrome_box(h=5.0, n_lattice=3)  # multi-wavelength sd maps
    out.append(cfg0)
    return out


def test_meshprep_batch_matches_meshprep():
    import redbirdpy_amd as rb

    cfgs = _meshes()
    single = [rb.meshprep(copy.deepcopy(c)) for c in cfgs]
    batch = rb.meshprep_batch([copy.deepcopy(c) for c in cfgs])
    assert len(batch) == len(cfgs)
    for (c1, sd1), (c2, sd2) in zip(single, batch):
        for k in ("elem", "face", "rows", "cols", "idxcount", "idxsum"):
            assert np.array_equal(np.asarray(c1[k]), np.asarray(c2[k])), k
        for k in ("evol", "area", "nvol", "deldotdel"):
            a, b = np.asarray(c1[k]), np.asarray(c2[k])
            assert a.shape == b.shape and np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(a)), k
        assert c1["reff"] == c2["reff"] if not isinstance(c1["reff"], dict) else c1["reff"] == c2["reff"]
        if isinstance(sd1, dict):
            assert sd1.keys() == sd2.keys() and all(np.array_equal(sd1[w], sd2[w]) for w in sd1)
        else:
            assert np.array_equal(sd1, sd2)
        if "widesrc" in c1:
            assert np.max(np.abs(c1["widesrc"] - c2["widesrc"])) <= 1e-12 * np.max(np.abs(c1["widesrc"]))
    # the flipped tets were reoriented: every signed volume positive
    c = batch[2][0]
    p = c["node"][c["elem"][:, :4] - 1]
    det = np.einsum("ij,ij->i", np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), p[:, 3] - p[:, 0])
    assert np.all(det > 0)


def test_meshprep_batch_then_forward_batch():
    """The batch path end to end: GPU mesh preparation, then one block-diagonal solve."""
    import redbirdpy_amd as rb
    from oracle import rb_oracle as orc
    from conftest import box_cfg

    cfgs = [box_cfg((12, 10, 6), 2.0), box_cfg((20, 16, 8), 2.0)]
    prepped = [c for c, _ in rb.meshprep_batch([copy.deepcopy(c) for c in cfgs])]
    res = rb.forward_batch(prepped)
    for c, r in zip(cfgs, res):
        oc, _ = orc.meshprep(copy.deepcopy(c))
        det, phi = orc.runforward(oc)
        assert np.max(np.abs(r.detval - det)) <= 1e-9 * np.max(np.abs(det))
