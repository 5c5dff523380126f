"""The MDS .smb reader (sprg_smb_counts / sprg_smb_read), CPU only.

Pinned by:
  - the reference's own part files (tests/golden/cube0.smb, sphere0.smb,
    pipe0.smb, copied from its python_wrappers/input and smoke_test_meshes):
    header counts, and an independent reader (tests/smb_util.py) that derives
    each tet's edges and vertices from its triangles as set intersections;
  - geometry: every tet of the reference's parts comes out positively
    oriented, and each tet edge joins the two vertices apf::tet_edge_verts
    names (a wrong derivation table permutes vertices and flips tets);
  - round trips: box meshes (mds/apfBox.cc numbering) and Delaunay meshes
    written in the reference's layout come back with the same vertex order
    and edge numbering;
  - parts with other element types (pipe0: prisms and quads besides tets)
    are refused with a message that names them.
"""
import os

import numpy as np
import pytest

from core_amd import _lib, mesh
import smb_util

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def golden(name):
    return os.path.join(GOLDEN, name)


@pytest.mark.parametrize("name,version,counts", [
    ("cube0.smb", 0, [9, 25, 28, 0, 0, 0, 0, 11]),
    ("sphere0.smb", 5, [6, 13, 12, 0, 0, 0, 0, 4]),
    ("pipe0.smb", 3, [784, 3332, 2842, 1596, 0, 1022, 0, 867]),
])
def test_counts(name, version, counts):
    got, v = mesh.smb_counts(golden(name))
    assert v == version
    assert list(got.values()) == counts
    n, _, _, _, _ = smb_util.read_smb(golden(name)) if name != "pipe0.smb" else (counts, 0, 0, 0, 0)
    assert list(n) == counts


@pytest.mark.parametrize("name", ["cube0.smb", "sphere0.smb"])
def test_reader_matches_independent_restatement(name):
    h = mesh.load_smb(golden(name))
    n, tv, te, ev, xyz = smb_util.read_smb(golden(name))
    assert np.array_equal(h.tv, tv)
    assert np.array_equal(h.te, te)
    assert np.array_equal(h.ev, ev)
    assert np.array_equal(h.xyz, xyz)


@pytest.mark.parametrize("name", ["cube0.smb", "sphere0.smb"])
def test_reference_parts_are_consistent(name):
    h = mesh.load_smb(golden(name))
    X = h.xyz[h.tv]
    assert np.all(np.linalg.det(X[:, 1:] - X[:, :1]) > 0)
    for k, (a, b) in enumerate(mesh.TET_EDGE_VERTS):
        assert np.all(np.sort(h.ev[h.te[:, k]], 1) == np.sort(h.tv[:, [a, b]], 1))
    # every vertex and edge is used, every tet's vertices are distinct
    assert len(np.unique(h.tv)) == h.nv and len(np.unique(h.te)) == h.ne
    assert all(len(set(r)) == 4 for r in h.tv.tolist())


def test_mixed_part_is_refused():
    with pytest.raises(_lib.SPRError) as e:
        mesh.load_smb(golden("pipe0.smb"))
    assert e.value.code == _lib.SPRG_EUNSUPPORTED
    assert "1596 quad" in str(e.value) and "1022 prism" in str(e.value)


def test_missing_or_truncated_file(tmp_path):
    with pytest.raises(_lib.SPRError):
        mesh.smb_counts(str(tmp_path / "none0.smb"))
    data = open(golden("cube0.smb"), "rb").read()
    p = tmp_path / "cut0.smb"
    p.write_bytes(data[:200])
    with pytest.raises(_lib.SPRError) as e:
        mesh.load_smb(str(p))
    assert "ends early" in str(e.value)


@pytest.mark.parametrize("order", [1, 2])
def test_box_round_trip(tmp_path, order):
    b = mesh.box(3, order=1, jitter=0.2, seed=7)
    p = str(tmp_path / "box0.smb")
    smb_util.write_smb(p, b.tv, b.xyz, edge_order=b.ev.tolist())
    h = mesh.load_smb(p, order=order)
    assert np.array_equal(h.tv, b.tv)
    assert np.array_equal(h.te, b.te)
    assert np.array_equal(h.xyz[: b.nv], b.xyz)
    if order == 2:  # midpoints as the box generator's Lagrange-2 nodes
        b2 = mesh.box(3, order=2, jitter=0.2, seed=7)
        assert np.array_equal(h.xyz, b2.xyz)


def test_delaunay_round_trip(tmp_path):
    tv, xyz = smb_util.delaunay(300, 3)
    p = str(tmp_path / "del0.smb")
    ev = smb_util.write_smb(p, tv, xyz)
    h = mesh.load_smb(p)
    assert np.array_equal(h.tv, tv)
    assert np.array_equal(h.ev, ev)
    _, tv2, te2, _, _ = smb_util.read_smb(p)
    assert np.array_equal(h.te, te2) and np.array_equal(h.tv, tv2)
