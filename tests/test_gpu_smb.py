"""GPU parity on unstructured meshes read from MDS .smb part files.

  cube0, sphere0   the reference's own tet parts (tests/golden), coordinate
                   order 1 and 2 (edge nodes at the midpoints)
  delaunay         a seeded Delaunay tetrahedralisation of 3,000 points in
                   the unit cube, written as an .smb part: vertex valences up
                   to ~100 tets, so vertex patches above 192 sample rows at 4
                   points per tet take the large-set exact kernel

Against the FP64 oracle on the same arrays: adjacency, patch sets and the
FP64 form's recovered field bit for bit; size fields to 1e-13 (the FP64
form) and 1e-6 (the fast forms).  Where the reference would abort ("all
hope is lost" on a part too small to fit), the library must report
SPRG_EPATCH.
"""
import os

import numpy as np
import pytest

from core_amd import _lib, mesh, spr, workloads
from oracle import oracle as O
import smb_util

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _omesh(h):
    return O.Mesh(h.tv, h.xyz, h.order, te=h.te, ev=h.ev)


@pytest.fixture(scope="module")
def delaunay_path(tmp_path_factory):
    tv, xyz = smb_util.delaunay(3000, 11)
    p = str(tmp_path_factory.mktemp("smb") / "delaunay0.smb")
    smb_util.write_smb(p, tv, xyz)
    return p


def _path(name, delaunay_path):
    return delaunay_path if name == "delaunay" else os.path.join(HERE, "golden", name + "0.smb")


CASES = [("cube", 1), ("cube", 2), ("sphere", 1), ("sphere", 2), ("delaunay", 1), ("delaunay", 2)]


@pytest.fixture(scope="module", params=CASES, ids=["%s-p%d" % c for c in CASES])
def case(request, gpu, delaunay_path):
    name, order = request.param
    m = spr.Mesh.load_smb(_path(name, delaunay_path), order)
    h = m.host
    val = np.bincount(h.tv.ravel(), minlength=h.nv)
    print("%s p%d: %d tets, %d vertices, %d edges; vertex valence max %d, %d above 48"
          % (name, order, h.nt, h.nv, h.ne, val.max(), (val > 48).sum()))
    return name, h, m, _omesh(h)


def test_smb_adjacency_and_patch_sets(case):
    name, h, m, om = case
    off, idx = m.adjacency(0)
    ooff, oidx = O.vertex_tets(om)
    assert np.array_equal(off, ooff) and np.array_equal(idx, oidx)
    for q, ipo in ((1, 1), (4, 2)):
        try:
            ooff, oidx = O.patch_sets(om, q)
        except RuntimeError:
            with pytest.raises(_lib.SPRError) as e:
                m.patch_sets(ipo)
            assert e.value.code == _lib.SPRG_EPATCH
            continue
        off, idx = m.patch_sets(ipo)
        assert np.array_equal(off, ooff.astype(np.int64)) and np.array_equal(idx, oidx), q


@pytest.mark.parametrize("ip_order", [1, 2])
def test_smb_recover_and_size(case, ip_order):
    name, h, m, om = case
    ip = workloads.strain_field(h.ip_points(ip_order))
    f = m.createIPField("eps%d" % ip_order, 9, ip_order, ip)
    try:
        ref = O.recover(om, ip)
    except RuntimeError:  # the reference aborts: all hope is lost
        with pytest.raises(_lib.SPRError) as e:
            spr.recoverField(f)
        assert e.value.code == _lib.SPRG_EPATCH
        m.destroyField(f)
        return
    r = spr.recoverField(f)
    stats = m.last_stats()
    print("%s p%d ip%d routes %s" % (name, h.order, ip_order, stats))
    assert np.array_equal(r.numpy(), ref)
    m.destroyField(r)
    oref, _ = O.spr_size(om, ip_order, ip, 0.1)
    for form, tol in (("fp64", 1e-13), ("fp64f", 1e-6), ("fp32r", 1e-6)):
        s = spr.getSPRSizeField(f, 0.1, form=form)
        err = np.max(np.abs(s.numpy()[:, 0] / oref - 1))
        m.destroyField(s)
        print("%s p%d ip%d %s size max rel %.3e" % (name, h.order, ip_order, form, err))
        assert err <= tol, form
    m.destroyField(f)


def test_delaunay_large_sets_take_the_large_kernel(gpu, delaunay_path):
    m = spr.Mesh.load_smb(delaunay_path, 1)
    h = m.host
    assert np.bincount(h.tv.ravel()).max() * 4 > 192  # vertex patches above 192 rows
    f = m.createIPField("eps", 9, 2, workloads.strain_field(h.ip_points(2)))
    r = spr.recoverField(f)
    assert m.last_stats()["large_sets"] > 0
    m.destroyField(r)
    m.close()


@pytest.mark.parametrize("form", ["fp64", "fp32r", "fp64f"])
def test_vertex_beyond_a_tile_union_is_fitted_outside_the_tiles(gpu, form):
    # the centre of star(70) has about 136 tets, more than a tile's union
    # holds (128): its patch takes the thread-per-patch pass in the fast
    # forms; at one point per tet the exact form fits it as well
    tv, xyz = smb_util.star(70)
    h = mesh.HostMesh(tv, xyz, 1)
    assert np.bincount(h.tv.ravel()).max() > 128
    m = spr.Mesh(h)
    ip = workloads.strain_field(h.ip_points(1))
    f = m.createIPField("eps", 9, 1, ip)
    got = spr.recoverField(f, form=form).numpy()
    ref = O.recover(_omesh(h), ip)
    if form == "fp64":
        assert np.array_equal(got, ref)
    else:
        scale = np.abs(ref).max(0)
        keep = scale > 0
        assert np.max(np.abs(got - ref)[:, keep] / scale[keep]) <= 1e-9
    m.close()
