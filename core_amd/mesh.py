"""Host arrays of one tetrahedral mesh part, and the box generator.

`BoxMesh` numbers entities as apf::makeMdsBox followed by a Lagrange-2 shape
change does (mds/apfBox.cc:178-336, apf/apfMesh.cc:736-762): vertices in grid
order, axis edges, face diagonals, body diagonals, six Kuhn tets per hex.  The
arrays are built by sprg_box_build in libsprgpu.so (host code).
"""
import numpy as np

from . import _lib

# apf::tet_edge_verts (apf/apfMesh.cc:53-60)
TET_EDGE_VERTS = np.array([[0, 1], [1, 2], [2, 0], [0, 3], [1, 3], [2, 3]], dtype=np.int64)


def _ptr(a):
    return None if a is None else a.ctypes.data


class HostMesh:
    """tv [nt,4] int32, te [nt,6] int32 (order 2), ev [ne,2] int32,
    xyz [nodes,3] float64 (vertex nodes, then one node per edge for order 2)."""

    def __init__(self, tv, xyz, order, te=None, ev=None, ne=None):
        self.tv = np.ascontiguousarray(tv, dtype=np.int32)
        self.te = None if te is None else np.ascontiguousarray(te, dtype=np.int32)
        self.ev = None if ev is None else np.ascontiguousarray(ev, dtype=np.int32)
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        self.order = int(order)
        self.nt = self.tv.shape[0]
        if ne is None:
            if self.ev is not None:
                ne = self.ev.shape[0]
            elif self.te is not None and self.te.size:
                ne = int(self.te.max()) + 1
            else:
                ne = 0
        self.ne = int(ne)
        self.nv = self.xyz.shape[0] - (self.ne if self.order == 2 else 0)

    @property
    def nnodes(self):
        return self.nv + (self.ne if self.order == 2 else 0)

    @property
    def npatches(self):
        """Recovered entities: vertices, plus edges for a Lagrange-2 mesh."""
        return self.nnodes

    def ip_points(self, ip_order):
        """Global positions [nt,q,3] of the tet rule points of ip_order under
        the affine map of the tet's vertices (equal to the Lagrange-2 map of a
        straight-sided mesh up to rounding)."""
        xi, _ = tet_rule(ip_order)
        lam = np.stack([1 - xi.sum(1), xi[:, 0], xi[:, 1], xi[:, 2]], 1)
        P = np.zeros((self.nt, lam.shape[0], 3))
        for a in range(4):
            P += lam[None, :, a, None] * self.xyz[self.tv[:, a]][:, None, :]
        return P


def tet_rule(order):
    """Tet integration rules N1..N3 of apf/apfIntegrate.cc:316-361."""
    a, b = 0.138196601125011, 0.585410196624969
    if order <= 1:
        return np.array([[0.25, 0.25, 0.25]]), np.array([1.0 / 6.0])
    if order == 2:
        xi = np.array([[a, a, a], [b, a, a], [a, b, a], [a, a, b]])
        return xi, np.full(4, 0.25 / 6.0)
    if order == 3:
        s, h = 1.0 / 6.0, 1.0 / 2.0
        xi = np.array([[0.25, 0.25, 0.25], [s, s, s], [s, s, h], [s, h, s], [h, s, s]])
        return xi, np.array([-2.0 / 15.0] + [3.0 / 40.0] * 4)
    raise ValueError("integration order %d is not supported" % order)


SMB_TYPES = ("vertex", "edge", "triangle", "quad", "hex", "prism", "pyramid", "tet")


def smb_counts(path):
    """Entity counts of an MDS .smb part file by SMB type, and its version."""
    counts = np.zeros(8, np.int64)
    version = np.zeros(1, np.int32)
    _lib.check(_lib.lib().sprg_smb_counts(str(path).encode(), _ptr(counts), _ptr(version)))
    return dict(zip(SMB_TYPES, counts.tolist())), int(version[0])


def load_smb(path, order=1):
    """One part of tetrahedra from an MDS .smb file (apf::loadMdsPart) in
    the file's entity order.  order 2: Lagrange-2 coordinates with edge
    nodes at the edge midpoints (apf::changeShape's projection)."""
    counts, _ = smb_counts(path)
    nv, ne, nt = counts["vertex"], counts["edge"], counts["tet"]
    tv = np.empty((nt, 4), np.int32)
    te = np.empty((nt, 6), np.int32)
    ev = np.empty((ne, 2), np.int32)
    xyz = np.empty((nv, 3), np.float64)
    _lib.check(_lib.lib().sprg_smb_read(str(path).encode(), _ptr(tv), _ptr(te), _ptr(ev), _ptr(xyz)))
    if order == 2:
        mid = np.zeros((ne, 3))
        mid += xyz[ev[:, 0]] * 0.5
        mid += xyz[ev[:, 1]] * 0.5
        xyz = np.vstack([xyz, mid])
    return HostMesh(tv, xyz, order, te=te, ev=ev, ne=ne)


def box_counts(n):
    nv = np.zeros(1, np.int64)
    nt = np.zeros(1, np.int64)
    ne = np.zeros(1, np.int64)
    _lib.check(_lib.lib().sprg_box_counts(n, _ptr(nv), _ptr(nt), _ptr(ne)))
    return int(nv[0]), int(nt[0]), int(ne[0])


def box(n, order=1, jitter=0.0, seed=0, grade=False):
    """Unit-cube Kuhn box of n^3 hexes x 6 tets.  jitter: interior-vertex
    displacement bound in lattice units (seeded splitmix64), grade: x -> x^2."""
    nv, nt, ne = box_counts(n)
    tv = np.empty((nt, 4), np.int32)
    te = np.empty((nt, 6), np.int32)
    ev = np.empty((ne, 2), np.int32)
    xyz = np.empty((nv + (ne if order == 2 else 0), 3), np.float64)
    _lib.check(_lib.lib().sprg_box_build(n, order, float(jitter), int(seed) & (2**64 - 1),
                                         1 if grade else 0, _ptr(tv), _ptr(te), _ptr(ev),
                                         _ptr(xyz)))
    return HostMesh(tv, xyz, order, te=te, ev=ev, ne=ne)
