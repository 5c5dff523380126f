"""Test helpers for MDS .smb part files (test infrastructure).

write_smb    writes a part of tetrahedra in the layout of mds/mds_smb.c:617-637
             (big-endian, version 6, no tags, linear coordinate field), each
             tet stored as the 4 triangles from which MDS derives the given
             vertex order (faces (0,1,2), (0,1,3), (1,2,3), (0,2,3); mds/mds.c
             tables T12/T21/T10)
read_smb     an independent reader: parses the header and connectivity with
             numpy and derives each tet's edges and vertices as
             intersections of sets
delaunay     a seeded Delaunay tetrahedralisation of random points in the
             unit cube (scipy), positively oriented: an unstructured mesh
             with vertex valences up to about 100 tets
"""
import struct

import numpy as np

TET_EDGE_VERTS = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]
# the triangles of a tet (vertex positions), in the order MDS stores them
TET_TRI_VERTS = [(0, 1, 2), (0, 1, 3), (1, 2, 3), (0, 2, 3)]
# a triangle's edges as tet edge numbers
TET_TRI_EDGES = [(0, 1, 2), (0, 4, 3), (1, 5, 4), (2, 3, 5)]


def _u32(a):
    return np.asarray(a, dtype=">u4").tobytes()


def _f64(a):
    return np.asarray(a, dtype=">f8").tobytes()


def write_smb(path, tv, xyz, edge_order=None):
    """tv [nt, 4] int, xyz [nv, 3].  edge_order: optional list of vertex
    pairs fixing the edge numbering (edges found in the tets otherwise get
    numbers in order of first appearance)."""
    tv = np.asarray(tv, dtype=np.int64)
    nt = tv.shape[0]
    nv = xyz.shape[0]
    edges = {}
    if edge_order is not None:
        for a, b in edge_order:
            edges[(min(a, b), max(a, b))] = len(edges)
    ev = []
    for t in range(nt):
        for a, b in TET_EDGE_VERTS:
            key = (min(tv[t, a], tv[t, b]), max(tv[t, a], tv[t, b]))
            if key not in edges:
                edges[key] = len(edges)
    ev = np.zeros((len(edges), 2), np.int64)
    for (a, b), i in edges.items():
        ev[i] = (a, b)
    tris = {}
    tri_edges = []
    tet_tris = np.zeros((nt, 4), np.int64)
    for t in range(nt):
        te = [edges[(min(tv[t, a], tv[t, b]), max(tv[t, a], tv[t, b]))] for a, b in TET_EDGE_VERTS]
        for f, (verts, fe) in enumerate(zip(TET_TRI_VERTS, TET_TRI_EDGES)):
            key = tuple(sorted(tv[t, list(verts)]))
            if key not in tris:
                tris[key] = len(tris)
                tri_edges.append([te[k] for k in fe])
            tet_tris[t, f] = tris[key]
    n = [nv, len(edges), len(tris), 0, 0, 0, 0, nt]
    out = bytearray()
    out += _u32([0, 6, 3, 1])  # magic, version, dimension, parts
    out += _u32(n)
    out += _u32(ev.ravel())
    out += _u32(np.asarray(tri_edges).ravel())
    out += _u32(tet_tris.ravel())
    out += _f64(np.asarray(xyz, dtype=np.float64).ravel())
    out += _f64(np.zeros(2 * nv))  # parametric coordinates
    out += _u32([0])  # remote copies: no peers
    for k in n:  # classification: model entity (id, dim) per entity
        out += _u32(np.zeros(2 * k))
    out += _u32([0])  # no tags; then, per type, the (empty) tag sizes
    for _ in range(8):  # matches: no peers
        out += _u32([0])
    out += b"Linear\0" + struct.pack(">i", -3) + struct.pack(">i", 0) + struct.pack(">i", 0)
    with open(path, "wb") as fh:
        fh.write(bytes(out))
    return ev


def read_smb(path):
    """Counts, tv, te, ev, xyz of a tet part, parsed and derived here."""
    data = open(path, "rb").read()
    head = np.frombuffer(data, dtype=">u4", count=12)
    n = [int(x) for x in head[4:12]]
    off = 48
    down = [0, 2, 3, 4, 6, 5, 5, 4]
    conn = []
    for t in range(1, 8):
        k = down[t] * n[t]
        conn.append(np.frombuffer(data, dtype=">u4", count=k, offset=off).astype(np.int64).reshape(n[t], down[t])
                    if k else np.zeros((0, down[t]), np.int64))
        off += 4 * k
    xyz = np.frombuffer(data, dtype=">f8", count=3 * n[0], offset=off).astype(np.float64).reshape(n[0], 3)
    ev, fe, tf = conn[0], conn[1], conn[6]
    nt = n[7]
    tv = np.zeros((nt, 4), np.int64)
    te = np.zeros((nt, 6), np.int64)
    # tet edge i: the edge the tet's triangles (0,1), (0,2), (0,3), (1,3),
    # (1,2), (2,3) share; vertex i: the vertex edges (2,0), (0,1), (1,2),
    # (3,4) share
    pairs_f = [(0, 1), (0, 2), (0, 3), (1, 3), (1, 2), (2, 3)]
    pairs_e = [(2, 0), (0, 1), (1, 2), (3, 4)]
    for t in range(nt):
        fs = [set(fe[f].tolist()) for f in tf[t]]
        for i, (a, b) in enumerate(pairs_f):
            (te[t, i],) = fs[a] & fs[b]
        es = [set(ev[e].tolist()) for e in te[t]]
        for i, (a, b) in enumerate(pairs_e):
            (tv[t, i],) = es[a] & es[b]
    return n, tv, te, ev, xyz


def delaunay(npts, seed):
    """tv [nt, 4] (positive volume), xyz [nv, 3]: the Delaunay
    tetrahedralisation of the unit cube's corners and npts seeded random
    points in it."""
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)
    corners = np.array([[i, j, k] for i in (0, 1) for j in (0, 1) for k in (0, 1)], float)
    P = np.vstack([corners, rng.random((npts, 3))])
    T = Delaunay(P).simplices.astype(np.int64)
    X = P[T]
    det = np.linalg.det(X[:, 1:] - X[:, :1])
    flip = det < 0
    T[flip] = T[flip][:, [1, 0, 2, 3]]
    return T, P


def star(n):
    """tv [nt, 4], xyz [n+1, 3]: the fan of tets joining the centre of the
    unit sphere to the hull triangles of n Fibonacci points on it, so the
    centre's vertex patch holds every tet (about 2n)."""
    from scipy.spatial import ConvexHull

    i = np.arange(n) + 0.5
    phi = np.arccos(1 - 2 * i / n)
    th = np.pi * (1 + 5 ** 0.5) * i
    P = np.stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi), np.cos(phi)], 1)
    tri = ConvexHull(P).simplices
    T = np.hstack([tri, np.full((len(tri), 1), n)]).astype(np.int64)
    X = np.vstack([P, [[0.0, 0.0, 0.0]]])
    det = np.linalg.det(X[T][:, 1:] - X[T][:, :1])
    flip = det < 0
    T[flip] = T[flip][:, [1, 0, 2, 3]]
    return T, X
