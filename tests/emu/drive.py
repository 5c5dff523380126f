"""Runs the CPU emulation of the library (tests/emu/build.py) on tiny P1
meshes and checks it against the oracle: adjacency, patch sets, gradients
and recovered fields bit for bit, size fields to 1e-13, the fast forms to
1e-6 on the size field, and a two-mesh batch (recorded graph, replayed)
equal to the per-mesh calls.  (Each block's lanes are OS threads meeting at
condition-variable barriers, so P2 meshes take many minutes here; the GPU
tests cover them.)  Started by tests/test_emu_cpu.py in a child
process with AddressSanitizer preloaded; prints one line per check and exits
non-zero on the first failure."""
import ctypes
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from core_amd import _lib  # noqa: E402

_lib.LIB_PATH = os.path.join(HERE, "_build", "libsprgpu_emu.so")
from core_amd import mesh, workloads  # noqa: E402
from oracle import oracle as O  # noqa: E402

sys.path.insert(0, os.path.dirname(HERE))
import smb_util  # noqa: E402

L = _lib.lib()
vp = ctypes.c_void_p


def P(a):
    return vp(a.ctypes.data)


def check(name, ok):
    print(("PASS " if ok else "FAIL ") + name, flush=True)
    if not ok:
        sys.exit(1)


# persistent grids of a few blocks: the emulator walks every block of a
# grid, so the default grids (thousands of blocks) only cost time on these
# meshes, and small grids give every block several work items
SMALL_SHAPE = (ctypes.c_int64 * 5)(64, 24, 16, 8, 12)


def device_mesh(ctx, h):
    m = vp()
    _lib.check(L.sprg_mesh_create(ctx, h.nv, h.nt, h.ne if h.te is not None else 0, P(h.tv),
                                  None if h.te is None else P(h.te), P(h.xyz), h.order,
                                  ctypes.byref(m)))
    _lib.check(L.sprg_mesh_set_shape(m, SMALL_SHAPE, 5))
    return m


def exact_sequence(ctx, name, h, shapes=True, fast=True):
    om = O.Mesh(h.tv, h.xyz, h.order, te=h.te, ev=h.ev)
    m = device_mesh(ctx, h)
    nodes = h.nv + (h.ne if h.order == 2 else 0)
    for dim in ([0, 1] if h.order == 2 else [0]):
        off, idx = vp(), vp()
        _lib.check(L.sprg_mesh_adjacency(m, dim, ctypes.byref(off), ctypes.byref(idx)))
        n = h.nv if dim == 0 else h.ne
        o = np.frombuffer((ctypes.c_int32 * (n + 1)).from_address(off.value), np.int32).copy()
        i = np.frombuffer((ctypes.c_int32 * int(o[-1])).from_address(idx.value), np.int32).copy()
        ro, ri = O.vertex_tets(om) if dim == 0 else O.edge_tets(om)
        check("%s adjacency %d" % (name, dim), np.array_equal(o, ro) and np.array_equal(i, ri))
    for ipo, q in ((1, 1), (2, 4)):
        off = np.zeros(nodes + 1, np.int64)
        total = np.zeros(1, np.int64)
        _lib.check(L.sprg_patch_sets(m, ipo, P(off), None, 0, P(total)))
        idx = np.zeros(max(int(total[0]), 1), np.int32)
        _lib.check(L.sprg_patch_sets(m, ipo, P(off), P(idx), idx.size, P(total)))
        ro, ri = O.patch_sets(om, q)
        check("%s patch sets q=%d" % (name, q),
              np.array_equal(off, ro.astype(np.int64)) and np.array_equal(idx[:total[0]], ri))
    Pn = workloads.node_points(h)
    u = np.ascontiguousarray(np.stack([workloads.sin_field(Pn), workloads.exp_field(Pn),
                                       Pn[:, 0] * Pn[:, 1]], 1))
    g = np.full((h.nt, 4, 9), np.nan)
    _lib.check(L.sprg_grad_ip(m, h.order, 3, P(u), 2, P(g)))
    check("%s gradient" % name, np.array_equal(g, O.grad_ip(om, h.order, u, 2)))
    ip = np.ascontiguousarray(workloads.strain_field(h.ip_points(2)))
    out = np.full((nodes, 9), np.nan)
    _lib.check(L.sprg_recover(m, 2, 9, P(ip), P(out), 0))
    check("%s recovered field" % name, np.array_equal(out, O.recover(om, ip)))
    ref, _ = O.spr_size(om, 2, ip, 0.05)
    sizes = {}
    for form, tol in ((0, 1e-13), (1, 1e-6), (2, 1e-6)) if fast else ((0, 1e-13),):
        size = np.full(h.nv, np.nan)
        diag = np.zeros(4)
        _lib.check(L.sprg_spr_size(m, 2, 9, P(ip), 0.05, P(size), form, P(diag)))
        check("%s size field form %d" % (name, form), np.max(np.abs(size / ref - 1)) <= tol)
        sizes[form] = size
    tref, _ = O.target_size(om, 2, g, 2 * h.nt, 0.25, 2.0)
    for form, tol in ((0, 1e-13), (2, 1e-6)) if fast else ((0, 1e-13),):
        size = np.full(h.nv, np.nan)
        diag = np.zeros(4)
        _lib.check(L.sprg_target_spr_size(m, 2, 9, P(g), 2 * h.nt, 0.25, 2.0, P(size), form, P(diag)))
        check("%s target size field form %d" % (name, form), np.max(np.abs(size / tref - 1)) <= tol)
    # other launch shapes (small odd grids: every persistent loop wraps)
    # change no result
    for form, shape in ((1, [64, 3, 5, 7, 2]), (0, [128, 2, 3, 1, 3])) if shapes else ():
        sh = np.array(shape, np.int64)
        _lib.check(L.sprg_mesh_set_shape(m, P(sh), 5))
        size = np.full(h.nv, np.nan)
        diag = np.zeros(4)
        _lib.check(L.sprg_spr_size(m, 2, 9, P(ip), 0.05, P(size), form, P(diag)))
        check("%s form %d with launch shape %s" % (name, form, shape), np.array_equal(size, sizes[form]))
    L.sprg_mesh_destroy(m)


def leftover_check(ctx):
    # a vertex with more tets than a tile union holds: its patch is fitted
    # outside the tiles, by the thread-per-patch pass (one point per tet, so
    # the exact form takes it too)
    tv, xyz = smb_util.star(70)
    h = mesh.HostMesh(tv, xyz, 1)
    check("star(70) has a vertex of more than 128 tets", np.bincount(h.tv.ravel()).max() > 128)
    om = O.Mesh(h.tv, h.xyz, 1)
    m = device_mesh(ctx, h)
    ip = np.ascontiguousarray(workloads.strain_field(h.ip_points(1)))
    ref = O.recover(om, ip)
    scale = np.abs(ref).max(0)
    keep = scale > 0
    for form in (0, 1, 2):
        out = np.full((h.nv, 9), np.nan)
        _lib.check(L.sprg_recover(m, 1, 9, P(ip), P(out), form))
        if form == 0:
            ok = np.array_equal(out, ref)
        else:
            ok = np.max(np.abs(out - ref)[:, keep] / scale[keep]) <= 1e-9
        check("leftover vertex recovered, form %d" % form, ok)
    L.sprg_mesh_destroy(m)


def batch_check(ctx):
    # two P1 meshes in one recorded graph, first with the target estimator,
    # then with the adapt-ratio estimator; between replays, eager calls
    # that grow a mesh's workspaces (27 components) must make the batch
    # record again instead of replaying kernels bound to freed buffers, and
    # so must a new launch shape (the first mesh's regular pass switched
    # between tiles and thread per patch: another summation order)
    hs, ms, eps = [], [], []
    for i, n in enumerate([2, 3]):
        h = mesh.box(n, order=1, jitter=0.2, seed=workloads.C4 + i)
        m = device_mesh(ctx, h)
        u = np.ascontiguousarray(workloads.sin_field(h.xyz).reshape(-1, 1))
        g = np.full((h.nt, 4, 3), np.nan)
        _lib.check(L.sprg_grad_ip(m, 1, 1, P(u), 2, P(g)))
        hs.append(h), ms.append(m), eps.append(g)

    def eager(targets):
        out = []
        for m, h, g, t in zip(ms, hs, eps, targets if targets is not None else [0, 0]):
            s = np.full(h.nv, np.nan)
            diag = np.zeros(4)
            if targets is not None:
                _lib.check(L.sprg_target_spr_size(m, 2, 3, P(g), t, 0.25, 2.0, P(s), 2, P(diag)))
            else:
                _lib.check(L.sprg_spr_size(m, 2, 3, P(g), 0.1, P(s), 2, P(diag)))
            out.append(s)
        return out

    def grow_workspaces():
        for m, h in zip(ms, hs):
            ip27 = np.ascontiguousarray(np.tile(workloads.strain_field(h.ip_points(2)), (1, 1, 3)))
            s = np.full(h.nv, np.nan)
            diag = np.zeros(4)
            _lib.check(L.sprg_spr_size(m, 2, 27, P(ip27), 0.1, P(s), 2, P(diag)))

    arr = (vp * 2)(*[m.value for m in ms])
    b = vp()
    _lib.check(L.sprg_batch_create(ctx, 2, arr, ctypes.byref(b)))
    outs = [np.full(h.nv, -1.0) for h in hs]
    ips = (vp * 2)(*[e.ctypes.data for e in eps])
    ops = (vp * 2)(*[o.ctypes.data for o in outs])
    t = np.zeros(1)
    for targets in ([2 * h.nt for h in hs], None):
        ref = eager(targets)
        tga = None if targets is None else np.array(targets, np.uint64)
        for k in range(4):
            if k == 2:
                grow_workspaces()
            if k == 3:
                sh = np.array([256, 8192, 8192, 4096, 0 if targets is not None else 8192], np.int64)
                _lib.check(L.sprg_mesh_set_shape(ms[0], P(sh), 5))
                ref = eager(targets)
            for o in outs:
                o[:] = -1.0
            _lib.check(L.sprg_batch_size(b, 2, 3, ips, None if tga is None else P(tga), 0.1, 0.25, 2.0,
                                         ops, 2, P(t)))
            check("batch (%s) run %d equals the per-mesh calls" % ("target" if targets else "ratio", k),
                  all(np.array_equal(o, e) for o, e in zip(outs, ref)))
    L.sprg_batch_destroy(b)
    for m in ms:
        L.sprg_mesh_destroy(m)


def main():
    ctx = vp()
    _lib.check(L.sprg_ctx_create(0, ctypes.byref(ctx)))
    t0 = time.time()
    only = {"batch": batch_check, "leftover": leftover_check}
    if len(sys.argv) == 2 and sys.argv[1] in only:  # one check alone
        only[sys.argv[1]](ctx)
        print("ALL PASS (%.0f s)" % (time.time() - t0), flush=True)
        return

    def timed(label, fn, *args, **kw):
        t = time.time()
        fn(*args, **kw)
        print("TIME %s %.1f s" % (label, time.time() - t), flush=True)

    timed("p1 n=2", exact_sequence, ctx, "p1 n=2 jittered", mesh.box(2, order=1, jitter=0.25, seed=11))
    timed("p2 n=2", exact_sequence, ctx, "p2 n=2 jittered", mesh.box(2, order=2, jitter=0.25, seed=12),
          shapes=False)
    # one vertex holding 116 tets: 464 sample rows at 4 points per tet,
    # fitted by the large-set kernel
    tv, xyz = smb_util.star(60)
    timed("star", exact_sequence, ctx, "star", mesh.HostMesh(tv, xyz, 1), shapes=False)
    # the unjittered P1 box at 4 points per tet (every tet's samples collapse
    # to its centroid: most patches take the rank test and the expansion),
    # in every form: the sequence of the GPU tests held after an unexplained
    # fault
    timed("p1 n=4", exact_sequence, ctx, "p1 n=4", mesh.box(4, order=1), shapes=False)
    timed("leftover", leftover_check, ctx)
    timed("batch", batch_check, ctx)
    print("ALL PASS (%.0f s)" % (time.time() - t0), flush=True)


if __name__ == "__main__":
    main()
