"""The FP64 oracle's getSPRSizeField run over a partitioned mesh: one forked
process per core, each owning one RCB part with a ghost halo.

TEST INFRASTRUCTURE ONLY: the CPU baseline of the benchmark (and its tests).
The numerical work is the oracle's (spr_oracle.c, which cites the reference
lines it restates); this file only partitions the mesh and runs the step.

Parts (mesh setup, not timed):
  - recursive coordinate bisection of the vertices: each cut is at the
    median of the longest extent, with vertex counts in proportion to the
    part counts on either side;
  - a vertex belongs to its part; an edge and a tet belong to the part of
    their smallest vertex id;
  - each process builds its local mesh from the tets within `layers` vertex
    layers of its owned vertices (layer 1: every tet touching an owned
    vertex, i.e. every owned patch; layer 2: every tet touching a vertex of
    layer 1, i.e. one round of patch expansion), numbered in ascending global
    order so that every patch lists its tets in the serial order and the
    recovered values are the serial run's bit for bit.

The timed step (spr/sprEstimateError.cc:301-321), per process:
  A  recover the owned vertex and edge patches on the local mesh and publish
     the owned values in shared memory (the counterpart of synchronising the
     recovered field to its ghost copies);
  B  read the ghost values, then the self product and the element errors over
     the owned tets, and the two partial sums;
  C  all-reduce of the two sums (every process adds the parts' partials in
     part order), the size factor, and the owned tets' element sizes,
     published;
  D  the vertex average over each owned vertex's tets.
Barriers separate the phases; the step's time is the parent's wall clock
between the start and end barriers.
"""
import ctypes
import mmap
import multiprocessing
import os
import threading
import time

import numpy as np

from oracle import oracle as O

_p = ctypes.c_void_p
_i64 = ctypes.c_int64


def _lib():
    L = O.lib()
    if not getattr(L, "_partition_sigs", False):
        L.om_patcher_new.restype = _p
        L.om_patcher_new.argtypes = [_p, ctypes.c_int, ctypes.c_int]
        L.om_patcher_recover.argtypes = [_p, _p, _p, _i64, _i64]
        L.om_patcher_free.argtypes = [_p]
        L.om_self_product_range.argtypes = [_p, ctypes.c_int, ctypes.c_int, _p, _i64, _i64, _p]
        L.om_element_error_range.argtypes = [_p, ctypes.c_int, ctypes.c_int, _p, _p, _p, _i64, _i64]
        L.om_current_size_range.argtypes = [_p, _p, _i64, _i64]
        L.om_csr_mean.argtypes = [_i64, _p, _p, _p, _p]
        L._partition_sigs = True
    return L


def rcb(x, nparts):
    """Part label per point (x [n, 3]) by recursive coordinate bisection."""
    lab = np.zeros(len(x), np.int32)

    def split(idx, base, k):
        if k == 1:
            lab[idx] = base
            return
        k1 = k // 2
        pts = x[idx]
        ax = int(np.argmax(pts.max(0) - pts.min(0)))
        order = np.argsort(pts[:, ax], kind="stable")
        cut = len(idx) * k1 // k
        split(idx[order[:cut]], base, k1)
        split(idx[order[cut:]], base + k1, k - k1)

    split(np.arange(len(x)), 0, nparts)
    return lab


def _shared(shape, dtype):
    """A zeroed array in an anonymous shared mapping (seen by forked
    children)."""
    n = int(np.prod(shape)) * np.dtype(dtype).itemsize
    buf = mmap.mmap(-1, max(n, 1))
    return np.frombuffer(buf, dtype=dtype, count=int(np.prod(shape))).reshape(shape)


class Part:
    """One part's local mesh (owned entities first) and its maps to the
    global numbering."""

    def __init__(self, g, lab, k, layers):
        tv, nv = g["tv"], g["nv"]
        order = g["order"]
        own_v = lab == k
        tmask = own_v[tv].any(1)
        for _ in range(layers - 1):
            vm = np.zeros(nv, bool)
            vm[tv[tmask]] = True
            tmask = vm[tv].any(1)
        H = np.flatnonzero(tmask)
        own_t = np.flatnonzero(lab[g["tmin"]] == k)
        used = np.zeros(nv, bool)
        used[tv[H]] = True
        ov = np.flatnonzero(own_v)
        gv = np.concatenate([ov, np.flatnonzero(used & ~own_v)])
        vmap = np.full(nv, -1, np.int32)
        vmap[gv] = np.arange(len(gv), dtype=np.int32)
        self.nvo, self.nto = len(ov), len(own_t)
        tvL, tvO = vmap[tv[H]], vmap[tv[own_t]]
        if order == 2:
            te, ev = g["te"], g["ev"]
            own_e = lab[g["emin"]] == k
            eused = np.zeros(len(ev), bool)
            eused[te[H]] = True
            oe = np.flatnonzero(own_e)
            ge = np.concatenate([oe, np.flatnonzero(eused & ~own_e)])
            emap = np.full(len(ev), -1, np.int32)
            emap[ge] = np.arange(len(ge), dtype=np.int32)
            teL, teO, evL = emap[te[H]], emap[te[own_t]], vmap[ev[ge]]
            xyz = np.concatenate([g["xyz"][gv], g["xyz"][nv + ge]])
            self.neo = len(oe)
            self.nodes = np.concatenate([gv, nv + ge])  # global node of each local node
            self.own_nodes_local = np.concatenate([np.arange(self.nvo), len(gv) + np.arange(self.neo)])
        else:
            teL = teO = evL = None
            xyz = g["xyz"][gv]
            self.neo = 0
            self.nodes = gv
            self.own_nodes_local = np.arange(self.nvo)
        self.own_nodes_global = self.nodes[self.own_nodes_local]
        self.L = O.Mesh(tvL, xyz, order, te=teL, ev=evL)
        self.O = O.Mesh(tvO, self.L.xyz, order, te=teO, ev=evL)
        self.ipL = np.ascontiguousarray(g["ip"][H])
        self.ipO = np.ascontiguousarray(g["ip"][own_t])
        self.own_t = own_t
        self.ov = ov
        # each owned vertex's tets, ascending global id (the serial order)
        loc = tvL.ravel()
        sel = np.flatnonzero(loc < self.nvo)
        vert = loc[sel]
        tets = H[sel // 4].astype(np.int32)
        o = np.argsort(vert, kind="stable")
        self.avg_idx = np.ascontiguousarray(tets[o])
        self.avg_off = np.zeros(self.nvo + 1, np.int32)
        np.cumsum(np.bincount(vert, minlength=self.nvo), out=self.avg_off[1:])
        self.nlocal_tets = len(H)


def _step(part, k, sh, bar, ip_order, nc, ratio, times):
    """Phases A-D of one part (barriers between phases when bar is set);
    times[k] = each phase's own time, barrier waits excluded."""
    L = _lib()
    P = part
    d, p = 3.0, float(P.L.order)
    t = np.zeros(4)
    t0 = time.perf_counter()
    # A: owned patches, published
    recL = np.zeros((P.L.nnodes, nc))
    rc = L.om_patcher_recover(P.patcher, O._ptr(P.ipL), O._ptr(recL), 0, P.nvo)
    if rc == 0 and P.neo:
        rc = L.om_patcher_recover(P.patcher, O._ptr(P.ipL), O._ptr(recL), P.L.nv, P.L.nv + P.neo)
    O._check(rc)
    sh["rec"][P.own_nodes_global] = recL[P.own_nodes_local]
    t[0] = time.perf_counter() - t0
    if bar is not None:
        bar.wait()
    t0 = time.perf_counter()
    # B: ghost values, self product and element errors of the owned tets
    recL = np.ascontiguousarray(sh["rec"][P.nodes])
    self_k = np.zeros(1)
    O._check(L.om_self_product_range(P.O.c(), ip_order, nc, O._ptr(recL), 0, P.nto, O._ptr(self_k)))
    s = np.zeros(P.nto)
    O._check(L.om_element_error_range(P.O.c(), ip_order, nc, O._ptr(P.ipO), O._ptr(recL),
                                      O._ptr(s), 0, P.nto))
    sh["partial"][k] = (self_k[0], float(np.sum(np.power(np.sqrt(s), (2 * d) / (2 * p + d)))))
    t[1] = time.perf_counter() - t0
    if bar is not None:
        bar.wait()
    t0 = time.perf_counter()
    # C: all-reduce, size factor, element sizes
    self_all = 0.0
    errsum = 0.0
    for j in range(len(sh["partial"])):
        self_all += sh["partial"][j, 0]
        errsum += sh["partial"][j, 1]
    eps_star = np.sqrt(self_all)
    a = ratio * ratio * eps_star * eps_star
    factor = np.power(a / errsum, 1.0 / (2.0 * p))
    h = np.zeros(P.nto)
    O._check(L.om_current_size_range(P.O.c(), O._ptr(h), 0, P.nto))
    sh["esize"][P.own_t] = h * np.power(np.sqrt(s), -(2 / (2 * p + d))) * factor
    t[2] = time.perf_counter() - t0
    if bar is not None:
        bar.wait()
    t0 = time.perf_counter()
    # D: vertex average over the owned vertices
    out = np.zeros(P.nvo)
    L.om_csr_mean(P.nvo, O._ptr(P.avg_off), O._ptr(P.avg_idx), O._ptr(sh["esize"]), O._ptr(out))
    sh["size"][P.ov] = out
    t[3] = time.perf_counter() - t0
    times[k] = t


def _worker(g, lab, k, layers, sh, outer, inner, ip_order, nc, ratio, times, err):
    try:
        part = Part(g, lab, k, layers)
        part.patcher = _lib().om_patcher_new(part.L.c(), part.ipL.shape[1], nc)
        if not part.patcher:
            raise RuntimeError(O.lib().om_last_error().decode())
    except BaseException as e:  # noqa: BLE001 -- reported by the parent
        err[k] = 1
        print("partition part %d: %r" % (k, e), flush=True)
        outer.abort()
        inner.abort()
        return
    outer.wait()  # start
    try:
        _step(part, k, sh, inner, ip_order, nc, ratio, times)
    except BaseException as e:  # noqa: BLE001
        err[k] = 1
        print("partition part %d: %r" % (k, e), flush=True)
        inner.abort()
    outer.wait()  # end


def globals_of(host, ip):
    """Global arrays of a core_amd.mesh.HostMesh (int32 indices) and its IP
    field, as the parts read them."""
    tv = np.ascontiguousarray(host.tv, dtype=np.int32)
    g = {"tv": tv, "nv": int(host.nv), "nt": int(host.nt), "order": int(host.order),
         "xyz": np.ascontiguousarray(host.xyz, dtype=np.float64),
         "ip": np.ascontiguousarray(ip, dtype=np.float64), "tmin": tv.min(1)}
    if host.order == 2:
        g["te"] = np.ascontiguousarray(host.te, dtype=np.int32)
        g["ev"] = np.ascontiguousarray(host.ev, dtype=np.int32)
        g["emin"] = g["ev"].min(1)
        g["nnodes"] = g["nv"] + len(g["ev"])
    else:
        g["nnodes"] = g["nv"]
    return g


def spr_size(host, ip, ip_order, ratio, nparts, layers=2, serial_part=None):
    """getSPRSizeField over `nparts` RCB parts, one forked process per part.

    Returns (size [nv], recovered [nnodes, nc], info).  info: "wall_s" (the
    timed step, start to end barrier), "phase_s" [nparts, 4] (each part's A-D
    times), per-part owned counts and local tet counts; with serial_part=k,
    also part k's whole step run again alone in one process after the
    parallel step ("serial_s", "serial_phase_s"), reading the ghost values the
    parallel step published, as a one-core run of a stated fraction.

    Call before the GPU is initialised in this process: the parts are forked.
    """
    g = globals_of(host, ip)
    nc = g["ip"].shape[2]
    lab = rcb(g["xyz"][: g["nv"]], nparts)
    sh = {"rec": _shared((g["nnodes"], nc), np.float64), "esize": _shared((g["nt"],), np.float64),
          "size": _shared((g["nv"],), np.float64), "partial": _shared((nparts, 2), np.float64)}
    times = _shared((nparts, 4), np.float64)
    err = _shared((nparts,), np.int32)
    ctx = multiprocessing.get_context("fork")
    outer, inner = ctx.Barrier(nparts + 1), ctx.Barrier(nparts)
    procs = [ctx.Process(target=_worker, args=(g, lab, k, layers, sh, outer, inner, ip_order, nc,
                                               ratio, times, err)) for k in range(nparts)]
    for pr in procs:
        pr.start()
    try:
        outer.wait()
        t0 = time.perf_counter()
        outer.wait()
        wall = time.perf_counter() - t0
    except threading.BrokenBarrierError as e:
        for pr in procs:
            pr.join()
        raise RuntimeError("partitioned run failed: a part reported an error") from e
    for pr in procs:
        pr.join()
    if err.any() or any(pr.exitcode != 0 for pr in procs):
        raise RuntimeError("partitioned run failed in parts %s" % np.flatnonzero(err).tolist())
    info = {"wall_s": wall, "phase_s": np.array(times), "nparts": nparts, "layers": layers}
    if serial_part is not None:
        k = serial_part
        part = Part(g, lab, k, layers)
        part.patcher = _lib().om_patcher_new(part.L.c(), part.ipL.shape[1], nc)
        st = np.zeros((nparts, 4))
        # the all-reduce reads every part's partials: keep the parallel
        # step's, as this part's own are rewritten with the same values
        t0 = time.perf_counter()
        _step(part, k, sh, None, ip_order, nc, ratio, st)
        info["serial_s"] = time.perf_counter() - t0
        info["serial_phase_s"] = st[k]
        info["serial_counts"] = {"vertices": part.nvo, "edges": part.neo, "tets": part.nto,
                                 "local_tets": part.nlocal_tets}
        _lib().om_patcher_free(part.patcher)
    return np.array(sh["size"]), np.array(sh["rec"]), info


def cores():
    """(host cores, cores this process may run on)."""
    n = os.cpu_count() or 1
    try:
        a = len(os.sched_getaffinity(0))
    except AttributeError:
        a = n
    return n, a
