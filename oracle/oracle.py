"""ctypes wrapper of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  Imported by
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the
product package.  Each function mirrors one function of spr_oracle.c, which
cites the reference file:line it restates.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_i64 = ctypes.c_int64
_p = ctypes.c_void_p


class _Mesh(ctypes.Structure):
    _fields_ = [("nv", _i64), ("nt", _i64), ("ne", _i64), ("tv", _p),
                ("te", _p), ("ev", _p), ("xyz", _p), ("order", ctypes.c_int)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run make -C oracle")
        _LIB = ctypes.CDLL(path)
        _LIB.om_last_error.restype = ctypes.c_char_p
    return _LIB


def ref_qr():
    """The reference's own mth/mthQR.cc built by oracle/Makefile, or None."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libmthqr_ref.so")
        if not os.path.exists(path):
            return None
        _REF = ctypes.CDLL(path)
    return _REF


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_p)


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle error %d: %s" % (rc, lib().om_last_error().decode()))


class Mesh:
    """Plain arrays of one tet mesh part (see spr_oracle.h)."""

    def __init__(self, tv, xyz, order, te=None, ev=None):
        self.tv = np.ascontiguousarray(tv, dtype=np.int32)
        self.te = None if te is None else np.ascontiguousarray(te, dtype=np.int32)
        self.ev = None if ev is None else np.ascontiguousarray(ev, dtype=np.int32)
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        self.order = int(order)
        self.nt = self.tv.shape[0]
        self.ne = 0 if self.te is None else int(self.te.max()) + 1
        if self.ev is not None:
            self.ne = self.ev.shape[0]
        nnodes = self.xyz.shape[0]
        self.nv = nnodes - (self.ne if order == 2 else 0)
        self._c = _Mesh(self.nv, self.nt, self.ne, _ptr(self.tv), _ptr(self.te),
                        _ptr(self.ev), _ptr(self.xyz), self.order)

    @property
    def nnodes(self):
        return self.nv + (self.ne if self.order == 2 else 0)

    def c(self):
        return ctypes.byref(self._c)


def rule(order):
    n = lib().om_rule_npoints(order)
    if n < 0:
        raise ValueError("unsupported order")
    xi = np.zeros((n, 3))
    w = np.zeros(n)
    _check(lib().om_rule(order, _ptr(xi), _ptr(w)))
    return xi, w


def decompose_qr(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    m, n = a.shape
    q = np.zeros((m, m))
    r = np.zeros((m, n))
    rank = lib().om_decompose_qr(m, n, _ptr(a), _ptr(q), _ptr(r))
    return rank, q, r


def solve_from_qr(q, r, b):
    m, n = r.shape
    x = np.zeros(n)
    b = np.ascontiguousarray(b, dtype=np.float64)
    lib().om_solve_from_qr(m, n, _ptr(np.ascontiguousarray(q)),
                           _ptr(np.ascontiguousarray(r)), _ptr(b), _ptr(x))
    return x


def vertex_tets(mesh):
    off = np.zeros(mesh.nv + 1, np.int32)
    idx = np.zeros(4 * mesh.nt, np.int32)
    _check(lib().om_vertex_tets(mesh.c(), _ptr(off), _ptr(idx)))
    return off, idx


def edge_tets(mesh):
    off = np.zeros(mesh.ne + 1, np.int32)
    idx = np.zeros(6 * mesh.nt, np.int32)
    _check(lib().om_edge_tets(mesh.c(), _ptr(off), _ptr(idx)))
    return off, idx


def grad_ip(mesh, field_order, nodal, ip_order):
    nodal = np.ascontiguousarray(nodal, dtype=np.float64)
    nc_in = 1 if nodal.ndim == 1 else nodal.shape[1]
    q = lib().om_rule_npoints(ip_order)
    ip = np.zeros((mesh.nt, q, 3 * nc_in))
    _check(lib().om_grad_ip(mesh.c(), field_order, nc_in, _ptr(nodal), ip_order, _ptr(ip)))
    return ip


def patch_sets(mesh, q):
    npatch = mesh.nnodes if mesh.order == 2 else mesh.nv
    off = np.zeros(npatch + 1, np.int32)
    cap = 40 * mesh.nt + 64
    used = _i64(0)
    while True:
        idx = np.zeros(cap, np.int32)
        rc = lib().om_patch_sets(mesh.c(), q, _ptr(off), _ptr(idx), _i64(cap), ctypes.byref(used))
        if rc == 2 and used.value > cap:  # capacity: retry with the size it needs
            cap = used.value
            continue
        break
    _check(rc)
    return off, idx[:used.value].copy()


def recover(mesh, ip, p0=0, p1=None):
    ip = np.ascontiguousarray(ip, dtype=np.float64)
    nt, q, nc = ip.shape
    npatch = mesh.nnodes if mesh.order == 2 else mesh.nv
    out = np.zeros((npatch, nc))
    if p1 is None:
        p1 = npatch
    _check(lib().om_recover_range(mesh.c(), q, nc, _ptr(ip), _ptr(out), _i64(p0), _i64(p1)))
    return out


def spr_size(mesh, ip_order, ip, adapt_ratio):
    ip = np.ascontiguousarray(ip, dtype=np.float64)
    nc = ip.shape[2]
    size = np.zeros(mesh.nv)
    diag = np.zeros(4)
    lib().om_spr_size.argtypes = [_p, ctypes.c_int, ctypes.c_int, _p, ctypes.c_double, _p, _p]
    _check(lib().om_spr_size(mesh.c(), ip_order, nc, _ptr(ip), adapt_ratio, _ptr(size), _ptr(diag)))
    return size, diag


def spr_size_rec(mesh, ip_order, ip, rec, adapt_ratio):
    """getSPRSizeField's estimator stage from a given recovered field."""
    ip = np.ascontiguousarray(ip, dtype=np.float64)
    rec = np.ascontiguousarray(rec, dtype=np.float64)
    nc = ip.shape[2]
    size = np.zeros(mesh.nv)
    diag = np.zeros(4)
    f = lib().om_spr_size_rec
    f.argtypes = [_p, ctypes.c_int, ctypes.c_int, _p, _p, ctypes.c_double, _p, _p]
    _check(f(mesh.c(), ip_order, nc, _ptr(ip), _ptr(rec), adapt_ratio, _ptr(size), _ptr(diag)))
    return size, diag


def target_size(mesh, ip_order, ip, target, alpha=0.25, beta=2.0):
    ip = np.ascontiguousarray(ip, dtype=np.float64)
    nc = ip.shape[2]
    size = np.zeros(mesh.nv)
    diag = np.zeros(4)
    lib().om_target_size.argtypes = [_p, ctypes.c_int, ctypes.c_int, _p, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double, _p, _p]
    _check(lib().om_target_size(mesh.c(), ip_order, nc, _ptr(ip), float(target), alpha, beta,
                                _ptr(size), _ptr(diag)))
    return size, diag


def target_size_rec(mesh, ip_order, ip, rec, target, alpha=0.25, beta=2.0):
    """getTargetSPRSizeField's estimator stage from a given recovered field."""
    ip = np.ascontiguousarray(ip, dtype=np.float64)
    rec = np.ascontiguousarray(rec, dtype=np.float64)
    nc = ip.shape[2]
    size = np.zeros(mesh.nv)
    diag = np.zeros(4)
    f = lib().om_target_size_rec
    f.argtypes = [_p, ctypes.c_int, ctypes.c_int, _p, _p, ctypes.c_double, ctypes.c_double,
                  ctypes.c_double, _p, _p]
    _check(f(mesh.c(), ip_order, nc, _ptr(ip), _ptr(rec), float(target), alpha, beta,
             _ptr(size), _ptr(diag)))
    return size, diag


def element_error(mesh, ip_order, ip, rec):
    ip = np.ascontiguousarray(ip, dtype=np.float64)
    rec = np.ascontiguousarray(rec, dtype=np.float64)
    s = np.zeros(mesh.nt)
    _check(lib().om_element_error(mesh.c(), ip_order, ip.shape[2], _ptr(ip), _ptr(rec), _ptr(s)))
    return s


def current_size(mesh):
    h = np.zeros(mesh.nt)
    _check(lib().om_current_size(mesh.c(), _ptr(h)))
    return h


def element_error_range(mesh, ip_order, ip, rec, t0, t1):
    ip = np.ascontiguousarray(ip, dtype=np.float64)
    rec = np.ascontiguousarray(rec, dtype=np.float64)
    s = np.zeros(mesh.nt)
    f = lib().om_element_error_range
    f.argtypes = [_p, ctypes.c_int, ctypes.c_int, _p, _p, _p, _i64, _i64]
    _check(f(mesh.c(), ip_order, ip.shape[2], _ptr(ip), _ptr(rec), _ptr(s), t0, t1))
    return s
