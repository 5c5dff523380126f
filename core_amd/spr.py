"""Host-side mirror of SCOREC/core's `spr` interface (spr/spr.h:33-63) over
the libsprgpu C ABI.

The reference's calls take an apf::Field* and return a new apf::Field*
attached to the field's mesh.  Here a `Mesh` is the device mirror of one mesh
part and a `Field` is a named device array attached to it, with the apf
layout (nodal Lagrange [nodes, nc]; integration point [nt, q, nc]):

    getGradIPField(f, name, order)                    spr/sprGetGradIPField.cc:15
    recoverField(ip_field)                            spr/sprRecoverField.cc:419
    getSPRSizeField(f, adapt_ratio)                   spr/sprEstimateError.cc:311
    getTargetSPRSizeField(f, t, alpha=0.25, beta=2.0) spr/sprEstimateTargetError.cc:241

Error behaviour: where the reference aborts (PCU_ALWAYS_ASSERT, apf::fail)
these raise SPRError with the library's status code; a duplicate field name
(PCU_ALWAYS_ASSERT_VERBOSE(!m->findField(name)), apf/apf.cc:76) raises too.
Device memory comes from torch; the library runs on its own HIP stream and
every call returns after its work is complete.
"""
import json
import os

import numpy as np
import torch

from . import _lib
from ._lib import SPRError, SPRG_FP32R, SPRG_FP64, SPRG_FP64F
from .mesh import HostMesh, load_smb

SCALAR, VECTOR, MATRIX = 1, 3, 9
FORMS = {"fp64": SPRG_FP64, "fp32r": SPRG_FP32R, "fp64f": SPRG_FP64F}
_CTX = {}


def _sync_torch():
    if torch.cuda.is_available():
        torch.cuda.current_stream().synchronize()


def _read_json(path):
    if not os.path.exists(path):
        return {}
    with open(path) as fh:
        return json.load(fh)


def context(device=0):
    """The library context (one HIP stream) of a device."""
    if device not in _CTX:
        L = _lib.lib()
        h = _lib.ctypes.c_void_p()
        _lib.check(L.sprg_ctx_create(device, _lib.ctypes.byref(h)))
        _CTX[device] = h
    return _CTX[device]


def _dptr(t):
    return _lib.ctypes.c_void_p(t.data_ptr())


class Field:
    """A named field on a Mesh.  kind: 'lagrange' (nodal) or 'ip'."""

    def __init__(self, mesh, name, kind, order, ncomp, data):
        self.mesh = mesh
        self.name = name
        self.kind = kind
        self.order = order
        self.ncomp = ncomp
        self.data = data

    def numpy(self):
        return self.data.detach().cpu().numpy()


class Mesh:
    """Device mirror of one tet mesh part (sprg_mesh_create)."""

    def __init__(self, host, device=0, _handle=None):
        if not isinstance(host, HostMesh):
            raise TypeError("Mesh expects a HostMesh")
        self.host = host
        self.device = device
        self.ctx = context(device)
        self.order = host.order
        self.nv, self.nt, self.ne = host.nv, host.nt, host.ne
        self.nnodes = host.nnodes
        self.fields = {}
        if _handle is not None:
            self.handle = _handle
            return
        h = _lib.ctypes.c_void_p()
        L = _lib.lib()
        _lib.check(L.sprg_mesh_create(self.ctx, host.nv, host.nt, host.ne if host.te is not None else 0,
                                      host.tv.ctypes.data,
                                      None if host.te is None else host.te.ctypes.data,
                                      host.xyz.ctypes.data, host.order, _lib.ctypes.byref(h)))
        self.handle = h

    @classmethod
    def load_smb(cls, path, order=1, device=0):
        """apf::loadMdsPart + (order 2) changeShape to Lagrange-2: the device
        mirror read by sprg_mesh_load_smb; .host holds the same arrays."""
        host = load_smb(path, order)
        h = _lib.ctypes.c_void_p()
        _lib.check(_lib.lib().sprg_mesh_load_smb(context(device), str(path).encode(), order,
                                                 _lib.ctypes.byref(h)))
        return cls(host, device, _handle=h)

    def close(self):
        if self.handle:
            _lib.lib().sprg_mesh_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- field registry (apf::Mesh::findField / addField / removeField) --
    def _add(self, field):
        if field.name in self.fields:
            raise SPRError(_lib.SPRG_EINVAL, "field name %r already exists on the mesh" % field.name)
        self.fields[field.name] = field
        return field

    def findField(self, name):
        return self.fields.get(name)

    def destroyField(self, field):
        self.fields.pop(field.name, None)

    def _tensor(self, shape):
        return torch.empty(shape, dtype=torch.float64, device="cuda:%d" % self.device)

    def createLagrangeField(self, name, ncomp, order, values=None):
        if order not in (1, 2):
            raise SPRError(_lib.SPRG_EUNSUPPORTED, "Lagrange order %d" % order)
        if order == 2 and self.ne == 0:
            raise SPRError(_lib.SPRG_EINVAL, "an order-2 field needs mesh edges")
        n = self.nv if order == 1 else self.nv + self.ne
        data = self._tensor((n, ncomp))
        if values is not None:
            data.copy_(torch.as_tensor(np.asarray(values, dtype=np.float64).reshape(n, ncomp)))
        return self._add(Field(self, name, "lagrange", order, ncomp, data))

    def createIPField(self, name, ncomp, order, values=None):
        q = {1: 1, 2: 4, 3: 5}.get(order)
        if q is None:
            raise SPRError(_lib.SPRG_EUNSUPPORTED, "integration order %d" % order)
        data = self._tensor((self.nt, q, ncomp))
        if values is not None:
            data.copy_(torch.as_tensor(np.asarray(values, dtype=np.float64).reshape(self.nt, q, ncomp)))
        return self._add(Field(self, name, "ip", order, ncomp, data))

    # -- adjacency mirror (for parity checks) --
    def adjacency(self, dim):
        """(off, idx) numpy copies of getAdjacent(entity, 3) for dim 0 or 1."""
        L = _lib.lib()
        off = _lib.ctypes.c_void_p()
        idx = _lib.ctypes.c_void_p()
        _lib.check(L.sprg_mesh_adjacency(self.handle, dim, _lib.ctypes.byref(off), _lib.ctypes.byref(idx)))
        n = self.nv if dim == 0 else self.ne
        o = np.empty(n + 1, np.int32)
        _lib.check(L.sprg_download(self.ctx, o.ctypes.data, off, o.nbytes))
        i = np.empty(int(o[-1]), np.int32)
        if i.size:
            _lib.check(L.sprg_download(self.ctx, i.ctypes.data, idx, i.nbytes))
        return o, i

    def phase_times(self):
        """GPU ms of the last call's phases: fit, expand, expanded fit,
        element pass, factor + average (HIP events on the library stream);
        after getGradIPField, the gradient kernel first and zeros."""
        ms = np.zeros(5)
        _lib.lib().sprg_mesh_phase_times(self.handle, ms.ctypes.data, 5)
        return ms

    def last_stats(self):
        """Patch routes of the last recovery: patches, queued for the exact
        rank test / expansion, expanded sets, FP64 fallback refits, status,
        queued patches whose expansion needed the reference's rank test."""
        st = np.zeros(7, np.int64)
        _lib.lib().sprg_mesh_last_stats(self.handle, st.ctypes.data, 7)
        return dict(zip(["patches", "queued_exact", "expanded", "fallback", "status", "exact_tested",
                         "large_sets"], st.tolist()))

    SHAPE_KEYS = ("fast_threads", "fast_blocks", "exact_blocks", "expand_blocks", "tile_blocks")

    def autotune(self, ip_field, form="fp64f", cache=None):
        """Chooses this mesh's patch-kernel launch shapes by timing
        recoveries of `ip_field` (sprg_mesh_autotune) and returns them.
        With `cache` (a JSON file path), shapes are looked up and stored
        under the key "<gfx arch>/p<mesh order>/c<components>/q<points>/<form>"
        and a cached entry is applied without timing."""
        key = None
        if cache is not None:
            arch = torch.cuda.get_device_properties(self.device).gcnArchName.split(":")[0]
            key = "%s/p%d/c%d/q%d/%s" % (arch, self.order, ip_field.ncomp,
                                         ip_field.data.shape[1], form)
            table = _read_json(cache)
            if key in table:
                self.set_shape(table[key])
                return dict(table[key])
        sh = np.zeros(5, np.int64)
        _sync_torch()
        _lib.check(_lib.lib().sprg_mesh_autotune(self.handle, ip_field.order, ip_field.ncomp,
                                                 _dptr(ip_field.data), _form(form), sh.ctypes.data, 5))
        shape = dict(zip(self.SHAPE_KEYS, sh.tolist()))
        if key is not None:
            table = _read_json(cache)
            table[key] = shape
            with open(cache, "w") as fh:
                json.dump(table, fh, indent=1, sort_keys=True)
        return shape

    def set_shape(self, shape):
        """Applies launch shapes as autotune() returns them."""
        keys = [k for k in self.SHAPE_KEYS if k in shape]
        sh = np.array([int(shape[k]) for k in keys], np.int64)
        _lib.check(_lib.lib().sprg_mesh_set_shape(self.handle, sh.ctypes.data, len(keys)))

    def patch_sets(self, ip_order):
        """(off, idx) of the tet set every recovered entity is fitted on."""
        L = _lib.lib()
        npatch = self.nnodes if self.order == 2 else self.nv
        off = torch.empty(npatch + 1, dtype=torch.int64, device="cuda:%d" % self.device)
        total = np.zeros(1, np.int64)
        _sync_torch()
        _lib.check(L.sprg_patch_sets(self.handle, ip_order, _dptr(off), None, 0, total.ctypes.data))
        idx = torch.empty(max(int(total[0]), 1), dtype=torch.int32, device=off.device)
        _lib.check(L.sprg_patch_sets(self.handle, ip_order, _dptr(off), _dptr(idx), idx.numel(),
                                     total.ctypes.data))
        return off.cpu().numpy(), idx[: int(total[0])].cpu().numpy()


def _form(form):
    if isinstance(form, str):
        if form not in FORMS:
            raise ValueError("form must be one of %s" % sorted(FORMS))
        return FORMS[form]
    return int(form)


def getGradIPField(f, name, order):
    """spr::getGradIPField: gradient of a scalar/vector/matrix Lagrange field
    at the integration points of `order`."""
    if f is None or f.kind != "lagrange":
        raise SPRError(_lib.SPRG_EINVAL, "getGradIPField needs a nodal (Lagrange) field")
    if f.ncomp not in (SCALAR, VECTOR, MATRIX):
        raise SPRError(_lib.SPRG_EINVAL, "getGradIPField: value type must be scalar, vector or matrix")
    m = f.mesh
    out = m.createIPField(name, 3 * f.ncomp, order)
    _sync_torch()
    try:
        _lib.check(_lib.lib().sprg_grad_ip(m.handle, f.order, f.ncomp, _dptr(f.data), order,
                                           _dptr(out.data)))
    except SPRError:
        m.destroyField(out)
        raise
    return out


def recoverField(ip_field, form="fp64"):
    """spr::recoverField: patch-recovered Lagrange field "spr_<name>" of the
    mesh's coordinate order."""
    if ip_field is None or ip_field.kind != "ip":
        raise SPRError(_lib.SPRG_EINVAL, "recoverField needs an integration-point field")
    m = ip_field.mesh
    out = m.createLagrangeField("spr_" + ip_field.name, ip_field.ncomp, m.order)
    _sync_torch()
    try:
        _lib.check(_lib.lib().sprg_recover(m.handle, ip_field.order, ip_field.ncomp,
                                           _dptr(ip_field.data), _dptr(out.data), _form(form)))
    except SPRError:
        m.destroyField(out)
        raise
    return out


def _check_temp_names(m, f):
    for name in ("spr_" + f.name, "esize", "size"):
        if m.findField(name) is not None:
            raise SPRError(_lib.SPRG_EINVAL, "field name %r already exists on the mesh" % name)


def getSPRSizeField(f, adapt_ratio, form="fp64", diag=None):
    """spr::getSPRSizeField: isotropic size field "size" at the vertices."""
    if f is None or f.kind != "ip":
        raise SPRError(_lib.SPRG_EINVAL, "getSPRSizeField needs an integration-point field")
    m = f.mesh
    _check_temp_names(m, f)
    out = m.createLagrangeField("size", 1, 1)
    d = np.zeros(4)
    _sync_torch()
    try:
        _lib.check(_lib.lib().sprg_spr_size(m.handle, f.order, f.ncomp, _dptr(f.data),
                                            float(adapt_ratio), _dptr(out.data), _form(form),
                                            d.ctypes.data))
    except SPRError:
        m.destroyField(out)
        raise
    if diag is not None:
        diag.update(self_product=d[0], error_sum=d[1], size_factor=d[2], wall_ms=d[3])
    return out


def getTargetSPRSizeField(f, t, alpha=0.25, beta=2.0, form="fp64", diag=None):
    """spr::getTargetSPRSizeField: size field aiming at t output elements
    (t is the reference's size_t: a positive integer)."""
    if f is None or f.kind != "ip":
        raise SPRError(_lib.SPRG_EINVAL, "getTargetSPRSizeField needs an integration-point field")
    if int(t) != t or t < 0:
        raise SPRError(_lib.SPRG_EINVAL, "getTargetSPRSizeField: the target is an element count")
    m = f.mesh
    _check_temp_names(m, f)
    out = m.createLagrangeField("size", 1, 1)
    d = np.zeros(4)
    _sync_torch()
    try:
        _lib.check(_lib.lib().sprg_target_spr_size(m.handle, f.order, f.ncomp, _dptr(f.data),
                                                   int(t), float(alpha), float(beta),
                                                   _dptr(out.data), _form(form), d.ctypes.data))
    except SPRError:
        m.destroyField(out)
        raise
    if diag is not None:
        diag.update(global_error=d[0], size_factor=d[1], wall_ms=d[3])
    return out


def averageToVertex(element_values, mesh, name="size"):
    """spr::averageToVertex over every vertex: a vertex field (Lagrange-1
    scalar) holding the mean of an element scalar over the vertex's tets."""
    if element_values.shape[0] != mesh.nt:
        raise SPRError(_lib.SPRG_EINVAL, "one value per tet expected")
    ev = element_values.reshape(mesh.nt).to(device="cuda:%d" % mesh.device, dtype=torch.float64).contiguous()
    out = mesh.createLagrangeField(name, 1, 1)
    _sync_torch()
    try:
        _lib.check(_lib.lib().sprg_average_to_vertex(mesh.handle, _dptr(ev), _dptr(out.data)))
    except SPRError:
        mesh.destroyField(out)
        raise
    return out


class Batch:
    """Several meshes whose size-field pipelines run as one recorded HIP
    graph (sprg_batch_*): the first call records, identical calls replay."""

    def __init__(self, meshes):
        meshes = list(meshes)
        if not meshes:
            raise SPRError(_lib.SPRG_EINVAL, "a batch needs at least one mesh")
        self.meshes = meshes
        self._arr = (_lib.ctypes.c_void_p * len(meshes))(*[m.handle.value for m in meshes])
        h = _lib.ctypes.c_void_p()
        _lib.check(_lib.lib().sprg_batch_create(meshes[0].ctx, len(meshes), self._arr,
                                                _lib.ctypes.byref(h)))
        self.handle = h
        self.last_ms = None

    def close(self):
        if self.handle:
            _lib.lib().sprg_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def size_fields(self, fields, targets=None, adapt_ratio=0.1, alpha=0.25, beta=2.0,
                    form="fp64", outputs=None):
        """For every mesh i: getSPRSizeField(fields[i], adapt_ratio) when
        targets is None, else getTargetSPRSizeField(fields[i], targets[i],
        alpha, beta).  Returns the "size" tensors; pass the previous
        `outputs` back to replay the recorded graph."""
        n = len(self.meshes)
        if len(fields) != n:
            raise SPRError(_lib.SPRG_EINVAL, "one field per mesh")
        f0 = fields[0]
        for f, m in zip(fields, self.meshes):
            if f.kind != "ip" or f.mesh is not m or f.order != f0.order or f.ncomp != f0.ncomp:
                raise SPRError(_lib.SPRG_EINVAL, "batch fields: IP fields of the batch's meshes, "
                                                 "one order and component count")
        if outputs is None:
            outputs = [m._tensor((m.nv,)) for m in self.meshes]
        tg = None
        if targets is not None:
            if len(targets) != n or any(int(t) != t or t <= 0 for t in targets):
                raise SPRError(_lib.SPRG_EINVAL, "batch targets: one positive element count per mesh")
            tg = np.asarray([int(t) for t in targets], dtype=np.uint64)
        ips = (_lib.ctypes.c_void_p * n)(*[f.data.data_ptr() for f in fields])
        outs = (_lib.ctypes.c_void_p * n)(*[o.data_ptr() for o in outputs])
        ms = np.zeros(1)
        _sync_torch()
        _lib.check(_lib.lib().sprg_batch_size(self.handle, f0.order, f0.ncomp, ips,
                                              None if tg is None else tg.ctypes.data,
                                              float(adapt_ratio), float(alpha), float(beta), outs,
                                              _form(form), ms.ctypes.data))
        self.last_ms = float(ms[0])
        return outputs
