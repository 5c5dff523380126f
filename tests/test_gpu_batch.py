"""Many meshes in one recorded HIP graph (sprg_batch_*).

The multi_bump workload: 256 jittered P2 cubes, n_i = round(exp(U[ln 6,
ln 40])) (seed 0x5EED0004), three Gaussian bumps each, gradient at IP order
2, getTargetSPRSizeField(eps, 2 x tets, 0.25, 2.0) per mesh.  The batch's
first call records every pipeline into one graph; later calls replay it.
Every replay must equal the per-mesh calls bit for bit (the same kernels in
the same order), with ordinary eager work and synchronisation between the
replays; the smallest meshes are also checked against the oracle.
"""
import time

import numpy as np
import pytest
import torch

from core_amd import spr, workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _build(sizes):
    meshes, fields = [], []
    for i, n in enumerate(sizes):
        h, u = workloads.multi_bump(i, n)
        m = spr.Mesh(h)
        f = m.createLagrangeField("u", 1, 2, u)
        eps = spr.getGradIPField(f, "eps", 2)
        meshes.append(m)
        fields.append((h, u, eps))
    return meshes, fields


@pytest.mark.parametrize("form", ["fp64", "fp64f"])
def test_multi_bump_batch(gpu, form):
    sizes = workloads.multi_bump_sizes()
    t0 = time.time()
    meshes, fields = _build(sizes)
    print("built %d meshes, %d tets (%.1fs)" % (len(meshes), sum(m.nt for m in meshes), time.time() - t0))
    targets = [2 * m.nt for m in meshes]
    # per-mesh eager calls
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eager = []
    for (h, u, eps), t in zip(fields, targets):
        s = spr.getTargetSPRSizeField(eps, t, 0.25, 2.0, form=form)
        eager.append(s.data[:, 0].clone())
        eps.mesh.destroyField(s)
    torch.cuda.synchronize()
    eager_ms = (time.perf_counter() - t0) * 1e3
    batch = spr.Batch(meshes)
    outs = batch.size_fields([f[2] for f in fields], targets, alpha=0.25, beta=2.0, form=form)
    record_ms = batch.last_ms
    for k in range(3):
        for o in outs:
            o.fill_(-1.0)  # eager work between replays
        torch.cuda.synchronize()
        outs = batch.size_fields([f[2] for f in fields], targets, alpha=0.25, beta=2.0, form=form,
                                 outputs=outs)
        for i, (o, e) in enumerate(zip(outs, eager)):
            assert torch.equal(o, e), (k, i, sizes[i])
    print("%s: eager %.1f ms for %d meshes; graph replay %.1f ms (first run incl. recording %.1f ms)"
          % (form, eager_ms, len(meshes), batch.last_ms, record_ms))
    if form == "fp64":
        small = [i for i, n in enumerate(sizes) if n <= 8][:4]
        assert small
        for i in small:
            h, u, eps = fields[i]
            om = O.Mesh(h.tv, h.xyz, h.order, te=h.te, ev=h.ev)
            ref, _ = O.target_size(om, 2, eps.numpy(), targets[i], 0.25, 2.0)
            got = outs[i].cpu().numpy()
            assert np.max(np.abs(got / ref - 1)) <= 1e-13, i
    batch.close()


def test_batch_records_again_after_workspace_growth(gpu):
    """An eager call that grows a batch mesh's workspaces (27 components
    instead of 3) moves them to new device addresses; the next batch call
    must record again rather than replay kernels bound to the freed ones,
    and still equal the per-mesh calls."""
    meshes, fields = _build([3, 4, 5])
    targets = [2 * m.nt for m in meshes]
    eager = []
    for (h, u, eps), t in zip(fields, targets):
        s = spr.getTargetSPRSizeField(eps, t, 0.25, 2.0, form="fp64f")
        eager.append(s.data[:, 0].clone())
        eps.mesh.destroyField(s)
    batch = spr.Batch(meshes)
    outs = batch.size_fields([f[2] for f in fields], targets, form="fp64f")
    for o, e in zip(outs, eager):
        assert torch.equal(o, e)
    for m, (h, u, eps) in zip(meshes, fields):
        ip27 = np.tile(workloads.strain_field(h.ip_points(2)), (1, 1, 3))
        f27 = m.createIPField("eps27", 27, 2, ip27)
        s = spr.getSPRSizeField(f27, 0.1, form="fp64f")
        m.destroyField(s)
        m.destroyField(f27)
    for o in outs:
        o.fill_(-1.0)
    outs = batch.size_fields([f[2] for f in fields], targets, form="fp64f", outputs=outs)
    for o, e in zip(outs, eager):
        assert torch.equal(o, e)
    batch.close()
