"""GPU parity on the full-size synthetic workloads (core_amd/workloads.py).

  cube_sin    n=8 P1 cube, u = sin sin sin, IP order 1, getSPRSizeField(0.1):
              everything against the oracle (gradients and recovered field
              bit for bit, size field to 1e-13; fast forms to 1e-6).
  cube_exp    n=64 P2 lattice cube (1,572,864 tets), u = exp(-10|x-c|^2),
              IP order 2, recoverField.  Most of its patches are
              rank-deficient on the lattice and need expansion.
  graded_vec  n=64 P2 cube, jitter 0.3, x -> x^2 grading, vector field,
              9-component gradient at IP order 2, recoverField.
  strain_120  the bench mesh (10,368,000 tets), getSPRSizeField(0.05).

At full size the oracle runs on every tet for the gradients and on sampled
ranges of patches for the recovered field (first vertices, including the
boundary, middle vertices, the first and middle edges), where the FP64 form
must agree bit for bit; over every patch the fast forms are held to the FP64
GPU form (which equals the oracle on the ranges).
"""
import numpy as np
import pytest

from core_amd import spr, workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _ranges(h, width=2000):
    nv, np_ = h.nv, h.nv + (h.ne if h.order == 2 else 0)
    starts = [0, nv // 2, nv, nv + (np_ - nv) // 2] if h.order == 2 else [0, nv // 2, nv - width]
    return [(s, min(s + width, np_)) for s in starts]


def _rel_components(got, ref):
    scale = np.abs(ref).max(0)
    keep = scale > 0
    return np.abs(got - ref)[:, keep] / scale[keep]


def _omesh(h):
    return O.Mesh(h.tv, h.xyz, h.order, te=h.te, ev=h.ev)


def test_cube_sin(gpu):
    h, u = workloads.cube_sin(8)
    assert (h.nt, h.nv) == (3072, 729)
    om = _omesh(h)
    m = spr.Mesh(h)
    f = m.createLagrangeField("u", 1, 1, u)
    eps = spr.getGradIPField(f, "eps", 1)
    g = eps.numpy()
    assert np.array_equal(g, O.grad_ip(om, 1, u, 1))
    r = spr.recoverField(eps)
    assert np.array_equal(r.numpy(), O.recover(om, g))
    m.destroyField(r)
    ref, odiag = O.spr_size(om, 1, g, 0.1)
    for form, tol in (("fp64", 1e-13), ("fp64f", 1e-6), ("fp32r", 1e-6)):
        s = spr.getSPRSizeField(eps, 0.1, form=form)
        err = np.max(np.abs(s.numpy()[:, 0] / ref - 1))
        m.destroyField(s)
        print("cube_sin %s size max rel %.3e" % (form, err))
        assert err <= tol


@pytest.mark.parametrize("name", ["cube_exp", "graded_vec"])
def test_n64_recover(gpu, name):
    h, u = getattr(workloads, name)(64)
    assert h.nt == 1572864
    om = _omesh(h)
    m = spr.Mesh(h)
    nc = 1 if u.ndim == 1 else u.shape[1]
    f = m.createLagrangeField("u", nc, 2, u)
    eps = spr.getGradIPField(f, "eps", 2)
    g = eps.numpy()
    assert np.array_equal(g, O.grad_ip(om, 2, u, 2))
    r64 = spr.recoverField(eps)
    got = r64.numpy()
    stats = m.last_stats()
    print("%s routes %s" % (name, stats))
    assert stats["status"] == 0 and stats["expanded"] == stats["queued_exact"]
    for p0, p1 in _ranges(h):
        ref = O.recover(om, g, p0, p1)
        assert np.array_equal(got[p0:p1], ref[p0:p1]), (p0, p1)
    m.destroyField(r64)
    for form in ("fp64f", "fp32r"):
        rf = spr.recoverField(eps, form=form)
        err = _rel_components(rf.numpy(), got)
        m.destroyField(rf)
        print("%s %s recovered vs FP64: max %.3e rms %.3e" % (name, form, err.max(),
                                                            np.sqrt((err ** 2).mean())))
        assert err.max() <= 1e-9


def test_strain120(gpu):
    """The bench workload.  The FP64 recovered field is compared with the
    oracle bit for bit on sampled patch ranges; the estimator stage of the
    oracle then runs over all 10,368,000 tets on that recovered field, and
    the GPU size fields are held to it: the FP64 form to 1e-12 (summation
    order of the two global sums and of the vertex averages aside, it
    computes the same numbers), the fast forms to 1e-6.  The same for
    getTargetSPRSizeField aiming at as many tets as the mesh has (its
    alpha/beta clamps active on part of the mesh); then getGradIPField of a
    vector field on the same mesh, bit for bit on sampled tet ranges."""
    h = workloads.strain_mesh(120)
    assert (h.nt, h.nv, h.ne) == (10368000, 1771561, 12225960)
    ip = workloads.strain_ip(h)
    om = _omesh(h)
    m = spr.Mesh(h)
    eps = m.createIPField("eps", 9, 2, ip)
    r64 = spr.recoverField(eps)
    got = r64.numpy()
    m.destroyField(r64)
    for p0, p1 in _ranges(h, width=500):
        ref = O.recover(om, ip, p0, p1)
        assert np.array_equal(got[p0:p1], ref[p0:p1]), (p0, p1)
    oref, odiag = O.spr_size_rec(om, 2, ip, got, 0.05)
    assert np.all(oref > 0) and np.all(np.isfinite(oref))
    s64 = spr.getSPRSizeField(eps, 0.05)
    a64 = s64.numpy()[:, 0]
    m.destroyField(s64)
    err64 = np.abs(a64 / oref - 1)
    print("strain_120 fp64 size vs oracle: max %.3e rms %.3e" % (err64.max(), np.sqrt((err64 ** 2).mean())))
    assert err64.max() <= 1e-12
    for form in ("fp64f", "fp32r"):
        s = spr.getSPRSizeField(eps, 0.05, form=form)
        err = np.abs(s.numpy()[:, 0] / oref - 1)
        m.destroyField(s)
        print("strain_120 %s size vs oracle: max %.3e rms %.3e" % (form, err.max(), np.sqrt((err ** 2).mean())))
        assert err.max() <= 1e-6
    tref, tdiag = O.target_size_rec(om, 2, ip, got, float(h.nt))
    assert np.all(tref > 0) and np.all(np.isfinite(tref))
    for form, tol in (("fp64", 1e-12), ("fp64f", 1e-6)):
        diag = {}
        s = spr.getTargetSPRSizeField(eps, h.nt, form=form, diag=diag)
        err = np.abs(s.numpy()[:, 0] / tref - 1)
        m.destroyField(s)
        print("strain_120 %s target size vs oracle: max %.3e rms %.3e (G %.9e / %.9e)"
              % (form, err.max(), np.sqrt((err ** 2).mean()), diag["global_error"], tdiag[0]))
        assert abs(diag["global_error"] / tdiag[0] - 1) <= tol
        assert err.max() <= tol
    m.destroyField(eps)
    # getGradIPField at this size: a Lagrange-2 vector field's IP gradients
    # (10.4M tets x 4 points x 9 values; the last tets' outputs lie past 2^31
    # bytes) bit for bit on tet ranges at the start, middle and end, the
    # oracle running on the sub-mesh of each range (a tet's gradient depends
    # on its own nodes only)
    P = workloads.node_points(h)
    u = np.stack([workloads.sin_field(P), workloads.exp_field(P), P[:, 0] * P[:, 1]], 1)
    fu = m.createLagrangeField("u", 3, 2, u)
    gu = spr.getGradIPField(fu, "gu", 2)
    ggot = gu.numpy()
    m.destroyField(gu)
    m.destroyField(fu)
    assert ggot.shape == (h.nt, 4, 9)
    w = 100000
    for t0 in (0, h.nt // 2, h.nt - w):
        sub = O.Mesh(h.tv[t0:t0 + w], h.xyz, h.order, te=h.te[t0:t0 + w], ev=h.ev)
        assert np.array_equal(ggot[t0:t0 + w], O.grad_ip(sub, 2, u, 2)), t0
