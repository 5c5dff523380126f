"""GPU accuracy of the fast forms (FP32 + FP64 refinement, and FP64 normal
equations) against the FP64 oracle.

Recovered fields are compared per component, normalised by the component's
max |oracle| (components that are identically zero are skipped); size fields
pointwise (they are strictly positive).  The bars for the fast form:
<= 1e-6 relative on the size field; the recovered field
is held to 1e-9 of the component scale.
"""
import numpy as np
import pytest

from core_amd import mesh, spr, workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CASES = [
    ("p1_jitter", 6, 1, 0.25, 11, False, 1),
    ("p2", 4, 2, 0.0, 0, False, 2),
    ("p2_jitter", 8, 2, 0.25, workloads.C5, False, 2),
    ("p2_graded", 6, 2, 0.3, workloads.C3, True, 2),
]
FORMS = ["fp32r", "fp64f"]


def _rel_components(got, ref):
    scale = np.abs(ref).max(0)
    keep = scale > 0
    return np.abs(got - ref)[:, keep] / scale[keep]


@pytest.fixture(scope="module", params=CASES, ids=[c[0] for c in CASES])
def case(request, gpu):
    _, n, order, jit, seed, grade, ip_order = request.param
    h = mesh.box(n, order=order, jitter=jit, seed=seed, grade=grade)
    om = O.Mesh(h.tv, h.xyz, h.order, te=h.te, ev=h.ev)
    ip = workloads.strain_field(h.ip_points(ip_order))
    return h, spr.Mesh(h), om, ip, ip_order


@pytest.mark.parametrize("form", FORMS)
def test_fast_recover(case, form):
    h, m, om, ip, ip_order = case
    f = m.createIPField("eps_%s" % form, 9, ip_order, ip)
    r = spr.recoverField(f, form=form)
    got = r.numpy()
    m.destroyField(f)
    m.destroyField(r)
    ref = O.recover(om, ip)
    err = _rel_components(got, ref)
    print("%s recovered: max %.3e rms %.3e" % (form, err.max(), np.sqrt((err ** 2).mean())))
    assert err.max() <= 1e-9


@pytest.mark.parametrize("form", FORMS)
def test_fast_size(case, form):
    h, m, om, ip, ip_order = case
    f = m.createIPField("eps_s_%s" % form, 9, ip_order, ip)
    s = spr.getSPRSizeField(f, 0.05, form=form)
    got = s.numpy()[:, 0]
    m.destroyField(f)
    m.destroyField(s)
    ref, _ = O.spr_size(om, ip_order, ip, 0.05)
    err = np.abs(got / ref - 1)
    print("%s size: max %.3e rms %.3e" % (form, err.max(), np.sqrt((err ** 2).mean())))
    assert err.max() <= 1e-6


@pytest.mark.parametrize("form", FORMS)
def test_fast_patch_rank_decisions_match(case, form):
    # the fast forms route every patch they cannot certify to the exact
    # path, so their expanded patches must be the reference's
    h, m, om, ip, ip_order = case
    f = m.createIPField("eps_p_%s" % form, 9, ip_order, ip)
    r1 = spr.recoverField(f, form=form).numpy()
    m.destroyField(m.findField("spr_" + f.name))
    r0 = spr.recoverField(f, form="fp64").numpy()
    m.destroyField(m.findField("spr_" + f.name))
    m.destroyField(f)
    assert np.isfinite(r1).all()
    err = _rel_components(r1, r0)
    assert err.max() <= 1e-9
