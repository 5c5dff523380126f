"""The partitioned CPU run of getSPRSizeField (oracle/partition.py), the
benchmark's CPU baseline, against the serial oracle on the same mesh.

Owned patches are fitted on each part's local mesh, whose tets keep their
global order, so the recovered field must equal the serial oracle's bit for
bit; the size field differs only through the order in which the parts'
partial sums are added (1e-13).  Many thin parts (8 parts of a 6^3 box: a
part is about one cell thick) put most patches next to a part boundary, so
a halo too thin for a patch or its expansion shows up as a changed value.
"""
import numpy as np
import pytest

from core_amd import mesh, workloads
from oracle import oracle as O
from oracle import partition


def _case(n, order, seed):
    h = mesh.box(n, order=order, jitter=0.2, seed=seed)
    ip = workloads.strain_field(h.ip_points(2))
    om = O.Mesh(h.tv, h.xyz, h.order, te=h.te, ev=h.ev)
    return h, ip, om


@pytest.mark.parametrize("order,nparts", [(2, 1), (2, 3), (2, 8), (1, 8), (1, 5)])
def test_partitioned_equals_serial(order, nparts):
    h, ip, om = _case(6, order, 5 + order)
    rec = O.recover(om, ip)
    size, _ = O.spr_size(om, 2, ip, 0.05)
    psize, prec, info = partition.spr_size(h, ip, 2, 0.05, nparts)
    assert np.array_equal(prec, rec)
    assert np.max(np.abs(psize / size - 1)) <= 1e-13
    assert info["phase_s"].shape == (nparts, 4) and info["wall_s"] > 0


def test_rcb_balances_and_covers():
    x = np.random.default_rng(3).random((1001, 3)) * [4, 1, 2]
    for k in (1, 2, 3, 7, 16):
        lab = partition.rcb(x, k)
        cnt = np.bincount(lab, minlength=k)
        assert len(cnt) == k and cnt.sum() == 1001 and cnt.max() - cnt.min() <= 1


def test_one_layer_halo_is_too_thin():
    """A one-layer halo cuts the expanded patches at part boundaries: the
    check above would see it (the planted mistake is in the partitioning)."""
    h, ip, om = _case(6, 2, 7)
    rec = O.recover(om, ip)
    try:
        _, prec, _ = partition.spr_size(h, ip, 2, 0.05, 8, layers=1)
    except RuntimeError:  # a cut patch that cannot be fitted at all
        return
    assert not np.array_equal(prec, rec)


def test_serial_leg_repeats_one_part():
    h, ip, om = _case(5, 2, 9)
    size, _ = O.spr_size(om, 2, ip, 0.05)
    psize, _, info = partition.spr_size(h, ip, 2, 0.05, 4, serial_part=0)
    assert np.max(np.abs(psize / size - 1)) <= 1e-13
    c = info["serial_counts"]
    assert c["vertices"] > 0 and c["edges"] > 0 and c["tets"] > 0 and info["serial_s"] > 0
