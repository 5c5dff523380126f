"""Synthetic SPR workloads: seeded box meshes and the fields put on them.

  cube_sin      3,072-tet P1 cube (n=8), u = sin(pi x) sin(pi y) sin(pi z) at
                the vertices, IP order 1, getSPRSizeField(0.1)
  cube_exp      n=64 P2 cube, u = exp(-10|x-c|^2) at the P2 nodes, IP order 2
  graded_vec    n=64 P2 cube, jitter 0.3 (seed 0x5EED0003), x -> x^2 grading,
                v = (sin 2pi y, cos 2pi z, x y) at the P2 nodes (9-comp IP field)
  multi_bump    256 jittered cubes, n_i = round(exp(U[ln 6, ln 40])), seed
                0x5EED0004 + i, three Gaussian bumps each, target = 2 x tets
  strain_120    n=120 P2 cube (10,368,000 tets), jitter 0.25 (seed
                0x5EED0005), eps = sym(grad v) for v = (sin(pi x) y, z^2,
                cos(pi y) x) at the IP positions, 9 comps, getSPRSizeField(0.05)

Fields are plain numpy arrays in the apf layouts.  Uniforms come from
splitmix64 ((x >> 11) * 2^-53), which is portable across platforms.
"""
import math

import numpy as np

from . import mesh as _mesh

C1, C2, C3, C4, C5 = 0x5EED0001, 0x5EED0002, 0x5EED0003, 0x5EED0004, 0x5EED0005


class SplitMix64:
    def __init__(self, seed):
        self.s = int(seed) & (2**64 - 1)

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & (2**64 - 1)
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2**64 - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2**64 - 1)
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next() >> 11) * 2.0**-53


def node_points(m):
    """Coordinates of the Lagrange nodes of the mesh's order."""
    return m.xyz[: m.nnodes]


def sin_field(P):
    return np.sin(np.pi * P[:, 0]) * np.sin(np.pi * P[:, 1]) * np.sin(np.pi * P[:, 2])


def exp_field(P, c=(0.5, 0.5, 0.5)):
    d = P - np.asarray(c)
    return np.exp(-10.0 * (d * d).sum(1))


def vec_field(P):
    x, y, z = P[:, 0], P[:, 1], P[:, 2]
    return np.stack([np.sin(2 * np.pi * y), np.cos(2 * np.pi * z), x * y], 1)


def strain_field(P):
    """eps = sym(grad v), v = (sin(pi x) y, z^2, cos(pi y) x), row-major 3x3."""
    x, y, z = P[..., 0], P[..., 1], P[..., 2]
    pi = np.pi
    e = np.zeros(P.shape[:-1] + (3, 3))
    e[..., 0, 0] = pi * np.cos(pi * x) * y
    e[..., 0, 1] = e[..., 1, 0] = 0.5 * np.sin(pi * x)
    e[..., 0, 2] = e[..., 2, 0] = 0.5 * np.cos(pi * y)
    e[..., 1, 2] = e[..., 2, 1] = 0.5 * (2 * z - pi * np.sin(pi * y) * x)
    return e.reshape(P.shape[:-1] + (9,))


def bump_field(P, rng):
    """Three Gaussian bumps: centres U[0,1]^3, widths U[0.05,0.3], amplitude 1."""
    u = np.zeros(P.shape[0])
    for _ in range(3):
        c = np.array([rng.uniform() for _ in range(3)])
        w = 0.05 + 0.25 * rng.uniform()
        d = P - c
        u += np.exp(-(d * d).sum(1) / (w * w))
    return u


def cube_sin(n=8):
    m = _mesh.box(n, order=1)
    return m, sin_field(node_points(m))


def cube_exp(n=64):
    m = _mesh.box(n, order=2)
    return m, exp_field(node_points(m))


def graded_vec(n=64):
    m = _mesh.box(n, order=2, jitter=0.3, seed=C3, grade=True)
    return m, vec_field(node_points(m))


def multi_bump_sizes(count=256):
    rng = SplitMix64(C4)
    lo, hi = math.log(6), math.log(40)
    return [int(round(math.exp(lo + (hi - lo) * rng.uniform()))) for _ in range(count)]


def multi_bump(i, n):
    m = _mesh.box(n, order=2, jitter=0.25, seed=C4 + i)
    rng = SplitMix64(C4 + i + 0x1000)
    return m, bump_field(node_points(m), rng)


def strain_mesh(n=120, jitter=0.25, seed=C5):
    return _mesh.box(n, order=2, jitter=jitter, seed=seed)


def strain_ip(m, ip_order=2, chunk=1 << 20, progress=None):
    """The strain field at the IP positions, computed in chunks of tets."""
    xi, _ = _mesh.tet_rule(ip_order)
    lam = np.stack([1 - xi.sum(1), xi[:, 0], xi[:, 1], xi[:, 2]], 1)
    out = np.empty((m.nt, lam.shape[0], 9))
    for t0 in range(0, m.nt, chunk):
        tv = m.tv[t0:t0 + chunk]
        P = np.zeros((tv.shape[0], lam.shape[0], 3))
        for a in range(4):
            P += lam[None, :, a, None] * m.xyz[tv[:, a]][:, None, :]
        out[t0:t0 + chunk] = strain_field(P)
        if progress:
            progress(min(t0 + chunk, m.nt), m.nt)
    return out
