// box.cpp -- synthetic tetrahedral box meshes with the same numbering as
// apf::makeMdsBox (mds/apfBox.cc:178-336) followed by a Lagrange-2 shape
// change (apf/apfMesh.cc:736-762), plus seeded interior jitter and an
// x -> x^2 grading.  Host code; no device work.
//
// Numbering (what makeMdsBox produces on a fresh mesh):
//   vertices  grid order, x fastest: v = x + (n+1)*(y + (n+1)*z), at
//             (1/n)*index (BoxBuilder::buildCellVert)
//   edges     1. axis edges, per grid point in order, directions x,y,z
//                (buildCellEdges)
//             2. quad-face diagonals, per grid point, face pairs (x,y),
//                (y,z), (z,x), from the lower to the upper corner
//                (buildCellFaces -> buildTriangles: tri edge (2,0))
//             3. hex body diagonals, per hex (created by the first tet)
//   tets      per hex in grid order, 6 tets from the Kuhn table
//             (BoxBuilder::buildTets)
//   tet->edge apf::tet_edge_verts order (apf/apfMesh.cc:53-60)
//   P2 nodes  edge node = (0 + x0*0.5) + x1*0.5, the linear edge element
//             evaluated at xi = 0 (apf/apfFieldOf.cc:7-46)
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "sprgpu_internal.h"

namespace {

const int kuhn[6][4] = {{0, 1, 2, 6}, {0, 2, 3, 6}, {0, 3, 7, 6},
                        {0, 7, 4, 6}, {0, 4, 5, 6}, {0, 5, 1, 6}};
const int hexoff[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                          {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
const int tev[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};

struct SplitMix64 {
  uint64_t s;
  uint64_t next()
  {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};

double tet_det(const double* a, const double* b, const double* c,
               const double* d)
{
  double u[3], v[3], w[3];
  for (int j = 0; j < 3; ++j) {
    u[j] = b[j] - a[j];
    v[j] = c[j] - a[j];
    w[j] = d[j] - a[j];
  }
  return u[0] * (v[1] * w[2] - v[2] * w[1]) - u[1] * (v[0] * w[2] - v[2] * w[0]) +
         u[2] * (v[0] * w[1] - v[1] * w[0]);
}

}  // namespace

extern "C" int sprg_box_counts(int n, int64_t* nv, int64_t* nt, int64_t* ne)
{
  if (n < 1) return sprg_fail(SPRG_EINVAL, "box: n must be >= 1");
  int64_t N = n;
  *nv = (N + 1) * (N + 1) * (N + 1);
  *nt = 6 * N * N * N;
  *ne = 3 * N * (N + 1) * (N + 1) + 3 * N * N * (N + 1) + N * N * N;
  return 0;
}

extern "C" int sprg_box_build(int n, int order, double jitter, uint64_t seed,
                              int grade, int32_t* tv, int32_t* te, int32_t* ev,
                              double* xyz)
{
  int64_t nv, nt, ne;
  int rc = sprg_box_counts(n, &nv, &nt, &ne);
  if (rc) return rc;
  if (order != 1 && order != 2)
    return sprg_fail(SPRG_EINVAL, "box: order must be 1 or 2");
  if ((int64_t)nt > INT32_MAX / 8 || nv + ne > INT32_MAX)
    return sprg_fail(SPRG_EINVAL, "box: mesh too large for int32 indices");
  const int64_t g = n + 1;
  auto vid = [g](int64_t x, int64_t y, int64_t z) { return x + g * (y + g * z); };
  const double w = 1.0 / n;
  // vertices (BoxBuilder::buildCellVert)
  std::vector<double> base(3 * nv), disp(3 * nv, 0.0);
  for (int64_t i = 0; i < nv; ++i) {
    int64_t x = i % g, y = (i / g) % g, z = i / (g * g);
    base[3 * i + 0] = w * (double)x;
    base[3 * i + 1] = w * (double)y;
    base[3 * i + 2] = w * (double)z;
    xyz[3 * i + 0] = base[3 * i + 0];
    xyz[3 * i + 1] = base[3 * i + 1];
    xyz[3 * i + 2] = base[3 * i + 2];
  }
  // tets (BoxBuilder::buildCellRegion / buildTets)
  int64_t t = 0;
  for (int64_t z = 0; z < n; ++z)
    for (int64_t y = 0; y < n; ++y)
      for (int64_t x = 0; x < n; ++x)
        for (int k = 0; k < 6; ++k, ++t)
          for (int j = 0; j < 4; ++j) {
            const int* o = hexoff[kuhn[k][j]];
            tv[4 * t + j] = (int32_t)vid(x + o[0], y + o[1], z + o[2]);
          }
  // jitter interior vertices in lattice units, then grade, then repair
  if (jitter > 0 || grade) {
    SplitMix64 rng{seed};
    for (int64_t i = 0; i < nv; ++i) {
      int64_t c[3] = {i % g, (i / g) % g, i / (g * g)};
      double u[3] = {rng.uniform(), rng.uniform(), rng.uniform()};
      bool interior = true;
      for (int j = 0; j < 3; ++j) interior = interior && c[j] > 0 && c[j] < n;
      if (interior && jitter > 0)
        for (int j = 0; j < 3; ++j) disp[3 * i + j] = (2.0 * u[j] - 1.0) * jitter * w;
    }
    auto place = [&](int64_t i) {
      for (int j = 0; j < 3; ++j) xyz[3 * i + j] = base[3 * i + j] + disp[3 * i + j];
      if (grade) xyz[3 * i + 0] = xyz[3 * i + 0] * xyz[3 * i + 0];
    };
    for (int64_t i = 0; i < nv; ++i) place(i);
    // validity repair: halve the displacement of every vertex of a
    // non-positive tet until all tets are positive
    for (int pass = 0; pass < 64; ++pass) {
      int64_t bad = 0;
      for (int64_t k = 0; k < nt; ++k) {
        const int32_t* q = tv + 4 * k;
        double d = tet_det(xyz + 3 * q[0], xyz + 3 * q[1], xyz + 3 * q[2], xyz + 3 * q[3]);
        if (!(d > 0)) {
          ++bad;
          for (int j = 0; j < 4; ++j)
            for (int c = 0; c < 3; ++c) disp[3 * q[j] + c] *= 0.5;
        }
      }
      if (!bad) break;
      for (int64_t i = 0; i < nv; ++i) place(i);
      if (pass == 63) return sprg_fail(SPRG_EINVAL, "box: jitter repair did not converge");
    }
  }
  // edges: axis, face diagonals, body diagonals (creation order)
  std::vector<int32_t> axis_id(3 * nv, -1), face_id(3 * nv, -1);
  std::vector<int32_t> body_id((size_t)n * n * n, -1);
  int64_t e = 0;
  for (int64_t i = 0; i < nv; ++i) {
    int64_t c[3] = {i % g, (i / g) % g, i / (g * g)};
    for (int j = 0; j < 3; ++j) {
      if (c[j] == n) continue;
      int64_t o[3] = {c[0], c[1], c[2]};
      o[j] += 1;
      axis_id[3 * i + j] = (int32_t)e;
      if (ev) { ev[2 * e] = (int32_t)i; ev[2 * e + 1] = (int32_t)vid(o[0], o[1], o[2]); }
      ++e;
    }
  }
  for (int64_t i = 0; i < nv; ++i) {
    int64_t c[3] = {i % g, (i / g) % g, i / (g * g)};
    for (int jx = 0; jx < 3; ++jx) {
      int jy = (jx + 1) % 3;
      if (c[jx] == n || c[jy] == n) continue;
      int64_t o[3] = {c[0], c[1], c[2]};
      o[jx] += 1;
      o[jy] += 1;
      face_id[3 * i + jx] = (int32_t)e;
      if (ev) { ev[2 * e] = (int32_t)vid(o[0], o[1], o[2]); ev[2 * e + 1] = (int32_t)i; }
      ++e;
    }
  }
  for (int64_t h = 0; h < (int64_t)n * n * n; ++h) {
    int64_t x = h % n, y = (h / n) % n, z = h / ((int64_t)n * n);
    body_id[h] = (int32_t)e;
    if (ev) { ev[2 * e] = (int32_t)vid(x, y, z); ev[2 * e + 1] = (int32_t)vid(x + 1, y + 1, z + 1); }
    ++e;
  }
  if (e != ne) return sprg_fail(SPRG_EINTERNAL, "box: edge count mismatch");
  // tet -> edge in apf::tet_edge_verts order
  if (te) {
    for (int64_t k = 0; k < nt; ++k) {
      for (int l = 0; l < 6; ++l) {
        int64_t a = tv[4 * k + tev[l][0]], b = tv[4 * k + tev[l][1]];
        int64_t ca[3] = {a % g, (a / g) % g, a / (g * g)};
        int64_t cb[3] = {b % g, (b / g) % g, b / (g * g)};
        int64_t lo[3], d[3], nd = 0;
        for (int j = 0; j < 3; ++j) {
          lo[j] = ca[j] < cb[j] ? ca[j] : cb[j];
          d[j] = ca[j] < cb[j] ? cb[j] - ca[j] : ca[j] - cb[j];
          nd += d[j];
        }
        int64_t li = vid(lo[0], lo[1], lo[2]);
        int32_t id = -1;
        if (nd == 1) {
          for (int j = 0; j < 3; ++j)
            if (d[j]) id = axis_id[3 * li + j];
        } else if (nd == 2) {
          int jx = (d[0] && d[1]) ? 0 : (d[1] && d[2]) ? 1 : 2;
          id = face_id[3 * li + jx];
        } else if (nd == 3) {
          id = body_id[lo[0] + n * (lo[1] + (int64_t)n * lo[2])];
        }
        if (id < 0) return sprg_fail(SPRG_EINTERNAL, "box: tet edge lookup failed");
        te[6 * k + l] = id;
      }
    }
  }
  // Lagrange-2 edge nodes at midpoints of the final vertex positions
  if (order == 2) {
    if (!ev && !te) return sprg_fail(SPRG_EINVAL, "box: order 2 needs edges");
    std::vector<int32_t> ends;
    const int32_t* evp = ev;
    if (!evp) {
      ends.assign(2 * ne, -1);
      for (int64_t k = 0; k < nt; ++k)
        for (int l = 0; l < 6; ++l) {
          int32_t id = te[6 * k + l];
          ends[2 * id] = tv[4 * k + tev[l][0]];
          ends[2 * id + 1] = tv[4 * k + tev[l][1]];
        }
      evp = ends.data();
    }
    for (int64_t k = 0; k < ne; ++k) {
      const double* x0 = xyz + 3 * (int64_t)evp[2 * k];
      const double* x1 = xyz + 3 * (int64_t)evp[2 * k + 1];
      double* out = xyz + 3 * (nv + k);
      for (int j = 0; j < 3; ++j) {
        double c = 0;
        c += x0[j] * 0.5;
        c += x1[j] * 0.5;
        out[j] = c;
      }
    }
  }
  return 0;
}
