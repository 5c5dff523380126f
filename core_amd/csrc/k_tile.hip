// k_tile.hip -- the regular pass of the fast forms over tiles of patches.
//
// The fast forms recover every vertex (and, on a Lagrange-2 mesh, every edge)
// from the normal equations of its patch in a node-centred monomial basis
// (k_fast.hip states the algebra; spr/sprRecoverField.cc:193-353 is the
// reference's fit).  Each tet belongs to the patches of its 4 vertices and 6
// edges, so a thread per patch that gathers its own tets loads every tet's
// vertex ids, coordinates and sample points about ten times, through chains
// of dependent loads (set -> vertex ids -> coordinates).
//
// Here a block takes a tile: a run of patches in vertex order, each vertex's
// patch followed by the edges it owns (an edge's tets all contain its lower
// vertex, so they lie in that vertex's patch), possibly ending inside a
// vertex's edges.  The block loads the union of the tets of the vertices
// the tile touches once, with coalesced loads, computes the tets' sample
// points into LDS, and then fits every patch of the tile from LDS:
//
//   - a patch is fitted by a group of g = 1, 2, 4 or 8 lanes (chosen by its
//     tet count when the tiles were built), lane r of the group taking the
//     tets r, r+g, ...; the group's partial Gram matrices, refinement
//     residuals and weighted sums are summed by a butterfly of lane
//     shuffles, after which every lane of the group holds the same sums (an
//     FP addition is commutative, so a+b and b+a round alike);
//   - the decisions (rank certificate, conditioning, routes to the
//     expansion, the FP64 fallback or the exact refit) are those of the
//     thread-per-patch regular pass, made by the group's rank-0 lane;
//   - the IP values are read from HBM in the weighted sum, a tet's q x nc
//     block contiguous.
//
// The grouping of a patch is fixed per mesh, so results do not depend on
// the grid size.
#include "device.h"
#include "fit_common.h"
#include "spr_math.h"

#include <algorithm>

namespace sprg {
namespace {


// ---------------------------------------------------------------------------
// tile construction (once per mesh)

// One block per tile.  The v2t rows of consecutive vertices are contiguous
// in the CSR, so the candidate list of the tile's rows [cv0, cv1)
// (desc[TILE_DESC tile ..]: cv0, cv1, pv0, pv1, pe0, pe1) is
// v2t[v2t_off[cv0] .. v2t_off[cv1]).  It is sorted in LDS (bitonic network
// padded with INT32_MAX) and compacted to the union.  pass 0 writes the
// union size to tet_off[tile + 1]; pass 1 writes the union at tet_off[tile]
// and, for the tile's patches -- the vertices [pv0, pv1) and the edges
// own_e[pe0, pe1), whose owners' rows are candidates -- the position of
// each of their tets in the union (status bit 1: a tet of an owned edge is
// missing from the union, i.e. tet->edge and tet->vertex disagree).
__global__ __launch_bounds__(256) void k_tile_union(MeshView m, const int32_t* desc, const int32_t* own_e,
                                                    int pass, int32_t* tet_off, int32_t* tets,
                                                    uint16_t* v2t_loc, uint16_t* e2t_loc,
                                                    int32_t* status)
{
  __shared__ int32_t sA[TILE_SORT_CAP];
  __shared__ int32_t sScan[256];
  __shared__ int32_t sU[TILE_SORT_CAP];
  const int tid = threadIdx.x;
  const int64_t tile = blockIdx.x;
  const int32_t* const d = desc + TILE_DESC * tile;
  const int32_t k0 = m.v2t_off[d[0]], k1 = m.v2t_off[d[1]];
  const int n = k1 - k0;  // <= TILE_SORT_CAP (the host caps the sum of valences)
  int P = 1;
  while (P < n) P <<= 1;
  for (int i = tid; i < P; i += 256) sA[i] = i < n ? m.v2t[k0 + i] : INT32_MAX;
  __syncthreads();
  for (int k = 2; k <= P; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = tid; i < P; i += 256) {
        const int l = i ^ j;
        if (l > i) {
          const int32_t x = sA[i], y = sA[l];
          const bool up = (i & k) == 0;
          if ((x > y) == up) {
            sA[i] = y;
            sA[l] = x;
          }
        }
      }
      __syncthreads();
    }
  // first occurrences, scanned in chunks of 8 per thread
  constexpr int PER = TILE_SORT_CAP / 256;
  int flags[PER];
  int cnt = 0;
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int i = tid * PER + j;
    flags[j] = (i < n && (i == 0 || sA[i] != sA[i - 1])) ? 1 : 0;
    cnt += flags[j];
  }
  sScan[tid] = cnt;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const int t = tid >= off ? sScan[tid - off] : 0;
    __syncthreads();
    sScan[tid] += t;
    __syncthreads();
  }
  const int total = sScan[255];
  if (pass == 0) {
    if (tid == 0) tet_off[tile + 1] = total;
    return;
  }
  int pos = sScan[tid] - cnt;
#pragma unroll
  for (int j = 0; j < PER; ++j)
    if (flags[j]) sU[pos++] = sA[tid * PER + j];
  __syncthreads();
  const int32_t u0 = tet_off[tile];
  for (int i = tid; i < total; i += 256) tets[u0 + i] = sU[i];
  auto find = [&](int32_t t) -> int {
    int lo = 0, hi = total - 1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (sU[mid] < t) lo = mid + 1; else hi = mid;
    }
    return (total > 0 && sU[lo] == t) ? lo : -1;
  };
  for (int k = m.v2t_off[d[2]] + tid; k < m.v2t_off[d[3]]; k += 256) v2t_loc[k] = (uint16_t)find(m.v2t[k]);
  if (!own_e) return;
  for (int i = d[4] + tid; i < d[5]; i += 256) {
    const int32_t e = own_e[i];
    for (int32_t k = m.e2t_off[e]; k < m.e2t_off[e + 1]; ++k) {
      const int u = find(m.e2t[k]);
      if (u < 0) atomicOr(status, 1);
      e2t_loc[k] = (uint16_t)(u < 0 ? 0 : u);
    }
  }
}

// ---------------------------------------------------------------------------
// the tiled regular pass

// Sample point l of the mesh-order rule of tet t: the affine map of the
// vertices for straight-sided meshes, the Lagrange-2 map otherwise (the
// same expressions as the thread-per-patch pass).
template <bool P2MAP>
__device__ __forceinline__ void sample_x(const MeshView& m, int32_t t, int l, double* x)
{
  double xi[3];
  rule_point(m.order, l, xi);
  const int64_t tt = t;
  if constexpr (P2MAP) {
    double N[10];
    tet_values(2, xi, N);
    double s[3] = {0, 0, 0};
    for (int k = 0; k < 10; ++k) {
      const int64_t node = k < 4 ? (int64_t)m.tv[4 * tt + k] : m.nv + (int64_t)m.te[6 * tt + k - 4];
      const double* p = m.xyz + 3 * node;
      for (int j = 0; j < 3; ++j) s[j] = fma(N[k], p[j], s[j]);
    }
    x[0] = s[0]; x[1] = s[1]; x[2] = s[2];
  } else {
    const int4 v = *reinterpret_cast<const int4*>(m.tv + 4 * tt);
    const double* p0 = m.xyz + 3 * (int64_t)v.x;
    const double* p1 = m.xyz + 3 * (int64_t)v.y;
    const double* p2 = m.xyz + 3 * (int64_t)v.z;
    const double* p3 = m.xyz + 3 * (int64_t)v.w;
    const double l0 = 1 - xi[0] - xi[1] - xi[2];
    for (int j = 0; j < 3; ++j)
      x[j] = fma(xi[2], p3[j], fma(xi[1], p2[j], fma(xi[0], p1[j], l0 * p0[j])));
  }
}

// The tiles a block fits: the tiles are cut into eight contiguous ranges,
// one per XCD (the dispatcher hands block b to XCD b mod 8), and the blocks
// of an XCD take its range's tiles in turn.  Consecutive tiles hold
// neighbouring vertices, so the tiles an XCD has in flight at once form a
// compact region of the mesh whose tets its L2 keeps for the neighbouring
// tiles that share them.  Every tile has a block when the grid has at least
// eight blocks (the launch guarantees it).
struct TileWalk {
  int64_t first, end, step;
};

__device__ __forceinline__ TileWalk tile_walk(int64_t ntile)
{
  const int64_t g = gridDim.x, b = blockIdx.x, x = b % 8;
  TileWalk w;
  w.first = ntile * x / 8 + b / 8;
  w.end = ntile * (x + 1) / 8;
  w.step = (g - x + 7) / 8;  // the blocks b' < g with b' mod 8 == x
  return w;
}

// What a lane needs to start a tile, loaded while the previous tile is
// fitted: the tile's tet range and lane range, the lane's two union tets and
// their vertex ids (a lane stages tets lane and lane + 64), and the lane's
// slot record.
struct TileNext {
  int32_t u0 = 0, U = 0, s0 = 0, nslot = 0;
  int32_t ta = 0, tb = 0;
  int4 va = {0, 0, 0, 0}, vb = {0, 0, 0, 0};
  int4 sl = {0, 0, 0, 0};
};

// The prefetch in three steps, each placed after enough work for the loads
// of the previous step to have arrived.
__device__ __forceinline__ void next_meta(const TileView& tv, int64_t tile, TileNext& n)
{
  if (tile < tv.ntile) {
    n.u0 = tv.tet_off[tile];
    n.U = tv.tet_off[tile + 1] - n.u0;
    n.s0 = tv.slot_off[tile];
    n.nslot = tv.slot_off[tile + 1] - n.s0;
  } else {
    n.U = 0;
    n.nslot = 0;
  }
}

__device__ __forceinline__ void next_ids(const TileView& tv, TileNext& n)
{
  const int l = threadIdx.x;
  n.ta = l < n.U ? tv.tets[n.u0 + l] : 0;
  n.tb = l + TILE_B < n.U ? tv.tets[n.u0 + l + TILE_B] : 0;
  if (l < n.nslot) n.sl = tv.slot[n.s0 + l];
}

__device__ __forceinline__ void next_verts(const MeshView& m, TileNext& n)
{
  const int l = threadIdx.x;
  if (l < n.U) n.va = *reinterpret_cast<const int4*>(m.tv + 4 * (int64_t)n.ta);
  if (l + TILE_B < n.U) n.vb = *reinterpret_cast<const int4*>(m.tv + 4 * (int64_t)n.tb);
}

// Stages the tile's union into LDS: tet ids into sTet, the q sample points
// of tet u into sX[(u q + l) 3 ..].  A lane per tet (at most 2 TILE_B tets
// per tile: the union cap is below that for TILE_B = 64 lanes is checked by
// the builder).  Straight-sided meshes use the prefetched vertex ids; curved
// Lagrange-2 meshes map through all ten nodes.
template <bool P2MAP>
__device__ __forceinline__ void stage_tets(const MeshView& m, const TileNext& n, int q, int32_t* sTet,
                                           double* sX)
{
  const int l = threadIdx.x;
  if constexpr (P2MAP) {
    for (int k = 0; k < 2; ++k) {
      const int u = l + k * TILE_B;
      if (u >= n.U) break;
      const int32_t t = k == 0 ? n.ta : n.tb;
      sTet[u] = t;
      for (int r = 0; r < q; ++r) {
        double x[3];
        sample_x<true>(m, t, r, x);
        for (int j = 0; j < 3; ++j) sX[3 * (u * q + r) + j] = x[j];
      }
    }
  } else {
    // both tets' vertex coordinates in flight before either is mapped (the
    // vertex ids of a lane without a second tet are an earlier tile's, or
    // zero: valid vertices, so the loads need no guard)
    double P[2][4][3];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int4 v = k == 0 ? n.va : n.vb;
      const int32_t ids[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int j = 0; j < 3; ++j) P[k][a][j] = m.xyz[3 * (int64_t)ids[a] + j];
    }
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int u = l + k * TILE_B;
      if (u >= n.U) break;
      sTet[u] = k == 0 ? n.ta : n.tb;
      for (int r = 0; r < q; ++r) {
        double xi[3];
        rule_point(m.order, r, xi);
        const double l0 = 1 - xi[0] - xi[1] - xi[2];
#pragma unroll
        for (int j = 0; j < 3; ++j)
          sX[3 * (u * q + r) + j] =
              fma(xi[2], P[k][3][j], fma(xi[1], P[k][2][j], fma(xi[0], P[k][1][j], l0 * P[k][0][j])));
      }
    }
  }
}

// Lane l's value of x from lane l ^ ST (ST = 1, 2, 4), within rows of 16:
// DPP quad permutations for 1 and 2, a swizzle for 4.
template <int ST>
__device__ __forceinline__ int xor_lane(int x)
{
  static_assert(ST == 1 || ST == 2 || ST == 4, "lane exchange distance");
  // every lane reads a lane of its own quad, so no lane needs an old value
  if constexpr (ST == 1) return __builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, true);  // quad_perm 1,0,3,2
  if constexpr (ST == 2) return __builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, true);  // quad_perm 2,3,0,1
  return __builtin_amdgcn_ds_swizzle(x, 0x101F);                                          // xor 4 in 32 lanes
}

template <int ST>
__device__ __forceinline__ double xor_lane(double x)
{
  const int lo = xor_lane<ST>(__double2loint(x)), hi = xor_lane<ST>(__double2hiint(x));
  return __hiloint2double(hi, lo);
}

template <int ST>
__device__ __forceinline__ float xor_lane(float x)
{
  return __int_as_float(xor_lane<ST>(__float_as_int(x)));
}

// One butterfly level: lanes of groups wider than ST add their partner's
// value, the others add it times zero.  The partner of a lane outside such a
// group belongs to another group or to no patch, whose sums are finite (or
// the zeros they start from), so the product is +-0 and the value stays;
// inside a group, t * 1 + v rounds once, as v + t does.
template <int ST, class T, int N>
__device__ __forceinline__ void group_step(T* v, int g)
{
  const T w = g > ST ? (T)1 : (T)0;
#pragma unroll
  for (int k = 0; k < N; ++k) v[k] = fma(xor_lane<ST>(v[k]), w, v[k]);
}

// Sums v over the lanes of each group (g lanes, aligned); every lane of a
// group ends with the same sum.  gmax: the largest group of the wave
// (wave-uniform), so every lane runs the same exchanges; lanes of smaller
// groups keep their values.
template <class T, int N>
__device__ __forceinline__ void group_sum(T* v, int g, int gmax)
{
  if (gmax > 1) group_step<1, T, N>(v, g);
  if (gmax > 2) group_step<2, T, N>(v, g);
  if (gmax > 4) group_step<4, T, N>(v, g);
}

// 1 / sqrt(d) for d > 0: the hardware estimate and Newton steps
// r <- r + r (1/2 - (d/2) r^2), two in FP64 (each doubles the correct bits)
// and one in FP32.
__device__ __forceinline__ double rsqrt_nr(double d)
{
  double r = __builtin_amdgcn_rsq(d);
  const double h = 0.5 * d;
#pragma unroll
  for (int i = 0; i < 2; ++i) r = fma(r, fma(-h, r * r, 0.5), r);
  return r;
}

__device__ __forceinline__ float rsqrt_nr(float d)
{
  const float r = __builtin_amdgcn_rsqf(d);
  return fmaf(r, fmaf(-0.5f * d, r * r, 0.5f), r);
}

// body(u) for the lane's tets s = rank, rank + g, ... < ns of its patch set
// (u: the tet's position in the tile's union), the next position loaded
// while the current tet is processed.
template <class F>
__device__ __forceinline__ void for_lane_tets(const uint16_t* loc, int rank, int g, int ns, F body)
{
  if (rank >= ns) return;
  int u = loc[rank];
  for (int s = rank; s < ns; s += g) {
    const int un = loc[s + g < ns ? s + g : ns - 1];
    body(u);
    u = un;
  }
}

// body(l) for the q points of a tet (a rolled loop: unrolled, the points'
// loads would be hoisted next to the 55 Gram accumulators and spill them).
template <class F>
__device__ __forceinline__ void for_points(int q, F body)
{
  for (int l = 0; l < q; ++l) body(l);
}

// Forward and back substitution with the Cholesky factor whose diagonal
// holds the reciprocals of the pivots.
template <class T, int NT>
__device__ __forceinline__ void chol_solve_rcp(const T* L, const T* rhs, T* z)
{
  T y[NT];
#pragma unroll
  for (int k = 0; k < NT; ++k) {
    T s = rhs[k];
#pragma unroll
    for (int j = 0; j < k; ++j) s = fma(-L[tri(k, j)], y[j], s);
    y[k] = s * L[tri(k, k)];
  }
#pragma unroll
  for (int kk = 0; kk < NT; ++kk) {
    const int k = NT - 1 - kk;
    T s = y[k];
#pragma unroll
    for (int j = k + 1; j < NT; ++j) s = fma(-L[tri(j, k)], z[j], s);
    z[k] = s * L[tri(k, k)];
  }
}

template <class T, int NT, int NC, int NREF, bool P2MAP>
__global__ __launch_bounds__(TILE_B) void k_tile(FitArgs a, float tau_rank, float tau_acc)
{
  extern __shared__ double smem[];
  const MeshView& m = a.m;
  const TileView& tv = a.tiles;
  const int q = a.q;
  double* const sX = smem;                                                         // [umax*q][3]
  int32_t* const sTet = reinterpret_cast<int32_t*>(smem + 3 * tv.max_union * q);  // [umax]
  constexpr int NG = NT * (NT + 1) / 2;
  const TileWalk w = tile_walk(tv.ntile);
  TileNext nx;
  next_meta(tv, w.first < w.end ? w.first : tv.ntile, nx);
  next_ids(tv, nx);
  next_verts(m, nx);
  for (int64_t tile = w.first; tile < w.end; tile += w.step) {
    const TileNext cur = nx;
    // this lane's slot: patch, group size, rank in the group, set
    const bool has = (int)threadIdx.x < cur.nslot;
    const uint32_t sv = has ? (uint32_t)cur.sl.x : 0u;
    const int64_t p = sv & 0x7ffffffu;
    const int g = has ? 1 << ((sv >> 27) & 3u) : 1;
    const int rank = (int)(sv >> 29);
    const uint16_t* const loc = tv.loc + (has ? cur.sl.y : 0);
    const int ns = has ? cur.sl.z : 0;
    // fewer samples than terms (hasEnoughPoints): straight to expansion
    const bool short_set = has && ns * q < NT;
    const bool work = has && !short_set;
    // the patch's node, loaded while the tile is staged
    double xn[3] = {0, 0, 0};
    if (work) {
      xn[0] = m.xyz[3 * p];
      xn[1] = m.xyz[3 * p + 1];
      xn[2] = m.xyz[3 * p + 2];
    }
    stage_tets<P2MAP>(m, cur, q, sTet, sX);
    next_meta(tv, tile + w.step < w.end ? tile + w.step : tv.ntile, nx);  // arrives during the Gram sums
    __syncthreads();
    const uint64_t b2 = __ballot(g >= 2), b4 = __ballot(g >= 4), b8 = __ballot(g >= 8);
    const int gmax = b8 ? 8 : b4 ? 4 : b2 ? 2 : 1;
    // Gram matrix of the group's tets, packed lower triangle
    T G[NG];
    if constexpr (NT == 10) {
      // from the moments of the lane's rows, summed over the group
      T M[NMOM];
#pragma unroll
      for (int k = 0; k < NMOM; ++k) M[k] = 0;
      if (work)
        for_lane_tets(loc, rank, g, ns, [&](int u) {
          const double* X = sX + 3 * q * u;
          for_points(q, [&](int l) {
            const double d[3] = {X[3 * l] - xn[0], X[3 * l + 1] - xn[1], X[3 * l + 2] - xn[2]};
            moments_row<T>(d, M);
          });
        });
      next_ids(tv, nx);  // arrives during the reduction and the factorisation
      group_sum<T, NMOM>(M, g, gmax);
      M[0] = (T)(ns * q);
      gram_from_moments<T>(M, G);
    } else {
#pragma unroll
      for (int k = 0; k < NG; ++k) G[k] = 0;
      if (work)
        for_lane_tets(loc, rank, g, ns, [&](int u) {
          const double* X = sX + 3 * q * u;
          for_points(q, [&](int l) {
            const double d[3] = {X[3 * l] - xn[0], X[3 * l + 1] - xn[1], X[3 * l + 2] - xn[2]};
            T f[NT];
            basis<T, NT>(d, f);
#pragma unroll
            for (int i = 0; i < NT; ++i)
#pragma unroll
              for (int j = 0; j <= i; ++j) G[tri(i, j)] = fma(f[i], f[j], G[tri(i, j)]);
          });
        });
      next_ids(tv, nx);  // arrives during the reduction and the factorisation
      group_sum<T, NG>(G, g, gmax);
    }
    // Cholesky in place, with the rank certificate and the conditioning
    bool rank_ok = true;
    float min_ratio = 1.0f;
    if (work) {
#pragma unroll
      for (int k = 0; k < NT; ++k) {
        const T diag = G[tri(k, k)];
        T dk = diag;
#pragma unroll
        for (int j = 0; j < k; ++j) dk = fma(-G[tri(k, j)], G[tri(k, j)], dk);
        const float ratio = (diag > 0 && dk > 0) ? (float)(dk / diag) : 0.0f;
        min_ratio = fminf(min_ratio, ratio);
        // L_kk = sqrt(dk) must clear 1.05e-10: dk >= 1.05e-10^2
        if (!(ratio >= tau_rank) || !((double)dk >= 1.1025e-20)) rank_ok = false;
        const T rk = dk > 0 ? rsqrt_nr(dk) : (T)1;  // the diagonal keeps 1 / L_kk
        G[tri(k, k)] = rk;
#pragma unroll
        for (int i = k + 1; i < NT; ++i) {
          T s = G[tri(i, k)];
#pragma unroll
          for (int j = 0; j < k; ++j) s = fma(-G[tri(i, j)], G[tri(k, j)], s);
          G[tri(i, k)] = s * rk;
          asm volatile("" : "+v"(G[tri(i, k)]));  // keeps the column's temporaries few
        }
      }
    }
    // the regular pass's routes (k_fast.hip, SRC_REGULAR)
    const int bucket = bucket_of(ns * q);
    const int poor = sizeof(T) == 4        ? ACT_FALLBACK
                     : bucket < NBUCKET      ? ACT_REFIT
                     : ns * q <= BIG_MMAX    ? ACT_BIG
                                             : ACT_FALLBACK;
    int act = ACT_NONE;
    if (short_set)
      act = ACT_EXACT;
    else if (work && !rank_ok)
      act = sizeof(T) == 4 ? ACT_FALLBACK : ACT_EXACT;
    else if (work && !(min_ratio >= tau_acc))
      act = poor;
    else if (work)
      act = ACT_SOLVE;
    // z = G^-1 e_0 with NREF FP64 refinement steps (the group's lanes hold
    // identical G and z, and sum their tets' residual contributions)
    double z[NT];
#pragma unroll
    for (int k = 0; k < NT; ++k) z[k] = 0;
    if (act == ACT_SOLVE) {
      T e0[NT], zt[NT];
#pragma unroll
      for (int k = 0; k < NT; ++k) e0[k] = k == 0 ? 1 : 0;
      chol_solve_rcp<T, NT>(G, e0, zt);
#pragma unroll
      for (int k = 0; k < NT; ++k) z[k] = (double)zt[k];
    }
    if (NREF > 0) {
      double last_dz = 0, zmax = 0;
      const bool refine = act == ACT_SOLVE;
      for (int r = 0; r < NREF; ++r) {
        double u[NT];
#pragma unroll
        for (int k = 0; k < NT; ++k) u[k] = 0;
        if (refine)
          for_lane_tets(loc, rank, g, ns, [&](int ut) {
            const double* X = sX + 3 * q * ut;
            for_points(q, [&](int l) {
              const double d[3] = {X[3 * l] - xn[0], X[3 * l + 1] - xn[1], X[3 * l + 2] - xn[2]};
              double f[NT];
              basis<double, NT>(d, f);
              double t = 0;
#pragma unroll
              for (int k = 0; k < NT; ++k) t = fma(f[k], z[k], t);
#pragma unroll
              for (int k = 0; k < NT; ++k) u[k] = fma(f[k], t, u[k]);
            });
          });
        group_sum<double, NT>(u, g, gmax);
        if (refine) {
          T res[NT], dz[NT];
#pragma unroll
          for (int k = 0; k < NT; ++k) res[k] = (T)((k == 0 ? 1.0 : 0.0) - u[k]);
          chol_solve_rcp<T, NT>(G, res, dz);
          last_dz = 0;
          zmax = 0;
#pragma unroll
          for (int k = 0; k < NT; ++k) {
            z[k] += (double)dz[k];
            last_dz = fmax(last_dz, fabs((double)dz[k]));
            zmax = fmax(zmax, fabs(z[k]));
          }
        }
      }
      if (refine && !(last_dz <= 1e-13 * zmax)) act = poor;
    }
    // appends by the group's first lane; one ballot skips all nine for a
    // wave with nothing to append (most tiles)
    const int app = rank == 0 ? act : ACT_NONE;
    if (__ballot(app != ACT_SOLVE && app != ACT_NONE)) {  // wave-uniform
      wave_append(ACT_EXACT, app, a.L.count + 2 * NBUCKET, a.L.xq,
                  short_set ? (int32_t)p : (int32_t)(p | 0x80000000u));
      wave_append(ACT_FALLBACK, app, a.L.count + 2 * NBUCKET + 1, a.L.fbq, (int32_t)p);
      for (int b = 0; b < NBUCKET; ++b)
        wave_append(b, app == ACT_REFIT ? bucket : -1, a.L.count + b, a.L.items[b], (int32_t)p);
      big_append(app == ACT_BIG, a.L, (int32_t)p);
    }
    next_verts(m, nx);  // arrives during the weighted sum
    // weighted sum over the lane's tets, centred on the set's first sample
    const bool solve = act == ACT_SOLVE;
    double v[NC], b0[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) v[c] = 0;
    auto accumulate = [&](const double* X, const double* b) {
      const double d[3] = {X[0] - xn[0], X[1] - xn[1], X[2] - xn[2]};
      double f[NT];
      basis<double, NT>(d, f);
      double w = 0;
#pragma unroll
      for (int k = 0; k < NT; ++k) w = fma(f[k], z[k], w);
#pragma unroll
      for (int c = 0; c < NC; ++c) v[c] = fma(w, b[c] - b0[c], v[c]);
    };
    if (q == 4 && 4 * NC <= 36) {
      // the tet's whole q x nc block in flight at once (G is dead here).
      // The set's first sample b0 is the first point of the first tet of
      // the group's rank-0 lane: every lane loads its first tet's block,
      // and the group takes b0 from that lane (every lane of the wave takes
      // part in the exchange; a group is all solved or all not).  Every lane
      // of a group has a first tet: the tile builder gives a set of ns tets
      // g <= ns lanes (group_log2 in runtime.hip), so rank < ns here.
      double b[4 * NC];
      const int uf = solve ? loc[rank] : 0;
      int un = solve && rank + g < ns ? loc[rank + g] : 0;  // in flight with the first block
      if (solve) {
        const double* bt = a.ip + (int64_t)sTet[uf] * 4 * NC;
#pragma unroll
        for (int i = 0; i < 4 * NC; ++i) b[i] = bt[i];
      } else {
#pragma unroll
        for (int i = 0; i < 4 * NC; ++i) b[i] = 0;
      }
      const int src = (int)threadIdx.x - rank;
#pragma unroll
      for (int c = 0; c < NC; ++c) b0[c] = __shfl(b[c], src);
      if (solve) {
#pragma unroll
        for (int l = 0; l < 4; ++l) accumulate(sX + 12 * uf + 3 * l, b + l * NC);
        for (int s = rank + g; s < ns; s += g) {
          const int u = un;
          un = loc[s + g < ns ? s + g : ns - 1];
          const double* X = sX + 12 * u;
          const double* bt = a.ip + (int64_t)sTet[u] * 4 * NC;
#pragma unroll
          for (int i = 0; i < 4 * NC; ++i) b[i] = bt[i];
#pragma unroll
          for (int l = 0; l < 4; ++l) accumulate(X + 3 * l, b + l * NC);
        }
      }
    } else if (solve) {
      const double* bp = a.ip + (int64_t)sTet[loc[0]] * q * NC;
#pragma unroll
      for (int c = 0; c < NC; ++c) b0[c] = bp[c];
      for_lane_tets(loc, rank, g, ns, [&](int u) {
        const double* X = sX + 3 * q * u;
        const double* bt = a.ip + (int64_t)sTet[u] * q * NC;
        for (int l = 0; l < q; ++l) {
          double b[NC];
#pragma unroll
          for (int c = 0; c < NC; ++c) b[c] = bt[l * NC + c];
          accumulate(X + 3 * l, b);
        }
      });
    }
    group_sum<double, NC>(v, g, gmax);
    if (solve && rank == 0) {
      double* out = a.out + p * NC;
#pragma unroll
      for (int c = 0; c < NC; ++c) out[c] = b0[c] + v[c];
    }
    __syncthreads();  // the next tile overwrites the staged samples
  }
}

template <class T, int NT, int NREF, bool P2>
int launch_tile_nc(const FitArgs& a, float tau_rank, float tau_acc, hipStream_t s)
{
  const TileView& tv = a.tiles;
  const size_t lds = sizeof(double) * 3 * (size_t)tv.max_union * a.q + sizeof(int32_t) * tv.max_union;
  // at least 8 blocks: tile_walk gives each XCD's range to its own blocks
  const unsigned blocks = (unsigned)std::max<int64_t>(8, std::min<int64_t>(a.shape.tile_blocks, tv.ntile));
  switch (a.nc) {
    case 1: k_tile<T, NT, 1, NREF, P2><<<blocks, TILE_B, lds, s>>>(a, tau_rank, tau_acc); break;
    case 3: k_tile<T, NT, 3, NREF, P2><<<blocks, TILE_B, lds, s>>>(a, tau_rank, tau_acc); break;
    case 9: k_tile<T, NT, 9, NREF, P2><<<blocks, TILE_B, lds, s>>>(a, tau_rank, tau_acc); break;
    case 27: k_tile<T, NT, 27, NREF, P2><<<blocks, TILE_B, lds, s>>>(a, tau_rank, tau_acc); break;
    default: return sprg_fail(SPRG_EUNSUPPORTED, "tiled recovery: %d components", a.nc);
  }
  return hipGetLastError() == hipSuccess ? 0 : sprg_fail(SPRG_EHIP, "tiled fit launch failed");
}

template <class T, int NT, int NREF>
int launch_tile_p2(const FitArgs& a, float tau_rank, float tau_acc, hipStream_t s)
{
  if (a.m.order == 2 && !a.m.straight) return launch_tile_nc<T, NT, NREF, true>(a, tau_rank, tau_acc, s);
  return launch_tile_nc<T, NT, NREF, false>(a, tau_rank, tau_acc, s);
}

}  // namespace

int launch_tile_build(const MeshView& m, int64_t ntile, const int32_t* desc, const int32_t* own_e, int pass,
                      int32_t* tet_off, int32_t* tets, uint16_t* v2t_loc, uint16_t* e2t_loc, int32_t* status,
                      hipStream_t s)
{
  if (ntile > 0)
    k_tile_union<<<(unsigned)ntile, 256, 0, s>>>(m, desc, own_e, pass, tet_off, tets, v2t_loc, e2t_loc, status);
  return hipGetLastError() == hipSuccess ? 0 : sprg_fail(SPRG_EHIP, "tile build launch failed");
}

// The fast forms' regular pass over the tiles (the same tolerances as the
// thread-per-patch pass: FP32 Gram with three FP64 refinement steps, or an
// FP64 Gram).
int launch_tile_fit(const FitArgs& a, int nterms, int form, hipStream_t s)
{
  if (a.tiles.ntile == 0) return 0;
  if (form == SPRG_FP32R)
    return nterms == 4 ? launch_tile_p2<float, 4, 3>(a, 1e-4f, 1e-3f, s)
                       : launch_tile_p2<float, 10, 3>(a, 1e-4f, 1e-3f, s);
  return nterms == 4 ? launch_tile_p2<double, 4, 0>(a, 1e-6f, 1e-6f, s)
                     : launch_tile_p2<double, 10, 0>(a, 1e-6f, 1e-6f, s);
}

}  // namespace sprg
