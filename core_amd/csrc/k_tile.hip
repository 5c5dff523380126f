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
// Here a block takes a tile: a run of consecutive vertices with the edges
// they own (an edge's tets all contain its lower vertex, so they lie in that
// vertex's patch).  The block loads the union of the tile's vertex patches
// once, with coalesced loads, computes the tets' sample points into LDS, and
// then fits every patch of the tile from LDS:
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

constexpr int XCD_K_TILE = 4;  // xcd_block chunk: neighbouring tiles share tets
constexpr int SORT_CAP = 2048; // candidate tets a tile build sorts (sum of valences)

// ---------------------------------------------------------------------------
// tile construction (once per mesh)

// One block per tile.  The v2t rows of consecutive vertices are contiguous
// in the CSR, so the candidate list of the tile's vertex run [v0, v1)
// (tile_vert[2 tile], tile_vert[2 tile + 1]) is v2t[v2t_off[v0] ..
// v2t_off[v1]).  It
// is sorted in LDS (bitonic network padded with INT32_MAX) and compacted to
// the union.  pass 0 writes the union size to tet_off[tile + 1]; pass 1
// writes the union at tet_off[tile] and, for every patch of the tile, the
// position of each of its tets in the union (status bit 1: a tet of an
// owned edge is missing from the union, i.e. tet->edge and tet->vertex
// disagree).
__global__ __launch_bounds__(256) void k_tile_union(MeshView m, const int32_t* tile_vert,
                                                    const int32_t* own_off, const int32_t* own_e,
                                                    int pass, int32_t* tet_off, int32_t* tets,
                                                    uint16_t* v2t_loc, uint16_t* e2t_loc,
                                                    int32_t* status)
{
  __shared__ int32_t sA[SORT_CAP];
  __shared__ int32_t sScan[256];
  __shared__ int32_t sU[SORT_CAP];
  const int tid = threadIdx.x;
  const int64_t tile = blockIdx.x;
  const int32_t v0 = tile_vert[2 * tile], v1 = tile_vert[2 * tile + 1];
  const int32_t k0 = m.v2t_off[v0], k1 = m.v2t_off[v1];
  const int n = k1 - k0;  // <= SORT_CAP (the host caps the sum of valences)
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
  constexpr int PER = SORT_CAP / 256;
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
  for (int k = k0 + tid; k < k1; k += 256) v2t_loc[k] = (uint16_t)find(m.v2t[k]);
  if (!own_off) return;
  const int32_t e0 = own_off[v0], e1 = own_off[v1];
  for (int i = e0 + tid; i < e1; i += 256) {
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

// Stages tets[0..U) of a tile: ids into sTet, the q sample points of each
// into sX[(u q + l) 3 ..].  A lane per tet; the straight-sided path issues
// the vertex ids of two tets, then their eight vertices' coordinates, before
// it uses any of them, so each dependent load waits once per pair.
template <bool P2MAP>
__device__ __forceinline__ void stage_tets(const MeshView& m, const int32_t* tets, int U, int q,
                                           int32_t* sTet, double* sX)
{
  for (int u = threadIdx.x; u < U; u += 2 * TILE_B) {
    const int ub = u + TILE_B;
    const bool two = ub < U;
    const int32_t ta = tets[u], tb = two ? tets[ub] : ta;
    sTet[u] = ta;
    if (two) sTet[ub] = tb;
    if constexpr (P2MAP) {
      for (int l = 0; l < q; ++l) {
        double x[3];
        sample_x<true>(m, ta, l, x);
        for (int j = 0; j < 3; ++j) sX[3 * (u * q + l) + j] = x[j];
        if (two) {
          sample_x<true>(m, tb, l, x);
          for (int j = 0; j < 3; ++j) sX[3 * (ub * q + l) + j] = x[j];
        }
      }
    } else {
      const int4 va = *reinterpret_cast<const int4*>(m.tv + 4 * (int64_t)ta);
      const int4 vb = *reinterpret_cast<const int4*>(m.tv + 4 * (int64_t)tb);
      double P[2][4][3];
      const int32_t ids[2][4] = {{va.x, va.y, va.z, va.w}, {vb.x, vb.y, vb.z, vb.w}};
#pragma unroll
      for (int k = 0; k < 2; ++k)
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int j = 0; j < 3; ++j) P[k][a][j] = m.xyz[3 * (int64_t)ids[k][a] + j];
      for (int l = 0; l < q; ++l) {
        double xi[3];
        rule_point(m.order, l, xi);
        const double l0 = 1 - xi[0] - xi[1] - xi[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
          if (k == 1 && !two) break;
          const int uu = k == 0 ? u : ub;
#pragma unroll
          for (int j = 0; j < 3; ++j)
            sX[3 * (uu * q + l) + j] =
                fma(xi[2], P[k][3][j], fma(xi[1], P[k][2][j], fma(xi[0], P[k][1][j], l0 * P[k][0][j])));
        }
      }
    }
  }
}

// Lane l's value of x from lane l ^ st (st = 1, 2, 4), within rows of 16:
// DPP quad permutations for 1 and 2, a swizzle for 4.
__device__ __forceinline__ int xor_lane(int x, int st)
{
  if (st == 1) return __builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, false);  // quad_perm 1,0,3,2
  if (st == 2) return __builtin_amdgcn_update_dpp(0, x, 0x4E, 0xF, 0xF, false);  // quad_perm 2,3,0,1
  return __builtin_amdgcn_ds_swizzle(x, 0x101F);                                // xor 4 in 32 lanes
}

__device__ __forceinline__ double xor_lane(double x, int st)
{
  const int lo = xor_lane(__double2loint(x), st), hi = xor_lane(__double2hiint(x), st);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ float xor_lane(float x, int st)
{
  return __int_as_float(xor_lane(__float_as_int(x), st));
}

// Sums v over the lanes of each group (g lanes, aligned); every lane of a
// group ends with the same sum.  gmax: the largest group of the wave
// (wave-uniform), so every lane runs the same exchanges; lanes of smaller
// groups keep their values.
template <class T, int N>
__device__ __forceinline__ void group_sum(T* v, int g, int gmax)
{
  for (int st = 1; st < gmax; st <<= 1) {
    const bool in = g > st;
#pragma unroll
    for (int k = 0; k < N; ++k) {
      const T t = xor_lane(v[k], st);
      v[k] = in ? v[k] + t : v[k];
      // one exchange in flight at a time: otherwise the scheduler hoists all
      // N of them and holds N temporaries next to the N sums
      asm volatile("" : "+v"(v[k]));
    }
  }
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
  double* const sX = smem;                                           // [umax*q][3]
  int32_t* const sTet = reinterpret_cast<int32_t*>(smem + 3 * tv.max_union * q);  // [umax]
  constexpr int NG = NT * (NT + 1) / 2;
  for (int64_t tile = xcd_block(blockIdx.x, gridDim.x, XCD_K_TILE); tile < tv.ntile; tile += gridDim.x) {
    // stage the union's tet ids and sample points, a lane per tet
    const int32_t u0 = tv.tet_off[tile], U = tv.tet_off[tile + 1] - u0;
    stage_tets<P2MAP>(m, tv.tets + u0, U, q, sTet, sX);
    __syncthreads();
    // this lane's slot: patch, group size, rank in the group
    const int32_t s0 = tv.slot_off[tile];
    const bool has = threadIdx.x < tv.slot_off[tile + 1] - s0;
    const uint32_t sv = has ? (uint32_t)tv.slot[s0 + threadIdx.x] : 0u;
    const int64_t p = sv & 0x7ffffffu;
    const int g = has ? 1 << ((sv >> 27) & 3u) : 1;
    const int rank = (int)(sv >> 29);
    const uint64_t b2 = __ballot(g >= 2), b4 = __ballot(g >= 4), b8 = __ballot(g >= 8);
    const int gmax = b8 ? 8 : b4 ? 4 : b2 ? 2 : 1;
    const uint16_t* loc = nullptr;
    int ns = 0;
    if (has) {
      if (p < m.nv) {
        loc = tv.v2t_loc + m.v2t_off[p];
        ns = m.v2t_off[p + 1] - m.v2t_off[p];
      } else {
        loc = tv.e2t_loc + m.e2t_off[p - m.nv];
        ns = m.e2t_off[p - m.nv + 1] - m.e2t_off[p - m.nv];
      }
    }
    // fewer samples than terms (hasEnoughPoints): straight to expansion
    const bool short_set = has && ns * q < NT;
    const bool work = has && !short_set;
    double xn[3] = {0, 0, 0};
    if (work) {
      xn[0] = m.xyz[3 * p];
      xn[1] = m.xyz[3 * p + 1];
      xn[2] = m.xyz[3 * p + 2];
    }
    // Gram matrix of the lane's tets, packed lower triangle
    T G[NG];
#pragma unroll
    for (int k = 0; k < NG; ++k) G[k] = 0;
    if (work)
      for (int s = rank; s < ns; s += g) {
        const double* X = sX + 3 * q * (int)loc[s];
        for (int l = 0; l < q; ++l) {
          const double d[3] = {X[3 * l] - xn[0], X[3 * l + 1] - xn[1], X[3 * l + 2] - xn[2]};
          T f[NT];
          basis<T, NT>(d, f);
#pragma unroll
          for (int i = 0; i < NT; ++i)
#pragma unroll
            for (int j = 0; j <= i; ++j) G[tri(i, j)] = fma(f[i], f[j], G[tri(i, j)]);
        }
      }
    group_sum<T, NG>(G, g, gmax);
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
        const T lkk = dk > 0 ? sqrt(dk) : (T)0;
        if (!(ratio >= tau_rank) || !((double)lkk >= 1.05e-10)) rank_ok = false;
        const T rk = lkk > 0 ? (T)1 / lkk : (T)1;  // the diagonal keeps 1 / L_kk
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
    const int poor = sizeof(T) == 4 ? ACT_FALLBACK : bucket < NBUCKET ? ACT_REFIT : ACT_FALLBACK;
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
          for (int s = rank; s < ns; s += g) {
            const double* X = sX + 3 * q * (int)loc[s];
            for (int l = 0; l < q; ++l) {
              const double d[3] = {X[3 * l] - xn[0], X[3 * l + 1] - xn[1], X[3 * l + 2] - xn[2]};
              double f[NT];
              basis<double, NT>(d, f);
              double t = 0;
#pragma unroll
              for (int k = 0; k < NT; ++k) t = fma(f[k], z[k], t);
#pragma unroll
              for (int k = 0; k < NT; ++k) u[k] = fma(f[k], t, u[k]);
            }
          }
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
    // appends by the group's first lane
    const int app = rank == 0 ? act : ACT_NONE;
    wave_append(ACT_EXACT, app, a.L.count + 2 * NBUCKET, a.L.xq,
                short_set ? (int32_t)p : (int32_t)(p | 0x80000000u));
    wave_append(ACT_FALLBACK, app, a.L.count + 2 * NBUCKET + 1, a.L.fbq, (int32_t)p);
    for (int b = 0; b < NBUCKET; ++b)
      wave_append(b, app == ACT_REFIT ? bucket : -1, a.L.count + b, a.L.items[b], (int32_t)p);
    // weighted sum over the lane's tets, centred on the set's first sample
    const bool solve = act == ACT_SOLVE;
    double v[NC], b0[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) v[c] = 0;
    if (solve) {
      const double* bp = a.ip + (int64_t)sTet[loc[0]] * q * NC;
#pragma unroll
      for (int c = 0; c < NC; ++c) b0[c] = bp[c];
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
      for (int s = rank; s < ns; s += g) {
        const int u = loc[s];
        const double* X = sX + 3 * q * u;
        const double* bt = a.ip + (int64_t)sTet[u] * q * NC;
        if (q == 4 && 4 * NC <= 36) {
          // the tet's whole q x nc block in flight at once (G is dead here)
          double b[4][NC];
#pragma unroll
          for (int l = 0; l < 4; ++l)
#pragma unroll
            for (int c = 0; c < NC; ++c) b[l][c] = bt[l * NC + c];
#pragma unroll
          for (int l = 0; l < 4; ++l) accumulate(X + 3 * l, b[l]);
        } else {
          for (int l = 0; l < q; ++l) {
            double b[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) b[c] = bt[l * NC + c];
            accumulate(X + 3 * l, b);
          }
        }
      }
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
  const unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>(a.shape.tile_blocks, tv.ntile));
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

int launch_tile_build(const MeshView& m, int64_t ntile, const int32_t* tile_vert,
                      const int32_t* own_off, const int32_t* own_e, int pass, int32_t* tet_off,
                      int32_t* tets, uint16_t* v2t_loc, uint16_t* e2t_loc, int32_t* status,
                      hipStream_t s)
{
  if (ntile > 0)
    k_tile_union<<<(unsigned)ntile, 256, 0, s>>>(m, tile_vert, own_off, own_e, pass, tet_off, tets,
                                                  v2t_loc, e2t_loc, status);
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
