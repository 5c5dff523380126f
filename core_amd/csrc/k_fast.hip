// k_fast.hip -- the fast forms of the patch recovery: one thread per patch,
// normal equations in a node-centred monomial basis, and the recovered value
// as a weighted sum of the patch's samples.
//
// Recovered value at the patch node x_n (spr/sprRecoverField.cc:304-353):
//   v_c = phi(x_n)^T (A^T A)^-1 A^T b_c.
// With monomials of d = x - x_n, phi(x_n) = e_0, so with z = G^-1 e_0
// (G = A^T A) every component is v_c = sum_i w_i b_ic, w_i = phi_i . z,
// and since sum_i w_i = 1 it is evaluated as b_0c + sum_i w_i (b_ic - b_0c),
// whose rounding scales with the variation of b over the patch rather
// than with |b|.  Centring is a triangular change of the graded monomial
// basis, so the fit (and the diagonal of its Cholesky factor, which equals
// the reference's Householder R_kk on raw coordinates in exact arithmetic)
// is unchanged.
//
//   T = float : G and its Cholesky factor in FP32; NREF (3) residual
//               refinement steps r = e_0 - A^T (A z) in FP64;
//   T = double: G and the factor in FP64 (NREF = 0), or the FP64 fallback
//               for ill-conditioned patches (NREF = 2).
//
// Rank: the reference rejects a patch when a Householder column norm falls
// below 1e-10 (mth/mthQR.cc:31).  The Cholesky pivot L_kk is that norm; a
// patch is certified full rank here only if every pivot clears 1e-10 with a
// margin and was computed without heavy cancellation (pivot / diagonal >=
// tau_rank).  Anything else goes to the exact FP64 path (k_expand tests the
// initial set with the reference's arithmetic before growing it).  Certified
// patches whose smallest pivot ratio is below tau_acc go to the FP64
// fallback.  Three passes over the rows: Gram, refinement, weighted sum;
// sample positions are recomputed in each (affine in the four vertices for a
// straight-sided mesh, the Lagrange-2 map otherwise).
#include "device.h"
#include "fit_common.h"
#include "spr_math.h"

#include <algorithm>

namespace sprg {
namespace {

constexpr int NQ_MAX = 4;  // points per element the fast kernels handle
constexpr int XCD_K_FIT = 16;  // xcd_block chunk of the fit pass

template <bool P2MAP>
struct Geo {
  double x[P2MAP ? 10 : 4][3];
};

template <bool P2MAP>
__device__ __forceinline__ void load_geo(const MeshView& m, int32_t t, bool p2map, Geo<P2MAP>& g)
{
  const int64_t tt = t;
  for (int a = 0; a < 4; ++a) {
    const double* p = m.xyz + 3 * (int64_t)m.tv[4 * tt + a];
    g.x[a][0] = p[0]; g.x[a][1] = p[1]; g.x[a][2] = p[2];
  }
  if (P2MAP)
    for (int e = 0; e < 6; ++e) {
      const double* p = m.xyz + 3 * (m.nv + (int64_t)m.te[6 * tt + e]);
      g.x[4 + e][0] = p[0]; g.x[4 + e][1] = p[1]; g.x[4 + e][2] = p[2];
    }
}

// Calls f(t, geometry of t) for the tets of a set in order.  Straight-sided
// meshes run a three-stage load pipeline: while tet s is processed, the
// coordinates of tet s+1, the vertex ids of tet s+2 and the id of tet s+3
// are in flight, so each dependent load (set -> vertex ids -> coordinates)
// has a whole tet's arithmetic to arrive.  Ids past the end repeat a valid
// tet (loaded, never used).
template <bool P2MAP, class F>
__device__ __forceinline__ void for_each_tet(const MeshView& m, const int32_t* set, int ns, F&& f)
{
  if constexpr (P2MAP) {
    for (int s = 0; s < ns; ++s) {
      Geo<true> g;
      load_geo(m, set[s], true, g);
      f(set[s], g);
    }
  } else {
    auto ids = [&](int32_t t) { return *reinterpret_cast<const int4*>(m.tv + 4 * (int64_t)t); };
    auto coords = [&](const int4& v, Geo<false>& g) {
      const int32_t vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const double* p = m.xyz + 3 * (int64_t)vv[a];
        g.x[a][0] = p[0];
        g.x[a][1] = p[1];
        g.x[a][2] = p[2];
      }
    };
    int32_t t0 = set[0];
    int32_t t1 = ns > 1 ? set[1] : t0;
    int32_t t2 = ns > 2 ? set[2] : t1;
    int4 v1 = ids(t1);
    Geo<false> g0;
    coords(ids(t0), g0);
    for (int s = 0; s < ns; ++s) {
      Geo<false> g1;
      coords(v1, g1);
      const int4 v2 = ids(t2);
      const int32_t t3 = s + 3 < ns ? set[s + 3] : t2;
      f(t0, g0);
      t0 = t1;
      t1 = t2;
      t2 = t3;
      v1 = v2;
      g0 = g1;
    }
  }
}

// Sample point l of the mesh-order rule, minus the node position.
template <bool P2MAP>
__device__ __forceinline__ void sample_d(const MeshView& m, const Geo<P2MAP>& g, bool p2map,
                                         int l, const double* xn, double* d)
{
  double xi[3];
  rule_point(m.order, l, xi);
  double x[3];
  if constexpr (P2MAP) {
    double N[10];
    tet_values(2, xi, N);
    for (int j = 0; j < 3; ++j) {
      double s = 0;
      for (int k = 0; k < 10; ++k) s = fma(N[k], g.x[k][j], s);
      x[j] = s;
    }
  } else {
    const double l0 = 1 - xi[0] - xi[1] - xi[2];
    for (int j = 0; j < 3; ++j)
      x[j] = fma(xi[2], g.x[3][j], fma(xi[1], g.x[2][j], fma(xi[0], g.x[1][j], l0 * g.x[0][j])));
  }
  d[0] = x[0] - xn[0];
  d[1] = x[1] - xn[1];
  d[2] = x[2] - xn[2];
}

// SRC_PRE_FALLBACK: the FP32 form's FP64 pass over its fallback queue before
// the expansion (it may still send a set to the rank test); SRC_FALLBACK then
// walks only the entries queued after it.
// SRC_LEFTOVER: the initial sets of the patches outside every tile of the
// tiled pass (k_tile.hip), with the regular pass's decisions.
enum { SRC_REGULAR = 0, SRC_EXPANDED = 1, SRC_FALLBACK = 2, SRC_PRE_FALLBACK = 3, SRC_LEFTOVER = 4 };


// src SRC_REGULAR walks every patch id (no list); the others walk a list.
template <class T, int NT, int NC, int NREF, bool P2MAP>
__global__ __launch_bounds__(256) void k_fast(FitArgs a, int src, float tau_rank, float tau_acc)
{
  const MeshView& m = a.m;
  const int q = a.q;
  constexpr bool p2map = P2MAP;
  const bool regular = src == SRC_REGULAR || src == SRC_LEFTOVER;
  const int32_t* list = src == SRC_EXPANDED ? a.L.xitems[0] : src == SRC_LEFTOVER ? a.tiles.leftover : a.L.fbq;
  const int64_t count = src == SRC_REGULAR    ? a.npatch
                        : src == SRC_LEFTOVER ? a.tiles.nleftover
                        : src == SRC_EXPANDED ? a.L.count[NBUCKET]
                                              : a.L.count[2 * NBUCKET + 1];
  const int64_t first = src == SRC_FALLBACK ? a.L.status[2] : 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  // The trip count is the same for every lane of a block, so the list
  // appends below run on whole waves; lanes past the end carry ACT_NONE.
  for (int64_t base = first + xcd_block(blockIdx.x, gridDim.x, XCD_K_FIT) * blockDim.x; base < count;
       base += stride) {
    const int64_t it = base + threadIdx.x;
    const bool valid = it < count;
    const int32_t entry = !valid ? 0 : src == SRC_REGULAR ? (int32_t)it : list[it];
    const bool expanded =
        src == SRC_EXPANDED || ((src == SRC_FALLBACK || src == SRC_PRE_FALLBACK) && entry < 0);
    const int32_t id = entry & 0x7fffffff;
    int64_t p = 0;
    const int32_t* set = nullptr;
    int ns = 0;
    if (valid) {
      if (expanded) {
        p = a.L.x_patch[id];
        set = a.L.x_set + a.L.x_off[id];
        ns = a.L.x_len[id];
      } else {
        p = id;
        if (p < m.nv) {
          set = m.v2t + m.v2t_off[p];
          ns = m.v2t_off[p + 1] - m.v2t_off[p];
        } else {
          set = m.e2t + m.e2t_off[p - m.nv];
          ns = m.e2t_off[p - m.nv + 1] - m.e2t_off[p - m.nv];
        }
      }
    }
    // fewer samples than terms (hasEnoughPoints): straight to expansion
    const bool short_set = valid && regular && ns * q < NT;
    const bool work = valid && !short_set;
    constexpr int NG = NT * (NT + 1) / 2;
    T G[NG];
#pragma unroll
    for (int k = 0; k < NG; ++k) G[k] = 0;
    bool rank_ok = true;
    float min_ratio = 1.0f;
    double xn[3] = {0, 0, 0};
    if (work) {
      xn[0] = m.xyz[3 * p];
      xn[1] = m.xyz[3 * p + 1];
      xn[2] = m.xyz[3 * p + 2];
      // pass 1: Gram matrix, packed lower triangle
      for_each_tet<P2MAP>(m, set, ns, [&](int32_t, const Geo<P2MAP>& g) {
        for (int l = 0; l < q; ++l) {
          double d[3];
          sample_d(m, g, p2map, l, xn, d);
          T f[NT];
          basis<T, NT>(d, f);
#pragma unroll
          for (int i = 0; i < NT; ++i)
#pragma unroll
            for (int j = 0; j <= i; ++j) G[tri(i, j)] = fma(f[i], f[j], G[tri(i, j)]);
        }
      });
      // Cholesky in place, with the rank certificate and the conditioning
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
        G[tri(k, k)] = lkk > 0 ? lkk : (T)1;
#pragma unroll
        for (int i = k + 1; i < NT; ++i) {
          T s = G[tri(i, k)];
#pragma unroll
          for (int j = 0; j < k; ++j) s = fma(-G[tri(i, j)], G[tri(k, j)], s);
          G[tri(i, k)] = s / G[tri(k, k)];
        }
      }
    }
    // Short sets and uncertified ranks: the exact path decides (bit 31: test
    // the initial set before growing it).  Below tau_acc the FP32 regular
    // pass hands the patch to the FP64 fallback; an FP64 pass hands it to
    // the exact kernel (normal equations square the condition number; the
    // reference's Householder arithmetic reproduces its own answer), when the
    // set fits the exact kernel's rows.
    const int bucket = bucket_of(ns * q);
    const int poor = (regular && sizeof(T) == 4) ? ACT_FALLBACK
                     : bucket < NBUCKET                     ? ACT_REFIT
                     : ns * q <= BIG_MMAX                   ? ACT_BIG
                     : src != SRC_FALLBACK && src != SRC_PRE_FALLBACK ? ACT_FALLBACK
                     // more rows than the exact kernels take: the refined
                     // FP64 solve, unless hopeless
                     : min_ratio >= 1e-13f ? ACT_SOLVE : ACT_ILLCOND;
    int act = ACT_NONE;
    if (short_set)
      act = ACT_EXACT;
    else if (work && (regular || src == SRC_PRE_FALLBACK) && !rank_ok)
      // the FP32 pass leaves the rank to its FP64 pass, which certifies
      // most sets the FP32 certificate could not
      act = (regular && sizeof(T) == 4) ? ACT_FALLBACK : ACT_EXACT;
    else if (work && !(min_ratio >= tau_acc))
      act = poor;
    else if (work)
      act = ACT_SOLVE;
    // z = G^-1 e_0, refined against the FP64 residual; a refinement that has
    // not settled (last correction above 1e-13 of |z|) counts as poorly
    // conditioned
    double z[NT];
    if (act == ACT_SOLVE) {
      T e0[NT], zt[NT];
#pragma unroll
      for (int k = 0; k < NT; ++k) e0[k] = k == 0 ? 1 : 0;
      chol_solve<T, NT>(G, e0, zt);
#pragma unroll
      for (int k = 0; k < NT; ++k) z[k] = (double)zt[k];
      double last_dz = 0, zmax = 0;
      for (int r = 0; r < NREF; ++r) {
        double u[NT];
#pragma unroll
        for (int k = 0; k < NT; ++k) u[k] = 0;
        for_each_tet<P2MAP>(m, set, ns, [&](int32_t, const Geo<P2MAP>& g) {
          for (int l = 0; l < q; ++l) {
            double d[3], f[NT];
            sample_d(m, g, p2map, l, xn, d);
            basis<double, NT>(d, f);
            double t = 0;
#pragma unroll
            for (int k = 0; k < NT; ++k) t = fma(f[k], z[k], t);
#pragma unroll
            for (int k = 0; k < NT; ++k) u[k] = fma(f[k], t, u[k]);
          }
        });
        T res[NT], dz[NT];
#pragma unroll
        for (int k = 0; k < NT; ++k) res[k] = (T)((k == 0 ? 1.0 : 0.0) - u[k]);
        chol_solve<T, NT>(G, res, dz);
        last_dz = 0;
        zmax = 0;
#pragma unroll
        for (int k = 0; k < NT; ++k) {
          z[k] += (double)dz[k];
          last_dz = fmax(last_dz, fabs((double)dz[k]));
          zmax = fmax(zmax, fabs(z[k]));
        }
      }
      if (NREF > 0 && !(last_dz <= 1e-13 * zmax) && poor != ACT_SOLVE) act = poor;
    }
    wave_append(ACT_EXACT, act, a.L.count + 2 * NBUCKET, a.L.xq,
                short_set ? (int32_t)p : (int32_t)(p | 0x80000000u));
    wave_append(ACT_FALLBACK, act, a.L.count + 2 * NBUCKET + 1, a.L.fbq,
                expanded ? (int32_t)(id | 0x80000000u) : (int32_t)p);
    for (int b = 0; b < NBUCKET; ++b)
      wave_append(b, act == ACT_REFIT ? bucket : -1, a.L.count + b, a.L.items[b],
                  expanded ? (int32_t)(id | 0x80000000u) : (int32_t)p);
    big_append(act == ACT_BIG, a.L, expanded ? (int32_t)(id | 0x80000000u) : (int32_t)p);
    if (act == ACT_ILLCOND) atomicOr(a.L.status, ST_ILLCOND);
    if (act != ACT_SOLVE) continue;
    // weighted sum, centred on the first sample
    double v[NC], b0[NC];
    {
      const double* bp = a.ip + (int64_t)set[0] * q * NC;
#pragma unroll
      for (int c = 0; c < NC; ++c) { b0[c] = bp[c]; v[c] = 0; }
    }
    for_each_tet<P2MAP>(m, set, ns, [&](int32_t t, const Geo<P2MAP>& g) {
      for (int l = 0; l < q; ++l) {
        const double* bp = a.ip + ((int64_t)t * q + l) * NC;
        double b[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) b[c] = bp[c];  // issued before the weight
        double d[3], f[NT];
        sample_d(m, g, p2map, l, xn, d);
        basis<double, NT>(d, f);
        double w = 0;
#pragma unroll
        for (int k = 0; k < NT; ++k) w = fma(f[k], z[k], w);
#pragma unroll
        for (int c = 0; c < NC; ++c) v[c] = fma(w, b[c] - b0[c], v[c]);
      }
    });
    double* out = a.out + p * NC;
#pragma unroll
    for (int c = 0; c < NC; ++c) out[c] = b0[c] + v[c];
  }
}

template <class T, int NT, int NREF, bool P2>
int launch_nc2(const FitArgs& a, int src, float tau_rank, float tau_acc, hipStream_t s)
{
  const unsigned threads = (unsigned)a.shape.fast_threads;
  const int64_t work = src == SRC_REGULAR    ? a.npatch
                       : src == SRC_LEFTOVER ? a.tiles.nleftover + 1
                                             : a.npatch / 8 + 1;
  const unsigned blocks =
      (unsigned)std::max<int64_t>(1, std::min<int64_t>(a.shape.fast_blocks, (work + threads - 1) / threads));
  switch (a.nc) {
    case 1: k_fast<T, NT, 1, NREF, P2><<<blocks, threads, 0, s>>>(a, src, tau_rank, tau_acc); break;
    case 3: k_fast<T, NT, 3, NREF, P2><<<blocks, threads, 0, s>>>(a, src, tau_rank, tau_acc); break;
    case 9: k_fast<T, NT, 9, NREF, P2><<<blocks, threads, 0, s>>>(a, src, tau_rank, tau_acc); break;
    case 27: k_fast<T, NT, 27, NREF, P2><<<blocks, threads, 0, s>>>(a, src, tau_rank, tau_acc); break;
    default: return sprg_fail(SPRG_EUNSUPPORTED, "fast recovery: %d components", a.nc);
  }
  return hipGetLastError() == hipSuccess ? 0 : sprg_fail(SPRG_EHIP, "fast fit launch failed");
}

template <class T, int NT, int NREF>
int launch_nc(const FitArgs& a, int src, float tau_rank, float tau_acc, hipStream_t s)
{
  if (a.m.order == 2 && !a.m.straight)
    return launch_nc2<T, NT, NREF, true>(a, src, tau_rank, tau_acc, s);
  return launch_nc2<T, NT, NREF, false>(a, src, tau_rank, tau_acc, s);
}

template <class T, int NREF>
int launch_t(const FitArgs& a, int nterms, int src, float tau_rank, float tau_acc, hipStream_t s)
{
  return nterms == 4 ? launch_nc<T, 4, NREF>(a, src, tau_rank, tau_acc, s)
                     : launch_nc<T, 10, NREF>(a, src, tau_rank, tau_acc, s);
}

}  // namespace

// form SPRG_FP32R: FP32 Gram + 3 FP64 refinement steps; SPRG_FP64F: FP64
// Gram, no refinement.  src: 0 regular, 1 expanded,
// 2 the FP64 fallback queue (FP64 Gram + 2 refinement steps).
int launch_fast_fit(const FitArgs& a, int nterms, int form, int src, hipStream_t s)
{
  if (a.q > NQ_MAX) return sprg_fail(SPRG_EUNSUPPORTED, "fast recovery: %d points per element", a.q);
  // FP64 passes refit sets whose smallest pivot ratio is below 1e-6 with the
  // exact kernel.  Expanded sets (often barely full rank) always take the
  // FP64 pass.
  if (src == 2) return launch_t<double, 2>(a, nterms, src, 0.0f, 1e-6f, s);
  if (src == 3) return launch_t<double, 2>(a, nterms, src, 1e-6f, 1e-6f, s);
  if (src == 1) return launch_t<double, 0>(a, nterms, src, 0.0f, 1e-6f, s);
  // FP32: three FP64 refinement steps; each contracts the error by about
  // kappa(G) * 2^-24, and the size field needs the recovered field to ~1e-12
  if (form == SPRG_FP32R) return launch_t<float, 3>(a, nterms, src, 1e-4f, 1e-3f, s);
  // a pivot that keeps less than 1e-6 of its diagonal is not a rank
  // certificate in FP64 (its rounding can exceed 1e-11 of the diagonal after
  // earlier cancellation): such sets take the reference's rank test
  return launch_t<double, 0>(a, nterms, src, 1e-6f, 1e-6f, s);
}

}  // namespace sprg
