// k_exact.hip -- the FP64 patch recovery of spr::recoverField, reproducing the
// reference's arithmetic operation for operation.
//
// The reference fits each patch with mth::decomposeQR / solveFromQR
// (mth/mthQR.cc:19-163): Householder reflectors from R with sequential
// column sums, an EXPLICIT m x m Q formed by reflect_rows, y = Q^T b as a
// sequential dot product, back substitution, and evaluation of the monomials
// at the entity's node (spr/sprRecoverField.cc:283-353).  Because the size
// field downstream amplifies a 1-ulp change of the recovered field by ~1e4
// or more, this form keeps every rounding of that sequence: sums in the same
// order, no fused multiply-add (the library is built with -ffp-contract=off),
// correctly rounded sqrt and division.
//
// Work mapping (one 64-lane wavefront per block):
//  - a patch is owned by a group of LANES lanes (16, 32 or 64); a wave runs
//    64/LANES patches in lock step;
//  - the R recurrence keeps R (m x NT) and the reflectors V (NT x m) in LDS;
//    its sequential sums run on one lane per group, its updates on all;
//  - Q is never stored: every row of Q evolves independently under the
//    reflectors, so lane j keeps row j in registers (MMAX doubles, static
//    indexing) and only its first NT entries are kept (all that Q^T b needs
//    for the back substitution);
//  - rows of the patch matrix are recomputed from the mesh: sample points
//    are the mesh-order rule points mapped by the coordinate field.
#include "device.h"
#include "fit_common.h"
#include "spr_math.h"

namespace sprg {
namespace {

__device__ __forceinline__ double sgn(double x) { return (x < 0) ? -1 : 1; }

__device__ __forceinline__ int64_t node_of(const MeshView& m, int32_t t, int n)
{
  return n < 4 ? (int64_t)m.tv[4 * (int64_t)t + n]
               : m.nv + (int64_t)m.te[6 * (int64_t)t + (n - 4)];
}

// getSamplePoints (spr/sprRecoverField.cc:193-211): the rule of the MESH
// order, point l (the N1 rule ignores l), mapped by the coordinate field
// (apf/apfElement.cc:106-114 summation order).  The mesh order is fixed at
// compile time (a P1 mesh fits 4 terms, a P2 mesh 10), so the rule point and
// shape values stay in registers.
template <int ORDER>
__device__ __forceinline__ void sample_point(const MeshView& m, int32_t t, int l, double* x)
{
  constexpr int NN = ORDER == 1 ? 4 : 10;
  double xi[3], N[NN];
  if (ORDER == 1) {
    xi[0] = 0.25; xi[1] = 0.25; xi[2] = 0.25;
  } else {  // rule_point(2, l)
    const double a = 0.138196601125011, b = 0.585410196624969;
    xi[0] = l == 1 ? b : a;
    xi[1] = l == 2 ? b : a;
    xi[2] = l == 3 ? b : a;
  }
  tet_values(ORDER, xi, N);
  x[0] = 0; x[1] = 0; x[2] = 0;
#pragma unroll
  for (int ni = 0; ni < NN; ++ni) {
    const double* p = m.xyz + 3 * node_of(m, t, ni);
    x[0] += p[0] * N[ni];
    x[1] += p[1] * N[ni];
    x[2] += p[2] * N[ni];
  }
}

// Rows i = gl, gl+LANES, ... of the monomial matrix A of the tet set S.
template <int NT, int LANES>
__device__ __forceinline__ void build_rows(const MeshView& m, int q, const int32_t* S, int nrows,
                           double* R, int gl)
{
  for (int i = gl; i < nrows; i += LANES) {
    double x[3], t[10];
    sample_point<NT == 4 ? 1 : 2>(m, S[i / q], i % q, x);
    poly_terms(NT == 4 ? 1 : 2, x, t);
    for (int j = 0; j < NT; ++j) R[i * NT + j] = t[j];
  }
}

// s + prod(i0) + prod(i0 + 1) + ... + prod(i1 - 1) (s = 0 unless given),
// added one by one in that order (the reference's accumulation loops).  Full blocks of 16 form
// their products ahead of the additions, so that the loads behind them are
// in flight together instead of each waiting in the chain of dependent
// adds; the tail goes four, then one, at a time.  A chain usually runs on
// one lane of the group, but every instruction of it costs the whole wave,
// so the blocks carry no per-element guards.
template <class F>
__device__ __forceinline__ double seq_sum(int i0, int i1, F prod, double s = 0)
{
  int b = i0;
  for (; b + 16 <= i1; b += 16) {
    double p[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) p[u] = prod(b + u);
#pragma unroll
    for (int u = 0; u < 16; ++u) s += p[u];
  }
  for (; b + 4 <= i1; b += 4) {
    double p[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) p[u] = prod(b + u);
#pragma unroll
    for (int u = 0; u < 4; ++u) s += p[u];
  }
  for (; b < i1; ++b) s += prod(b);
  return s;
}

// Two independent sums of the form of seq_sum over the same range, their
// additions interleaved so that each chain's latency hides the other's.
template <class F0, class F1>
__device__ __forceinline__ void seq_sum2(int i0, int i1, F0 prod0, F1 prod1, double& s0, double& s1)
{
  s0 = 0;
  s1 = 0;
  int b = i0;
  for (; b + 8 <= i1; b += 8) {
    double p0[8], p1[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      p0[u] = prod0(b + u);
      p1[u] = prod1(b + u);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      s0 += p0[u];
      s1 += p1[u];
    }
  }
  for (; b < i1; ++b) {
    s0 += prod0(b);
    s1 += prod1(b);
  }
}

// Householder R recurrence of mth::decomposeQR restricted to what the solve
// uses (get_reflector + reflect_columns on columns >= k, mth/mthQR.cc:19-63).
// R is [m][NT] row-major, V is [NT][MS] (reflector k, zero outside k..m-1).
// Returns false as soon as a column norm falls below 1e-10 (the reference's
// rank test, mth/mthQR.cc:31).  Every lane of the block must call it.
template <int NT, int LANES, int MS>
__device__ __forceinline__ bool r_recurrence(double* R, double* V, double* sDot, double* sNorm,
                             int m, int gl, bool act)
{
  bool ok = act;
  for (int k = 0; k < NT; ++k) {
    // get_reflector's two norms on one lane: cnorm over column k, then
    // rnorm over v = column k with v_k += sign(R_kk) cnorm, summed from
    // v_k^2 (0 + v_k^2 is v_k^2) on down the column, as the reference
    // sums the copied vector
    if (ok && gl == 0) {
      const double cn = sqrt(seq_sum(k, m, [&](int i) { return R[i * NT + k] * R[i * NT + k]; }));
      sNorm[0] = cn;
      if (!(cn < 1e-10)) {
        const double vk = R[k * NT + k] + sgn(R[k * NT + k]) * cn;
        sNorm[2] = vk;
        sNorm[1] = sqrt(seq_sum(k + 1, m, [&](int i) { return R[i * NT + k] * R[i * NT + k]; }, vk * vk));
      }
    }
    __syncthreads();
    const double cn = ok ? sNorm[0] : 0;
    if (ok && cn < 1e-10) ok = false;
    if (ok) {
      // the normalised reflector, zero outside rows k..m-1
      const double rn = sNorm[1], vk = sNorm[2];
      for (int i = gl; i < MS; i += LANES)
        V[k * MS + i] = (i >= k && i < m) ? (i == k ? vk : R[i * NT + k]) / rn : 0.0;
    }
    __syncthreads();
    if (ok && gl >= k && gl < NT) {
      const double* v = V + k * MS;
      sDot[gl] = seq_sum(k, m, [&](int i) { return R[i * NT + gl] * v[i]; });
    }
    __syncthreads();
    if (ok) {
      // reflect_columns' update a(i,j) -= 2 dot_j v(i) for j >= k, with the
      // columns at compile-time positions (the loads of a row in flight
      // together) and 2 dot_j read once per reflector
      double d2[NT];
#pragma unroll
      for (int j = 0; j < NT; ++j) d2[j] = 2 * sDot[j];
      for (int i = k + gl; i < m; i += LANES) {
        const double vi = V[k * MS + i];
        double* Ri = R + i * NT;
#pragma unroll
        for (int j = 0; j < NT; ++j)
          if (j >= k) Ri[j] -= d2[j] * vi;
      }
    }
    __syncthreads();
  }
  return ok;
}

__device__ __forceinline__ void patch_set(const FitArgs& a, int64_t p,
                                          const int32_t** set, int* ns)
{
  const MeshView& m = a.m;
  if (p < m.nv) {
    *set = m.v2t + m.v2t_off[p];
    *ns = m.v2t_off[p + 1] - m.v2t_off[p];
  } else {
    int64_t e = p - m.nv;
    *set = m.e2t + m.e2t_off[e];
    *ns = m.e2t_off[e + 1] - m.e2t_off[e];
  }
}

// Sorts patches into row-count buckets; patches with fewer rows than terms
// go straight to the expansion queue (hasEnoughPoints, :355-363).  With
// a.fast set, every regular patch goes to list 0.
__global__ __launch_bounds__(256) void k_classify(FitArgs a, int nterms)
{
  int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  // -1 none, 0..NBUCKET-1 bucket, NBUCKET expansion queue, NBUCKET + 1 the
  // big list (the rank is decided there)
  int target = -1;
  if (p < a.npatch) {
    const int32_t* set;
    int ns;
    patch_set(a, p, &set, &ns);
    int m = ns * a.q;
    if (m < nterms) {
      target = NBUCKET;
    } else if (a.fast) {
      target = 0;
    } else {
      target = bucket_of(m);
      if (target >= NBUCKET) {
        target = m <= BIG_MMAX ? NBUCKET + 1 : -1;
        if (target < 0) atomicOr(a.L.status, ST_TOO_LARGE);
      }
    }
  }
  for (int b = 0; b < NBUCKET; ++b)
    wave_append(b, target, a.L.count + b, a.L.items[b], (int32_t)p);
  wave_append(NBUCKET, target, a.L.count + 2 * NBUCKET, a.L.xq, (int32_t)p);
  big_append(target == NBUCKET + 1, a.L, (int32_t)(p | 0x40000000));
}

// MODE: FIT_REGULAR walks items[bucket] (patch ids), FIT_EXPANDED walks
// xitems[bucket] (expanded-set ids), FIT_MIXED walks items[bucket] holding
// either (bit 31 set: an expanded-set id) -- the fast forms' refits.
enum { FIT_REGULAR = 0, FIT_EXPANDED = 1, FIT_MIXED = 2 };

// Waves per SIMD the fit of a bucket is compiled for: the kernel is bound by
// the latency of its sequential sums, so the buckets that hold most patches
// ask for two resident waves (the LDS of the 16- and 32-row buckets allows
// no more); the Q row of MMAX doubles per lane sets the floor of the
// register budget.
// (Only targets that compile without scratch spills are requested.)
constexpr int exact_waves(int nt, int mmax)
{
  return nt == 4 ? (mmax == 32 ? 2 : 1) : (mmax <= 32 || mmax == 96 ? 2 : 1);
}

template <int NT, int MMAX, int LANES, int MODE>
__global__ __launch_bounds__(64, exact_waves(NT, MMAX)) void k_exact_fit(FitArgs a, int bucket)
{
  constexpr bool EXP = MODE == FIT_EXPANDED;
  constexpr int G = 64 / LANES;
  __shared__ double sR[G][MMAX * NT];
  __shared__ double sV[G][NT * MMAX];
  __shared__ double sRt[G][NT * NT];
  __shared__ double sY[G][27 * NT];
  __shared__ double sDot[G][NT];
  __shared__ double sNorm[G][3];
  __shared__ int32_t sS[G][MMAX];
  const int g = threadIdx.x / LANES, gl = threadIdx.x % LANES;
  const MeshView& m = a.m;
  const int q = a.q, nc = a.nc;
  const int32_t* list = EXP ? a.L.xitems[bucket] : a.L.items[bucket];
  const int nitems = a.L.count[(EXP ? NBUCKET : 0) + bucket];
  for (int base = blockIdx.x * G; base < nitems; base += gridDim.x * G) {
    const int item = base + g;
    const bool act = item < nitems;
    int64_t p = 0;
    int ns = 0;
    const int32_t* set = nullptr;
    if (act) {
      const int32_t entry = list[item];
      if (EXP || (MODE == FIT_MIXED && entry < 0)) {
        const int32_t it = entry & 0x7fffffff;
        p = a.L.x_patch[it];
        set = a.L.x_set + a.L.x_off[it];
        ns = a.L.x_len[it];
      } else {
        p = entry;
        patch_set(a, p, &set, &ns);
      }
    }
    const int mrows = ns * q;
    for (int i = gl; i < ns; i += LANES) sS[g][i] = set[i];
    __syncthreads();
    build_rows<NT, LANES>(m, q, sS[g], mrows, sR[g], gl);
    __syncthreads();
    bool ok = r_recurrence<NT, LANES, MMAX>(sR[g], sV[g], sDot[g], sNorm[g], mrows, gl, act);
    if (act && !ok && gl == 0) {
      if (MODE != FIT_REGULAR) {  // sets whose rank was already decided
        atomicOr(a.L.status, ST_INTERNAL);
      } else {
        int k = atomicAdd(a.L.count + 2 * NBUCKET, 1);
        a.L.xq[k] = (int32_t)p;
      }
    }
    // keep the top NT x NT of R for the back substitution, reuse sR for Q
    for (int i = gl; i < NT * NT; i += LANES) sRt[g][i] = sR[g][i];
    __syncthreads();
    // Q rows: q_j = e_j, then reflect_rows for k = 0..NT-1 (mth/mthQR.cc:65-80)
#pragma unroll 1
    for (int rb = 0; rb < MMAX; rb += LANES) {
      const int j = rb + gl;
      double qr[MMAX];
#pragma unroll
      for (int l = 0; l < MMAX; ++l) qr[l] = (l == j) ? 1.0 : 0.0;
      // V is zero outside rows k..m-1, so both loops can run over all MMAX
      // entries: adding +-0 to the running dot (which starts at +0) and
      // subtracting +-0 from q leave every value as the reference's loops
      // from l = k produce it, and the indexing stays static.
#pragma unroll 1
      for (int k = 0; k < NT; ++k) {
        const double* v = sV[g] + k * MMAX;
        double dot = 0;
#pragma unroll
        for (int l = 0; l < MMAX; ++l) dot += qr[l] * v[l];
        const double t2 = 2 * dot;
        // read v again from LDS rather than keeping MMAX more registers live
        asm volatile("" ::: "memory");
#pragma unroll
        for (int l = 0; l < MMAX; ++l) qr[l] -= t2 * v[l];
      }
      if (ok && j < mrows)
#pragma unroll
        for (int i = 0; i < NT; ++i) sR[g][j * NT + i] = qr[i];
    }
    __syncthreads();
    // stage b = IP values of the patch rows in sV when it fits
    const bool staged = nc <= NT;
    if (ok && staged)
      // a row per lane: its nc values are contiguous in the IP field (row
      // r is point r mod q of set tet r / q)
      for (int r = gl; r < mrows; r += LANES) {
        const int s = q == 4 ? r >> 2 : r / q, l = q == 4 ? r & 3 : r % q;
        const double* src = a.ip + ((int64_t)sS[g][s] * q + l) * nc;
        for (int c = 0; c < nc; ++c) sV[g][r * nc + c] = src[c];
      }
    __syncthreads();
    // y = transpose(Q) b, sequential over rows (mth/mth_def.h:221-232)
    if (ok)
      // two (component, coefficient) sums per lane, w and w + LANES
      for (int w = gl; w < NT * nc; w += 2 * LANES) {
        const bool two = w + LANES < NT * nc;
        const int i0 = w % NT, c0 = w / NT;
        const int w1 = two ? w + LANES : w;
        const int i1 = w1 % NT, c1 = w1 / NT;
        // b from LDS or from the IP field, the branch outside the sums (a
        // select of the two addresses would make every load a generic one)
        double y0, y1;
        if (staged) {
          auto prod = [&](int i, int c, int r) { return sR[g][r * NT + i] * sV[g][r * nc + c]; };
          seq_sum2(0, mrows, [&](int r) { return prod(i0, c0, r); }, [&](int r) { return prod(i1, c1, r); }, y0, y1);
        } else {
          auto prod = [&](int i, int c, int r) {
            return sR[g][r * NT + i] * a.ip[((int64_t)sS[g][r / q] * q + r % q) * nc + c];
          };
          seq_sum2(0, mrows, [&](int r) { return prod(i0, c0, r); }, [&](int r) { return prod(i1, c1, r); }, y0, y1);
        }
        sY[g][c0 * NT + i0] = y0;
        if (two) sY[g][c1 * NT + i1] = y1;
      }
    __syncthreads();
    // backsubUT (mth/mthQR.cc:131-148) and evalPolynomial at the node
    if (ok) {
      double node[3], t[10];
      for (int j = 0; j < 3; ++j) node[j] = m.xyz[3 * p + j];
      poly_terms(NT == 4 ? 1 : 2, node, t);
      for (int c = gl; c < nc; c += LANES) {
        double x[NT];
#pragma unroll
        for (int ii = 0; ii < NT; ++ii) {
          const int i = NT - ii - 1;
          x[i] = sY[g][c * NT + i];
#pragma unroll
          for (int jj = i + 1; jj < NT; ++jj) x[i] -= sRt[g][i * NT + jj] * x[jj];
          x[i] /= sRt[g][i * NT + i];
        }
        double v = 0;
#pragma unroll
        for (int i = 0; i < NT; ++i) v += x[i] * t[i];
        a.out[p * nc + c] = v;
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------
// Patch expansion (expandAsNecessary, spr/sprRecoverField.cc:165-188,
// :365-386).  One wave per queued patch.  The set after shared_dim s is
// every tet that shares at least s+1 vertices with a tet of the set the
// round started from (a face, an edge or a vertex of a conforming tet
// mesh), which includes that set; after each step the exact rank test runs.

constexpr int XMAXR = 192;   // rows of a patch the expansion can test
constexpr int XSET = 192;    // tets in a set
constexpr int HSIZE = 1024;  // LDS hash of candidate tets (at most 319 entries)

__device__ __forceinline__ int shared_verts(const int32_t* a, const int32_t* b)
{
  int s = 0;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) s += (a[i] == b[j]);
  return s;
}

__device__ __forceinline__ double readlane_d(double v, int lane)
{
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ void term_factors(int k, int* a, int* b)
{
  // 1, x, y, z, xy, yz, zx, xx, yy, zz
  *a = k < 4 ? k : k < 7 ? k - 3 : k - 6;
  *b = k < 4 ? 0 : k == 4 ? 2 : k == 5 ? 3 : k == 6 ? 1 : k - 6;
}

// Rank screen of a candidate set, one wave: the FP64 Gram matrix of the
// monomials centred at the patch node and its Cholesky pivots.  The rows'
// centred sample points go to LDS (sD, 3 doubles per row); lane l < NT(NT+1)/2
// owns Gram entry l of the packed lower triangle, sums it over the rows and
// keeps it through a right-looking factorisation whose columns travel by
// lane shuffles (the fused multiply-adds of each entry run in the same order
// as a left-looking Cholesky's).  Centring is a unit-triangular change of the
// graded monomial basis, so in exact arithmetic pivot k equals the
// reference's Householder column norm R_kk (mth/mthQR.cc:31).  Returns true
// only when every pivot clears the reference's 1e-10 by 5% and keeps at least
// 1e-6 of its diagonal; then the reference's test passes too.  (The rounding
// of a pivot grows with the cancellation before it: on an exactly singular
// 10x10 P2 set whose previous pivot kept 1.6e-8 of its diagonal, the last
// pivot came out at +-5e-11 of its diagonal, so a bound near 1e-11 does not
// separate rank-deficient sets from full ones.)  A false return decides
// nothing: the exact test runs.
// Every lane of the block (one wave) must call it.
template <int NT>
__device__ bool certify_full_rank(const MeshView& m, int q, const int32_t* S, int nrows,
                                  int64_t p, int lane, double* sD)
{
  constexpr int NG = NT * (NT + 1) / 2;
  const double xn[3] = {m.xyz[3 * p], m.xyz[3 * p + 1], m.xyz[3 * p + 2]};
  for (int i = lane; i < nrows; i += 64) {
    double x[3];
    sample_point<NT == 4 ? 1 : 2>(m, S[i / q], i % q, x);
    for (int j = 0; j < 3; ++j) sD[3 * i + j] = x[j] - xn[j];
  }
  __syncthreads();
  const bool own = lane < NG;
  int r = 0;
  while (own && (r + 1) * (r + 2) / 2 <= lane) ++r;
  const int c = own ? lane - r * (r + 1) / 2 : 0;
  int ra, rb, ca, cb;
  term_factors(r, &ra, &rb);
  term_factors(c, &ca, &cb);
  // this lane's two monomials as v[a] * v[b] over v = (1, x, y, z) (the
  // products poly_terms forms; a unit factor is exact), each factor a
  // one-hot combination of v (the zero weights add +-0, so a factor is
  // exactly its component), the rows four at a time
  const double wa[4] = {ra == 0 ? 1.0 : 0.0, ra == 1 ? 1.0 : 0.0, ra == 2 ? 1.0 : 0.0, ra == 3 ? 1.0 : 0.0};
  const double wb[4] = {rb == 0 ? 1.0 : 0.0, rb == 1 ? 1.0 : 0.0, rb == 2 ? 1.0 : 0.0, rb == 3 ? 1.0 : 0.0};
  const double wc[4] = {ca == 0 ? 1.0 : 0.0, ca == 1 ? 1.0 : 0.0, ca == 2 ? 1.0 : 0.0, ca == 3 ? 1.0 : 0.0};
  const double wd[4] = {cb == 0 ? 1.0 : 0.0, cb == 1 ? 1.0 : 0.0, cb == 2 ? 1.0 : 0.0, cb == 3 ? 1.0 : 0.0};
  auto pick = [](const double* w, const double* d) { return fma(w[3], d[2], fma(w[2], d[1], fma(w[1], d[0], w[0]))); };
  double g = 0;
  if (own) {
    int i = 0;
    for (; i + 4 <= nrows; i += 4) {
      double ta[4], tb[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const double d[3] = {sD[3 * (i + u)], sD[3 * (i + u) + 1], sD[3 * (i + u) + 2]};
        ta[u] = pick(wa, d) * pick(wb, d);
        tb[u] = pick(wc, d) * pick(wd, d);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) g = fma(ta[u], tb[u], g);
    }
    for (; i < nrows; ++i) {
      const double d[3] = {sD[3 * i], sD[3 * i + 1], sD[3 * i + 2]};
      g = fma(pick(wa, d) * pick(wb, d), pick(wc, d) * pick(wd, d), g);
    }
  }
  __syncthreads();  // sD is free again
  const double diag_own = g;
  bool rank_ok = true;
  for (int k = 0; k < NT; ++k) {
    const int kk = k * (k + 1) / 2 + k;
    // the pivot's lane is the same for the whole wave: a lane read, not a
    // shuffle through LDS
    const double dk = readlane_d(g, kk);
    const double diag = readlane_d(diag_own, kk);
    const double lkk = dk > 0 ? sqrt(dk) : 0.0;
    if (!(diag > 0 && dk >= 1e-6 * diag) || !(lkk >= 1.05e-10)) rank_ok = false;
    const double piv = lkk > 0 ? lkk : 1.0;
    if (own && r > k && c == k) g = g / piv;  // L[r][k]
    const double lr = __shfl(g, own && r > k ? r * (r + 1) / 2 + k : 0);
    const double lc = __shfl(g, own && c > k ? c * (c + 1) / 2 + k : 0);
    if (own && c > k) g = fma(-lr, lc, g);
  }
  return rank_ok;
}

// EXACT = false: the screen alone, with no QR workspace in LDS (several
// times the occupancy); a patch whose growing set the screen cannot certify,
// or that must test its initial set first, is deferred whole to xq_exact.
// EXACT = true: walks xq_exact with the screen backed by the reference's
// test.  Each patch's expansion depends only on the mesh, so deferring it
// changes nothing but where it runs.
template <int NT, bool EXACT>
__global__ __launch_bounds__(64) void k_expand(FitArgs a)
{
  __shared__ double sR[EXACT ? XMAXR * NT : 1];
  __shared__ double sV[EXACT ? NT * XMAXR : 1];
  __shared__ double sDot[NT];
  __shared__ double sNorm[3];
  // the candidate hash during enumeration, the screen's sample points after
  // it (barriers separate the two uses)
  __shared__ double sD[3 * XMAXR];
  static_assert(HSIZE * sizeof(int32_t) <= sizeof(double) * 3 * XMAXR, "hash must fit the sample buffer");
  int32_t* const sHash = reinterpret_cast<int32_t*>(sD);
  __shared__ int32_t sCur[256];
  __shared__ int32_t sOld[XSET];
  __shared__ int32_t sTmp[XSET];
  __shared__ int sN, sOverflow;
  const int lane = threadIdx.x;
  const MeshView& m = a.m;
  const int q = a.q;
  int32_t* const nexact = a.L.status + 1;
  const int32_t* const queue = EXACT ? a.L.xq_exact : a.L.xq;
  const int nq = EXACT ? *nexact : a.L.count[2 * NBUCKET];
  // expanded sets are stored in chunks of x_set that each block reserves
  // with one atomic (a per-patch atomic on one address serialises the
  // launch); a chunk is at most 1024 entries and a quarter of x_set over the
  // grid
  __shared__ int64_t sBase, sLeft;
  const int64_t share = a.L.x_cap / (4 * (int64_t)gridDim.x);
  const int64_t chunk_size = share > 1024 ? 1024 : share < XSET ? XSET : share;
  if (lane == 0) { sBase = 0; sLeft = 0; }
  for (int k = blockIdx.x; k < nq; k += gridDim.x) {
    // the item id of a queued patch is its position in xq; the light pass
    // marks it empty first (x_len 0) and whichever pass finds the set fills it
    const int32_t it = EXACT ? queue[k] : k;
    const int32_t entry = a.L.xq[it];
    if (!EXACT && lane == 0) a.L.x_len[it] = 0;
    const int64_t p = entry & 0x7fffffff;
    const bool test_first = entry < 0;
    const int32_t* set0;
    int ns0;
    patch_set(a, p, &set0, &ns0);
    if (lane == 0) { sN = ns0; sOverflow = ns0 > XSET; }
    __syncthreads();
    if (sOverflow) {
      if (lane == 0) atomicOr(a.L.status, ST_TOO_LARGE);
      __syncthreads();
      continue;
    }
    if (!EXACT && test_first && ns0 * q > XMAXR) {
      // above the expansion's rows: the big kernel decides the rank of the
      // initial set (and fits it); this item keeps no set
      if (lane == 0) {
        if (ns0 * q <= BIG_MMAX) {
          a.L.big[atomicAdd(a.L.status + 3, 1)] = (int32_t)(p | 0x40000000);
        } else {
          atomicOr(a.L.status, ST_TOO_LARGE);
        }
      }
      __syncthreads();
      continue;
    }
    if (!EXACT && test_first && ns0 * q >= NT) {
      if (lane == 0) a.L.xq_exact[atomicAdd(nexact, 1)] = it;
      __syncthreads();
      continue;
    }
    for (int i = lane; i < ns0; i += 64) sCur[i] = set0[i];
    __syncthreads();
    bool done = false, failed = false, deferred = false;
    if (EXACT && test_first && ns0 * q >= NT) {
      // a fast form could not certify the rank: run the reference's test
      const int mrows = ns0 * q;
      if (mrows > XMAXR) {
        failed = true;
        if (lane == 0) sOverflow = 1;
      } else {
        build_rows<NT, 64>(m, q, sCur, mrows, sR, lane);
        __syncthreads();
        done = r_recurrence<NT, 64, XMAXR>(sR, sV, sDot, sNorm, mrows, lane, true);
      }
    }
    // every round grows the set (or fails) and a set holds at most XSET
    // tets (or fails), so the loop ends within XSET + 1 rounds
    for (int round = 0; round <= XSET && !done && !failed && !deferred; ++round) {
      const int nold = sN;
      for (int i = lane; i < nold; i += 64) sOld[i] = sCur[i];
      __syncthreads();
      for (int sd = 2; sd >= 0 && !done; --sd) {
        for (int i = lane; i < HSIZE; i += 64) sHash[i] = -1;
        if (lane == 0) sN = 0;
        __syncthreads();
        // items: (tet of the old set, one of its vertices); with P >= the
        // item count (a power of two, at most 64) each item gets R = 64 / P
        // lanes, which take its neighbour list's chunks of 8 in turn (the
        // same candidate pairs as one lane per item, walked in parallel)
        const int nitem = nold * 4;
        int P = 64;
        while (P > 1 && P / 2 >= nitem) P /= 2;
        const int R = 64 / P, sub = lane / P;
        for (int w = lane % P; w < nitem; w += P) {
          const int32_t t = sOld[w >> 2];
          const int32_t* tv = m.tv + 4 * (int64_t)t;
          const int32_t v = tv[w & 3];
          const int32_t k0 = m.v2t_off[v], k1 = m.v2t_off[v + 1];
          // neighbour ids and their vertices eight at a time, so the
          // dependent loads overlap
          for (int32_t k = k0 + 8 * sub; k < k1; k += 8 * R) {
            int32_t u[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) u[j] = (k + j < k1) ? m.v2t[k + j] : -1;
            int sh[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) sh[j] = u[j] >= 0 ? shared_verts(m.tv + 4 * (int64_t)u[j], tv) : 0;
            for (int j = 0; j < 8; ++j) {
              if (u[j] < 0 || sh[j] < sd + 1) continue;
              // past 256 distinct tets the set is too large anyway;
              // stopping here keeps the hash far from full so every probe
              // ends
              if (__atomic_load_n(&sN, __ATOMIC_RELAXED) >= 256) break;
              uint32_t h = ((uint32_t)u[j] * 2654435761u) & (HSIZE - 1);
              for (;;) {
                int32_t prev = atomicCAS(&sHash[h], -1, u[j]);
                if (prev == -1) {
                  int pos = atomicAdd(&sN, 1);
                  if (pos < 256) sCur[pos] = u[j];
                  break;
                }
                if (prev == u[j]) break;
                h = (h + 1) & (HSIZE - 1);
              }
            }
          }
        }
        __syncthreads();
        const int ns = sN;
        if (ns > XSET) {
          if (lane == 0) sOverflow = 1;
          __syncthreads();
          failed = true;
          break;
        }
        // sort sCur[0..ns) ascending by rank: the ids are distinct, so the
        // position of each is the number of smaller ids (ns <= XSET here)
        for (int i = lane; i < ns; i += 64) {
          const int32_t v = sCur[i];
          int r = 0;
          for (int j = 0; j < ns; ++j) r += sCur[j] < v ? 1 : 0;
          sTmp[r] = v;
        }
        __syncthreads();
        for (int i = lane; i < ns; i += 64) sCur[i] = sTmp[i];
        __syncthreads();
        const int mrows = ns * q;
        bool ok = false;
        if (mrows >= NT && mrows <= XMAXR) {
          // a certified full rank needs no exact test; anything else does
          ok = certify_full_rank<NT>(m, q, sCur, mrows, p, lane, sD);
          if (!ok && !EXACT) {
            if (lane == 0) a.L.xq_exact[atomicAdd(nexact, 1)] = it;
            deferred = true;
            break;
          }
          if (!ok) {
            build_rows<NT, 64>(m, q, sCur, mrows, sR, lane);
            __syncthreads();
            ok = r_recurrence<NT, 64, XMAXR>(sR, sV, sDot, sNorm, mrows, lane, true);
          }
        } else if (mrows > XMAXR) {
          if (lane == 0) sOverflow = 1;
          __syncthreads();
          failed = true;
          break;
        }
        if (ok) done = true;
      }
      if (!done && !failed && !deferred && sN <= nold) failed = true;
    }
    if (failed && lane == 0)
      atomicOr(a.L.status, sOverflow ? ST_TOO_LARGE : ST_PATCH_FAIL);
    // not reachable by the bound above; a patch left without a set must not
    // leave its recovered value unwritten silently
    if (!done && !failed && !deferred && lane == 0) atomicOr(a.L.status, ST_INTERNAL);
    if (done) {
      __shared__ int64_t sOff;
      const int ns = sN;
      if (lane == 0) {
        if (sLeft < ns) {
          const int64_t grab = ns > chunk_size ? (int64_t)ns : chunk_size;
          sBase = (int64_t)atomicAdd((unsigned long long*)a.L.x_cursor, (unsigned long long)grab);
          sLeft = grab;
        }
        sOff = sBase;
        sBase += ns;
        sLeft -= ns;
      }
      __syncthreads();
      if (sOff + ns > a.L.x_cap) {
        if (lane == 0) atomicOr(a.L.status, ST_XSET_OVERFLOW);
      } else {
        for (int i = lane; i < ns; i += 64) a.L.x_set[sOff + i] = sCur[i];
        if (lane == 0) {
          a.L.x_patch[it] = (int32_t)p;
          a.L.x_off[it] = sOff;
          a.L.x_len[it] = ns;
        }
      }
    }
    __syncthreads();
  }
}

// The expanded sets found (x_len > 0) into the lists the fits walk: by row
// bucket in the FP64 form, one list in the fast forms.  Wave-aggregated
// appends; also publishes the item count (one id per queued patch).
__global__ __launch_bounds__(256) void k_bucket_items(FitArgs a)
{
  const int nq = a.L.count[2 * NBUCKET];
  if (blockIdx.x == 0 && threadIdx.x == 0) a.L.x_cursor[1] = nq;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  // the trip count is the same for every lane of a block (convergent appends)
  for (int64_t base = blockIdx.x * (int64_t)blockDim.x; base < nq; base += stride) {
    const int64_t it = base + threadIdx.x;
    int target = -1;
    if (it < nq) {
      const int len = a.L.x_len[it];
      if (len > 0) {
        target = a.fast ? 0 : bucket_of(len * a.q);
        if (target >= NBUCKET) {
          atomicOr(a.L.status, ST_TOO_LARGE);
          target = -1;
        }
      }
    }
    for (int b = 0; b < NBUCKET; ++b)
      wave_append(b, target, a.L.count + NBUCKET + b, a.L.xitems[b], (int32_t)it);
  }
}

// ------------------------------------------------------------------------
// Sets of more than 192 rows (up to BIG_MMAX): one wave per set, the same
// arithmetic as k_exact_fit (mth/mthQR.cc:19-80, 132-163).  R and, later,
// the first NT columns of Q live in global workspace, the reflectors in LDS;
// lane j forms Q's row j (rows j, j+64, ...) in its own column of the
// workspace (entry l of the row at [l][lane], so that the wave's accesses
// are contiguous), with the reference's loops from k to m-1.

constexpr size_t BIG_WS_DOUBLES = (size_t)BIG_MMAX * 10 + (size_t)64 * BIG_MMAX;

template <int NT>
__global__ __launch_bounds__(64) void k_exact_big(FitArgs a, double* ws)
{
  __shared__ double sV[NT * BIG_MMAX];
  __shared__ double sRt[NT * NT];
  __shared__ double sY[27 * NT];
  __shared__ double sDot[NT];
  __shared__ double sNorm[3];
  __shared__ int32_t sS[BIG_MMAX];
  const int lane = threadIdx.x;
  const MeshView& m = a.m;
  const int q = a.q, nc = a.nc;
  double* const R = ws + blockIdx.x * BIG_WS_DOUBLES;  // [m][NT], then Q's first NT columns
  double* const Qw = R + (size_t)BIG_MMAX * 10;        // [BIG_MMAX][64]
  const int n = a.L.status[3];
  for (int k = blockIdx.x; k < n; k += gridDim.x) {
    const int32_t entry = a.L.big[k];
    const bool expanded = entry < 0, test = (entry & 0x40000000) != 0;
    const int32_t id = entry & 0x07ffffff;
    int64_t p;
    const int32_t* set;
    int ns;
    if (expanded) {
      p = a.L.x_patch[id];
      set = a.L.x_set + a.L.x_off[id];
      ns = a.L.x_len[id];
    } else {
      p = id;
      patch_set(a, p, &set, &ns);
    }
    const int mrows = ns * q;  // <= BIG_MMAX (checked where the entry was made)
    for (int i = lane; i < ns; i += 64) sS[i] = set[i];
    __syncthreads();
    build_rows<NT, 64>(m, q, sS, mrows, R, lane);
    __syncthreads();
    const bool ok = r_recurrence<NT, 64, BIG_MMAX>(R, sV, sDot, sNorm, mrows, lane, true);
    if (!ok) {
      // a set this large that is not of full rank would have to grow
      // beyond what the expansion holds; a certified one is inconsistent
      if (lane == 0) atomicOr(a.L.status, test ? ST_TOO_LARGE : ST_INTERNAL);
      __syncthreads();
      continue;
    }
    for (int i = lane; i < NT * NT; i += 64) sRt[i] = R[i];
    __syncthreads();
    // Q rows: q_j = e_j, then reflect_rows for k = 0..NT-1 (mth/mthQR.cc:65-80)
    for (int j = lane; j < mrows; j += 64) {
      for (int l = 0; l < mrows; ++l) Qw[(size_t)l * 64 + lane] = (l == j) ? 1.0 : 0.0;
      for (int r = 0; r < NT; ++r) {
        const double* v = sV + r * BIG_MMAX;
        double dot = 0;
        for (int l = r; l < mrows; ++l) dot += Qw[(size_t)l * 64 + lane] * v[l];
        const double t2 = 2 * dot;
        for (int l = r; l < mrows; ++l) Qw[(size_t)l * 64 + lane] -= t2 * v[l];
      }
      for (int i = 0; i < NT; ++i) R[j * NT + i] = Qw[(size_t)i * 64 + lane];
    }
    __syncthreads();
    // y = transpose(Q) b, sequential over rows (mth/mth_def.h:221-232)
    for (int w = lane; w < NT * nc; w += 64) {
      const int i = w % NT, c = w / NT;
      double y = 0;
      for (int r = 0; r < mrows; ++r) y += R[r * NT + i] * a.ip[((int64_t)sS[r / q] * q + r % q) * nc + c];
      sY[c * NT + i] = y;
    }
    __syncthreads();
    // backsubUT (mth/mthQR.cc:131-148) and evalPolynomial at the node
    double node[3], t[10];
    for (int j = 0; j < 3; ++j) node[j] = m.xyz[3 * p + j];
    poly_terms(NT == 4 ? 1 : 2, node, t);
    for (int c = lane; c < nc; c += 64) {
      double x[NT];
#pragma unroll
      for (int ii = 0; ii < NT; ++ii) {
        const int i = NT - ii - 1;
        x[i] = sY[c * NT + i];
#pragma unroll
        for (int jj = i + 1; jj < NT; ++jj) x[i] -= sRt[i * NT + jj] * x[jj];
        x[i] /= sRt[i * NT + i];
      }
      double v = 0;
#pragma unroll
      for (int i = 0; i < NT; ++i) v += x[i] * t[i];
      a.out[p * nc + c] = v;
    }
    __syncthreads();
  }
}

inline unsigned grid_for(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

template <int NT, int EXP>
int fit_all(const FitArgs& a, hipStream_t s)
{
  // Persistent grids: each block loops over its share of the bucket's
  // device-side item count.
  const unsigned blocks = (unsigned)a.shape.exact_blocks;
  const unsigned quarter = blocks / 4 > 0 ? blocks / 4 : 1;  // the 192-row bucket is rare
  k_exact_fit<NT, 16, 16, EXP><<<blocks, 64, 0, s>>>(a, 0);
  k_exact_fit<NT, 32, 32, EXP><<<blocks, 64, 0, s>>>(a, 1);
  k_exact_fit<NT, 64, 64, EXP><<<blocks, 64, 0, s>>>(a, 2);
  k_exact_fit<NT, 96, 64, EXP><<<blocks, 64, 0, s>>>(a, 3);
  k_exact_fit<NT, 128, 64, EXP><<<blocks, 64, 0, s>>>(a, 4);
  k_exact_fit<NT, 192, 64, EXP><<<quarter, 64, 0, s>>>(a, 5);
  return hipGetLastError() == hipSuccess ? 0 : sprg_fail(SPRG_EHIP, "exact fit launch failed");
}

}  // namespace

int launch_classify(const FitArgs& a, int nterms, hipStream_t s)
{
  if (a.npatch == 0) return 0;
  k_classify<<<grid_for(a.npatch, 256), 256, 0, s>>>(a, nterms);
  return hipGetLastError() == hipSuccess ? 0 : sprg_fail(SPRG_EHIP, "classify launch failed");
}

int launch_exact_fit(const FitArgs& a, int nterms, int mode, hipStream_t s)
{
  switch (mode) {
    case FIT_REGULAR: return nterms == 4 ? fit_all<4, FIT_REGULAR>(a, s) : fit_all<10, FIT_REGULAR>(a, s);
    case FIT_EXPANDED: return nterms == 4 ? fit_all<4, FIT_EXPANDED>(a, s) : fit_all<10, FIT_EXPANDED>(a, s);
    default: return nterms == 4 ? fit_all<4, FIT_MIXED>(a, s) : fit_all<10, FIT_MIXED>(a, s);
  }
}

size_t big_ws_bytes(int grid) { return sizeof(double) * BIG_WS_DOUBLES * (size_t)grid; }

int launch_exact_big(const FitArgs& a, int nterms, double* ws, int grid, hipStream_t s)
{
  if (grid <= 0) return 0;
  if (nterms == 4)
    k_exact_big<4><<<(unsigned)grid, 64, 0, s>>>(a, ws);
  else
    k_exact_big<10><<<(unsigned)grid, 64, 0, s>>>(a, ws);
  return hipGetLastError() == hipSuccess ? 0 : sprg_fail(SPRG_EHIP, "big exact fit launch failed");
}

int launch_expand(const FitArgs& a, int nterms, hipStream_t s)
{
  const unsigned blocks = (unsigned)a.shape.expand_blocks;
  // the deferred patches are few: a smaller persistent grid
  const unsigned exact_blocks = blocks < 1024 ? blocks : 1024;
  if (nterms == 4) {
    k_expand<4, false><<<blocks, 64, 0, s>>>(a);
    k_expand<4, true><<<exact_blocks, 64, 0, s>>>(a);
  } else {
    k_expand<10, false><<<blocks, 64, 0, s>>>(a);
    k_expand<10, true><<<exact_blocks, 64, 0, s>>>(a);
  }
  k_bucket_items<<<256, 256, 0, s>>>(a);
  return hipGetLastError() == hipSuccess ? 0 : sprg_fail(SPRG_EHIP, "expand launch failed");
}

}  // namespace sprg
