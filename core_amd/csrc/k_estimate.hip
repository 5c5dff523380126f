// k_estimate.hip -- the ZZ estimators behind spr::getSPRSizeField
// (spr/sprEstimateError.cc:98-321) and spr::getTargetSPRSizeField
// (spr/sprEstimateTargetError.cc:74-258): SelfProduct (:98-123),
// ElementError / Error / Error2 (:128-203), computeSizeFactor (:205-217),
// getCurrentSize (:219-228), getDesiredSize / getElementSizeField
// (:230-253), averageToVertex / averageSizeField (:255-299).
//
// The reference walks the elements three times (SelfProduct, Error, then
// Error2 per element inside getElementSizeField) and evaluates the recovered
// field at every integration point each time.  Here one thread per tet does
// it once: eps*(xi_p) and det J(xi_p) per point, the element error s_e, the
// self-product contribution, the element size h (max edge measure) and the
// per-element factor that only awaits the global size factor.  Block partial
// sums are reduced in a fixed order (deterministic), the size factor is
// formed on the device, and one thread per vertex averages the element
// sizes around it.  The per-element arithmetic keeps the reference's
// operation order; the two global sums differ from the reference's
// sequential element-order sums only by summation order.
#include "device.h"
#include "spr_math.h"

namespace sprg {
namespace {

constexpr int EST_T = 256;
constexpr int XCD_K_ELEM = 64;  // xcd_block chunk of the element pass

// The vertex averages give each XCD one contiguous eighth of the vertices
// (xcd_block with chunks of an eighth of the grid): the tets around a
// plane of vertices are also around the next plane's, which then reads
// them from the same L2 (0.198 -> 0.168 ms at 1.77M vertices, against
// chunks of 16 blocks).
__device__ __forceinline__ int avg_chunk(unsigned nblocks) { return nblocks >= 8 ? (int)(nblocks / 8) : 1; }

__device__ __forceinline__ int64_t node_id(const MeshView& m, int64_t t, int n)
{
  return n < 4 ? (int64_t)m.tv[4 * t + n] : m.nv + (int64_t)m.te[6 * t + (n - 4)];
}

// One thread per tet, the mesh order, component count and IP order fixed
// at compile time (shape values and rule points fold to constants).
//
//   - Recovered values at the points, v2[p][c] = (0 + r_0 N_0(xi_p)) +
//     r_1 N_1(xi_p) + ... in node order (Element::getComponents): nodes
//     outer, points inner, so each node's values are loaded once and every
//     point's sum keeps the reference's order.  (The element error measures
//     eps - eps*, about 1e-5 of eps, so these sums must round as the
//     reference's do.)
//   - dV = det J per point from the coordinate field's gradients in the
//     reference's order (apf/apfVectorElement.cc:44-66).  For a Lagrange-2
//     element the per-point sums cancel (terms ~|x|, result ~h), and the
//     determinant of a poorly shaped tet amplifies that rounding: forming J
//     once from the vertices instead moved s_e by up to 3e-10 relative on
//     the jittered n=120 mesh, so every point keeps the reference's J.  On
//     an order-1 mesh the gradients are 0 and +-1, and the same sums give
//     x_{i+1} - x_0 exactly, so J is formed once.
//   - The coordinates are used first (element size, Jacobians) and are dead
//     before the recovered values accumulate.
//   - A tet's Q x NC IP values are contiguous, so a lane that loads its own
//     tet's values touches a different cache line in every lane of every
//     load.  When the nodes go outer, the block loads its tets' IP values
//     instead as one contiguous run (16 bytes per lane, lanes in address
//     order), TB tets per pass, into LDS, and each lane reads its own tet's
//     values there.  A tet's values sit at an odd stride of RP doubles, so
//     the lanes of one read fall in different banks.  Only the source of the
//     values changes: the sums and their order are the same.
constexpr int elem_points(int ipo) { return ipo <= 1 ? 1 : ipo == 2 ? 4 : 5; }
// nodes outer needs Q x NC accumulators; above 36 the points go outer
constexpr bool elem_nodes_outer(int ipo, int nc) { return elem_points(ipo) * nc <= 36; }
// three blocks per CU (three waves per SIMD) when the nodes go outer: 168
// VGPRs, and the staging buffers below leave 3 x 52 KB of LDS
constexpr int elem_waves(int ipo, int nc) { return elem_nodes_outer(ipo, nc) ? 3 : 1; }

template <int ORDER, int NC, int IPO>
__global__ __launch_bounds__(EST_T, elem_waves(IPO, NC)) void k_element(EstArgs a)
{
  constexpr int NN = ORDER == 1 ? 4 : 10;
  constexpr int Q = elem_points(IPO);
  constexpr bool NODES_OUTER = elem_nodes_outer(IPO, NC);
  constexpr int R = Q * NC;  // IP values per tet
  constexpr int RP = R | 1;  // their LDS stride
  constexpr int STAGE_BYTES = 40960;
  constexpr int TB = EST_T * RP * 8 <= STAGE_BYTES         ? EST_T
                     : (EST_T / 2) * RP * 8 <= STAGE_BYTES ? EST_T / 2
                                                           : EST_T / 4;
  static_assert(!NODES_OUTER || TB * RP * 8 <= STAGE_BYTES, "staging pass too large");
  constexpr int NCHUNK = TB * R / 2;                      // 16-byte chunks per pass
  constexpr int CHUNKS = (NCHUNK + EST_T - 1) / EST_T;  // per thread
  __shared__ double red[2][EST_T];
  __shared__ double stage[NODES_OUTER ? TB * RP : 1];
  __shared__ double hold[Q + 1][EST_T];  // a thread's dV per point and h
  const MeshView& m = a.m;
  // logical block: its tets and its slot in the fixed-order partial sums
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x, XCD_K_ELEM);
  const int64_t t0 = blk * EST_T;
  const int64_t t = t0 + threadIdx.x;
  const bool live = t < m.nt;
  double self = 0, glob = 0;
  double h = 0;      // getCurrentSize
  double sum = 0;    // ElementError::sum
  double dV[Q];
  double v2[NODES_OUTER ? Q : 1][NC];
  auto point = [&](int p, const double* v2p, const double* v1) {
    const double w = rule_weight(IPO, p);
    // SelfProduct::atPoint: r += (v*v)*w*dV
    double vv = 0;
#pragma unroll
    for (int c = 0; c < NC; ++c) vv += v2p[c] * v2p[c];
    const double dVp = hold[p][threadIdx.x];
    self += vv * w * dVp;
    // ElementError::atPoint: diff = v1 - v2; sum += (diff*diff)*w*dV
    double dd = 0;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const double d = v1[c] - v2p[c];
      dd += d * d;
    }
    sum += dd * w * dVp;
  };
  if (live) {
    int64_t nid[NN];
    {
      const int4 vv = *reinterpret_cast<const int4*>(m.tv + 4 * t);
      nid[0] = vv.x; nid[1] = vv.y; nid[2] = vv.z; nid[3] = vv.w;
      if (ORDER == 2)
#pragma unroll
        for (int k = 0; k < 6; ++k) nid[4 + (k < 6 ? k : 0)] = m.nv + (int64_t)m.te[6 * t + k];
    }
    double X[NN][3];
#pragma unroll
    for (int n = 0; n < NN; ++n)
#pragma unroll
      for (int j = 0; j < 3; ++j) X[n][j] = m.xyz[3 * nid[n] + j];
    // getCurrentSize: max edge measure (spr/sprEstimateError.cc:219-228)
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const double* x0 = X[tet_edge_vert(k, 0)];
      const double* x1 = X[tet_edge_vert(k, 1)];
      const double* xm = ORDER == 2 ? X[4 + (ORDER == 2 ? k : 0)] : x0;
      const double e = edge_measure(ORDER, x0, x1, xm);
      h = (h < e) ? e : h;
    }
    hold[Q][threadIdx.x] = h;  // held in LDS until the end: registers are short
    if constexpr (ORDER == 1) {
      double J[9];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) J[3 * i + j] = X[i + 1][j] - X[0][j];
      const double d = det3(J);
#pragma unroll
      for (int p = 0; p < Q; ++p) dV[p] = d;
    } else {
#pragma unroll 1
      for (int p = 0; p < Q; ++p) {
        double xi[3], J[9];
        rule_point(IPO, p, xi);
#pragma unroll
        for (int n = 0; n < NN; ++n) {
          double g[3];
          tet_grad(ORDER, xi, n, g);
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
              const double v = X[n][j] * g[i];
              J[3 * i + j] = n == 0 ? v : J[3 * i + j] + v;
            }
        }
        dV[p] = det3(J);
      }
    }
#pragma unroll
    for (int p = 0; p < Q; ++p) hold[p][threadIdx.x] = dV[p];
    // the coordinates are dead from here; keep the value loads below it
    asm volatile("" ::: "memory");
    if constexpr (NODES_OUTER) {
#pragma unroll
      for (int n = 0; n < NN; ++n) {
        double r[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) r[c] = a.rec[nid[n] * NC + c];
        // at most two nodes' values in flight
        if (n & 1) asm volatile("" ::: "memory");
#pragma unroll
        for (int p = 0; p < Q; ++p) {
          double xi[3], N[NN];
          rule_point(IPO, p, xi);
          tet_values(ORDER, xi, N);
#pragma unroll
          for (int c = 0; c < NC; ++c) {
            const double v = r[c] * N[n];
            v2[p][c] = n == 0 ? 0 + v : v2[p][c] + v;
          }
        }
      }
    } else {
#pragma unroll 1
      for (int p = 0; p < Q; ++p) {
        double xi[3], N[NN], v2p[NC];
        rule_point(IPO, p, xi);
        tet_values(ORDER, xi, N);
#pragma unroll 1
        for (int n = 0; n < NN; ++n) {
          const double* r = a.rec + nid[n] * NC;
#pragma unroll
          for (int c = 0; c < NC; ++c) {
            const double v = r[c] * N[n];
            v2p[c] = n == 0 ? 0 + v : v2p[c] + v;
          }
        }
        point(p, v2p, a.ip + (t * Q + p) * NC);
      }
    }
  }
  if constexpr (NODES_OUTER) {
    // the staging loads start after the gathers above
    asm volatile("" ::: "memory");
#pragma unroll 1
    for (int pass = 0; pass < EST_T / TB; ++pass) {
      const int64_t base = t0 + (int64_t)pass * TB;  // first tet of the pass
      const int64_t left = m.nt - base;
      const int nval = left <= 0 ? 0 : left >= TB ? TB * R : (int)left * R;
      const double* src = a.ip + base * R;
      if (pass > 0) __syncthreads();  // the previous pass's reads are done
      // LDS slot of the pass's IP value d: tet d / R at stride RP
      auto slot = [](int d) { return d + (RP - R) * (d / R); };
      double x[CHUNKS][2];
      if (nval == TB * R) {  // a full pass: every block but the mesh's last
#pragma unroll
        for (int i = 0; i < CHUNKS; ++i) {
          const int d = 2 * (threadIdx.x + i * EST_T);
          if (CHUNKS * EST_T == NCHUNK || d < 2 * NCHUNK) {
            x[i][0] = src[d];
            x[i][1] = src[d + 1];
          }
        }
      } else {
#pragma unroll
        for (int i = 0; i < CHUNKS; ++i) {
          const int d = 2 * (threadIdx.x + i * EST_T);
          x[i][0] = d < nval ? src[d] : 0;
          x[i][1] = d + 1 < nval ? src[d + 1] : 0;
        }
      }
#pragma unroll
      for (int i = 0; i < CHUNKS; ++i) {
        const int d = 2 * (threadIdx.x + i * EST_T);
        if (CHUNKS * EST_T == NCHUNK || d < 2 * NCHUNK) {
          stage[slot(d)] = x[i][0];
          stage[slot(d + 1)] = x[i][1];
        }
      }
      __syncthreads();
      const int lt = (int)threadIdx.x - pass * TB;
      if (live && lt >= 0 && lt < TB) {
#pragma unroll
        for (int p = 0; p < Q; ++p) point(p, v2[p], stage + lt * RP + p * NC);
      }
    }
  }
  if (live) {
    h = hold[Q][threadIdx.x];
    const double d = 3, po = ORDER;
    glob = pow(sqrt(sum), ((2 * d) / (2 * po + d)));
    if (!a.target) {
      // Error2 + getDesiredSize: h * errorNorm, size_factor applied later
      double errorNorm = pow(sqrt(sum), -(2 / (2 * po + d)));
      a.elem[t] = h * errorNorm;
    } else {
      // getNewSize: r = pow(theta_e, -2/(2p+d)); clamping needs the factor
      const int pi = ORDER, di = 3;
      a.elem[t] = pow(sqrt(sum), ((-2.0) / (2.0 * pi + di)));
      a.elem[m.nt + t] = h;
    }
  }
  red[0][threadIdx.x] = self;
  red[1][threadIdx.x] = glob;
  __syncthreads();
  for (int s = EST_T / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    a.partial[2 * blk] = red[0][0];
    a.partial[2 * blk + 1] = red[1][0];
  }
}

// Fixed-order reduction of the block partials, then the size factor
// (computeSizeFactor, spr/sprEstimateError.cc:205-217;
// target::computeSizeFactor, spr/sprEstimateTargetError.cc:129-137).
__global__ __launch_bounds__(EST_T) void k_factor(EstArgs a)
{
  __shared__ double red[2][EST_T];
  double s0 = 0, s1 = 0;
  // each thread's partials in order, eight loads in flight at a time
  int i = threadIdx.x;
  for (; i + 7 * EST_T < a.nblocks; i += 8 * EST_T) {
    double p[16];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      p[2 * u] = a.partial[2 * (i + u * EST_T)];
      p[2 * u + 1] = a.partial[2 * (i + u * EST_T) + 1];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      s0 += p[2 * u];
      s1 += p[2 * u + 1];
    }
  }
  for (; i < a.nblocks; i += EST_T) {
    s0 += a.partial[2 * i];
    s1 += a.partial[2 * i + 1];
  }
  red[0][threadIdx.x] = s0;
  red[1][threadIdx.x] = s1;
  __syncthreads();
  for (int s = EST_T / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double self = red[0][0], glob = red[1][0];
    double sf;
    if (!a.target) {
      double epsStarNorm = sqrt(self);
      double aa = a.adapt_ratio * a.adapt_ratio * epsStarNorm * epsStarNorm;
      double b = aa / glob;
      double p = a.m.order;
      sf = pow(b, 1.0 / (2.0 * p));
    } else {
      int d = 3;
      sf = pow((glob / a.target_n), (1.0 / d));
    }
    a.scalars[0] = self;
    a.scalars[1] = glob;
    a.scalars[2] = sf;
  }
}

// averageToVertex / avgToVtx (spr/sprEstimateError.cc:255-265,
// spr/sprEstimateTargetError.cc:180-190), summing in ascending tet order.
__global__ void k_average(EstArgs a)
{
  const MeshView& m = a.m;
  const int64_t v = xcd_block(blockIdx.x, gridDim.x, avg_chunk(gridDim.x)) * blockDim.x + threadIdx.x;
  if (v >= m.nv) return;
  const double sf = a.scalars[2];
  const int32_t b = m.v2t_off[v], e = m.v2t_off[v + 1];
  auto elem_size = [&](int32_t t) {
    if (!a.target) return a.elem[t] * sf;
    double r = a.elem[t], h = a.elem[m.nt + t];
    double h_new = sf * r * h;
    if (h_new < a.alpha * h) h_new = a.alpha * h;
    if (h_new > a.beta * h) h_new = a.beta * h;
    return h_new;
  };
  // tets eight, then four, at a time so their loads overlap; the sum keeps
  // the adjacency order
  double s = 0;
  int32_t k = b;
  for (; k + 8 <= e; k += 8) {
    int32_t t[8];
    double es[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) t[j] = m.v2t[k + j];
#pragma unroll
    for (int j = 0; j < 8; ++j) es[j] = elem_size(t[j]);
#pragma unroll
    for (int j = 0; j < 8; ++j) s += es[j];
  }
  for (; k + 4 <= e; k += 4) {
    int32_t t[4];
    double es[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) t[j] = m.v2t[k + j];
#pragma unroll
    for (int j = 0; j < 4; ++j) es[j] = elem_size(t[j]);
#pragma unroll
    for (int j = 0; j < 4; ++j) s += es[j];
  }
  for (; k < e; ++k) s += elem_size(m.v2t[k]);
  s /= (double)(e - b);
  a.size[v] = s;
}

// spr::averageToVertex for every vertex (spr/sprEstimateError.cc:255-265):
// the mean of an element scalar over the vertex's tets, summed in ascending
// tet order.
__global__ void k_vertex_mean(MeshView m, const double* __restrict__ elem, double* __restrict__ out)
{
  const int64_t v = xcd_block(blockIdx.x, gridDim.x, avg_chunk(gridDim.x)) * blockDim.x + threadIdx.x;
  if (v >= m.nv) return;
  const int32_t b = m.v2t_off[v], e = m.v2t_off[v + 1];
  // as k_average: loads four tets at a time, the sum in adjacency order
  double s = 0;
  int32_t k = b;
  for (; k + 4 <= e; k += 4) {
    int32_t t[4];
    double x[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) t[j] = m.v2t[k + j];
#pragma unroll
    for (int j = 0; j < 4; ++j) x[j] = elem[t[j]];
#pragma unroll
    for (int j = 0; j < 4; ++j) s += x[j];
  }
  for (; k < e; ++k) s += elem[m.v2t[k]];
  s /= (double)(e - b);
  out[v] = s;
}

template <int ORDER, int NC>
void launch_element(EstArgs& a, hipStream_t s)
{
  switch (a.ip_order) {
    case 1: k_element<ORDER, NC, 1><<<a.nblocks, EST_T, 0, s>>>(a); break;
    case 2: k_element<ORDER, NC, 2><<<a.nblocks, EST_T, 0, s>>>(a); break;
    default:
      // the order-3 rule has 5 points; an order-2 mesh's rule has 4, so only
      // order-1 meshes take it (check_ip rejects the rest)
      if constexpr (ORDER == 1) k_element<ORDER, NC, 3><<<a.nblocks, EST_T, 0, s>>>(a);
      break;
  }
}

template <int ORDER>
int launch_element_ipo(EstArgs& a, hipStream_t s)
{
  if (a.ip_order < 1 || a.ip_order > (ORDER == 1 ? 3 : 2))
    return sprg_fail(SPRG_EUNSUPPORTED, "estimate: integration order %d on an order-%d mesh", a.ip_order, ORDER);
  switch (a.nc) {
    case 1: launch_element<ORDER, 1>(a, s); break;
    case 3: launch_element<ORDER, 3>(a, s); break;
    case 9: launch_element<ORDER, 9>(a, s); break;
    case 27: launch_element<ORDER, 27>(a, s); break;
    default: return sprg_fail(SPRG_EUNSUPPORTED, "estimate: %d components", a.nc);
  }
  return 0;
}

}  // namespace

int launch_vertex_mean(const MeshView& m, const double* elem, double* out, hipStream_t s)
{
  if (m.nv > 0) k_vertex_mean<<<(unsigned)((m.nv + 255) / 256), 256, 0, s>>>(m, elem, out);
  return hipGetLastError() == hipSuccess ? 0 : sprg_fail(SPRG_EHIP, "vertex mean launch failed");
}

int launch_estimate(EstArgs& a, hipStream_t s, const std::function<void(int)>& mark)
{
  if (a.m.nt == 0) return sprg_fail(SPRG_EINVAL, "estimate: empty mesh");
  a.nblocks = (int)((a.m.nt + EST_T - 1) / EST_T);
  mark(3);
  int rc = 0;
  rc = a.m.order == 1 ? launch_element_ipo<1>(a, s) : launch_element_ipo<2>(a, s);
  if (rc) return rc;
  mark(4);
  k_factor<<<1, EST_T, 0, s>>>(a);
  k_average<<<(unsigned)((a.m.nv + 255) / 256), 256, 0, s>>>(a);
  mark(5);
  return hipGetLastError() == hipSuccess ? 0 : sprg_fail(SPRG_EHIP, "estimate launch failed");
}

}  // namespace sprg
