// k_estimate.hip -- the ZZ estimators behind spr::getSPRSizeField
// (spr/sprEstimateError.cc:98-321) and spr::getTargetSPRSizeField
// (spr/sprEstimateTargetError.cc:74-258).
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
constexpr int XCD_K_AVG = 16;   // and of the vertex average

__device__ __forceinline__ int64_t node_id(const MeshView& m, int64_t t, int n)
{
  return n < 4 ? (int64_t)m.tv[4 * t + n] : m.nv + (int64_t)m.te[6 * t + (n - 4)];
}

template <int NC>
__global__ __launch_bounds__(EST_T) void k_element(EstArgs a)
{
  __shared__ double red[2][EST_T];
  const MeshView& m = a.m;
  // logical block: its tets and its slot in the fixed-order partial sums
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x, XCD_K_ELEM);
  const int64_t t = blk * EST_T + threadIdx.x;
  double self = 0, glob = 0;
  if (t < m.nt) {
    const int nn = tet_nodes(m.order);
    const int q = rule_npoints(a.ip_order);
    double sum = 0;  // ElementError::sum
    for (int p = 0; p < q; ++p) {
      double xi[3];
      rule_point(a.ip_order, p, xi);
      const double w = rule_weight(a.ip_order, p);
      double N[10];
      tet_values(m.order, xi, N);
      double J[9], v2[NC];
      for (int n = 0; n < nn; ++n) {
        double g[3];
        tet_grad(m.order, xi, n, g);
        const int64_t nid = node_id(m, t, n);
        const double* x = m.xyz + 3 * nid;
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) {
            double v = x[j] * g[i];
            J[3 * i + j] = n == 0 ? v : J[3 * i + j] + v;
          }
        // Element::getComponents: c = 0; c += data*N in node order
        const double* r = a.rec + nid * NC;
        for (int c = 0; c < NC; ++c) {
          double v = r[c] * N[n];
          v2[c] = n == 0 ? 0 + v : v2[c] + v;
        }
      }
      const double dV = det3(J);
      // SelfProduct::atPoint: r += (v*v)*w*dV
      double vv = 0;
      for (int c = 0; c < NC; ++c) vv += v2[c] * v2[c];
      self += vv * w * dV;
      // ElementError::atPoint: diff = v1 - v2; sum += (diff*diff)*w*dV
      const double* v1 = a.ip + (t * q + p) * NC;
      double dd = 0;
      for (int c = 0; c < NC; ++c) {
        double d = v1[c] - v2[c];
        dd += d * d;
      }
      sum += dd * w * dV;
    }
    // getCurrentSize: max edge measure (spr/sprEstimateError.cc:286-295)
    double h = 0;
    for (int k = 0; k < 6; ++k) {
      const double* x0 = m.xyz + 3 * (int64_t)m.tv[4 * t + tet_edge_vert(k, 0)];
      const double* x1 = m.xyz + 3 * (int64_t)m.tv[4 * t + tet_edge_vert(k, 1)];
      const double* xm = m.order == 2 ? m.xyz + 3 * (m.nv + (int64_t)m.te[6 * t + k]) : x0;
      double e = edge_measure(m.order, x0, x1, xm);
      h = (h < e) ? e : h;
    }
    const double d = 3, po = m.order;
    glob = pow(sqrt(sum), ((2 * d) / (2 * po + d)));
    if (!a.target) {
      // Error2 + getDesiredSize: h * errorNorm, size_factor applied later
      double errorNorm = pow(sqrt(sum), -(2 / (2 * po + d)));
      a.elem[t] = h * errorNorm;
    } else {
      // getNewSize: r = pow(theta_e, -2/(2p+d)); clamping needs the factor
      const int pi = m.order, di = 3;
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
// (computeSizeFactor, spr/sprEstimateError.cc:272-284;
// target::computeSizeFactor, spr/sprEstimateTargetError.cc:129-137).
__global__ __launch_bounds__(EST_T) void k_factor(EstArgs a)
{
  __shared__ double red[2][EST_T];
  double s0 = 0, s1 = 0;
  for (int i = threadIdx.x; i < a.nblocks; i += EST_T) {
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

// averageToVertex / avgToVtx (spr/sprEstimateError.cc:322-332,
// spr/sprEstimateTargetError.cc:180-190), summing in ascending tet order.
__global__ void k_average(EstArgs a)
{
  const MeshView& m = a.m;
  const int64_t v = xcd_block(blockIdx.x, gridDim.x, XCD_K_AVG) * blockDim.x + threadIdx.x;
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
  // tets four at a time so their loads overlap; the sum keeps the
  // adjacency order
  double s = 0;
  int32_t k = b;
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
  const int64_t v = xcd_block(blockIdx.x, gridDim.x, XCD_K_AVG) * blockDim.x + threadIdx.x;
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
  switch (a.nc) {
    case 1: k_element<1><<<a.nblocks, EST_T, 0, s>>>(a); break;
    case 3: k_element<3><<<a.nblocks, EST_T, 0, s>>>(a); break;
    case 9: k_element<9><<<a.nblocks, EST_T, 0, s>>>(a); break;
    case 27: k_element<27><<<a.nblocks, EST_T, 0, s>>>(a); break;
    default: return sprg_fail(SPRG_EUNSUPPORTED, "estimate: %d components", a.nc);
  }
  mark(4);
  k_factor<<<1, EST_T, 0, s>>>(a);
  k_average<<<(unsigned)((a.m.nv + 255) / 256), 256, 0, s>>>(a);
  mark(5);
  return hipGetLastError() == hipSuccess ? 0 : sprg_fail(SPRG_EHIP, "estimate launch failed");
}

}  // namespace sprg
