// k_grad.hip -- spr::getGradIPField (spr/sprGetGradIPField.cc:15-65).
//
// Scalar and vector fields: one thread per tet, templated on the mesh
// order, the field order, the integration order and the component count: the tet's node ids and
// coordinates are loaded once for all its points, and the rule points and
// local shape gradients are compile-time constants (a thread per (tet,
// point) with run-time orders loaded every node four times and evaluated
// the shape tables at run time).  The Jacobian of the coordinate field, its
// cofactor inverse and the global shape gradients follow
// apf/apfVectorElement.cc:44-66, apf/apfMatrix.h:165-173 and
// apf/apfElement.cc:77-90 operation for operation; the field gradient
// follows ScalarElement::grad / VectorElement::grad / MatrixElement::grad
// (apf/apfScalarElement.cc:18-26, apf/apfVectorElement.cc:44-59,
// apf/apfMatrixElement.cc:24-46).  Constant folding of the rule and shape
// expressions rounds as the run-time evaluation does (IEEE operations in
// the written order, no contraction), so the results are unchanged.
// HBM-bound: it reads the element's nodal values and coordinates and
// writes q*3*nc_in doubles per tet.
#include "device.h"
#include "spr_math.h"

namespace sprg {
namespace {

constexpr int XCD_K_GRAD = 64;  // xcd_block chunk: neighbouring tets share nodes

template <int MO, int FO, int IPO, int NCIN>
__global__ __launch_bounds__(256) void k_grad(MeshView m, const double* __restrict__ nodal,
                                              double* __restrict__ ip)
{
  constexpr int Q = IPO <= 1 ? 1 : IPO == 2 ? 4 : 5;
  constexpr int NNM = MO == 1 ? 4 : 10;
  constexpr int NNF = FO == 1 ? 4 : 10;
  constexpr int NN = NNM > NNF ? NNM : NNF;
  constexpr int NOUT = 3 * NCIN;
  const int64_t t = xcd_block(blockIdx.x, gridDim.x, XCD_K_GRAD) * blockDim.x + threadIdx.x;
  if (t >= m.nt) return;
  int64_t node[NN];
  const int4 tv = *reinterpret_cast<const int4*>(m.tv + 4 * t);
  node[0] = tv.x; node[1] = tv.y; node[2] = tv.z; node[3] = tv.w;
#pragma unroll
  for (int k = 4; k < NN; ++k) node[k] = m.nv + (int64_t)m.te[6 * t + (k - 4)];
  double x[NNM][3];
#pragma unroll
  for (int n = 0; n < NNM; ++n)
#pragma unroll
    for (int j = 0; j < 3; ++j) x[n][j] = m.xyz[3 * node[n] + j];
  double* const dst = ip + t * (Q * NOUT);
  // the inverse Jacobian at every point; the coordinates are dead after it
  double jinv[Q][9];
#pragma unroll
  for (int p = 0; p < Q; ++p) {
    double xi[3];
    rule_point(IPO, p, xi);
    // Jacobian: J[i][j] = x_0[j]*g_0[i], then J = J + x_n[j]*g_n[i]
    double J[9];
#pragma unroll
    for (int n = 0; n < NNM; ++n) {
      double g[3];
      tet_grad(MO, xi, n, g);
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          double v = x[n][j] * g[i];
          J[3 * i + j] = n == 0 ? v : J[3 * i + j] + v;
        }
    }
    invert3(J, jinv[p]);
  }
  // global shape gradient of field node n at point p
  auto shape_grad = [&](int p, int n, double* gg) {
    double xi[3], lg[3];
    rule_point(IPO, p, xi);
    tet_grad(FO, xi, n, lg);
#pragma unroll
    for (int i = 0; i < 3; ++i) gg[i] = dot3(jinv[p] + 3 * i, lg);
  };
  static_assert(NCIN <= 3, "matrix fields take k_grad_point");
  // node by node, each node's values loaded once, every point's sums
  // advanced (each sum still runs over the nodes in order)
  double out[Q][NOUT];
#pragma unroll
  for (int n = 0; n < NNF; ++n) {
    double u[NCIN];
#pragma unroll
    for (int j = 0; j < NCIN; ++j) u[j] = nodal[node[n] * NCIN + j];
#pragma unroll
    for (int p = 0; p < Q; ++p) {
      double gg[3];
      shape_grad(p, n, gg);
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < NCIN; ++j) {
          double v = NCIN == 1 ? gg[i] * u[0] : u[j] * gg[i];
          out[p][NCIN * i + j] = n == 0 ? v : out[p][NCIN * i + j] + v;
        }
    }
  }
#pragma unroll
  for (int p = 0; p < Q; ++p)
#pragma unroll
    for (int k = 0; k < NOUT; ++k) dst[p * NOUT + k] = out[p][k];
}

// Matrix fields (9 components): one thread per (tet, point) with run-time
// orders -- four points' 27 sums would not fit one thread's registers.  The
// same operation order as k_grad.
__device__ __forceinline__ int64_t node_id(const MeshView& m, int64_t t, int n)
{
  return n < 4 ? (int64_t)m.tv[4 * t + n] : m.nv + (int64_t)m.te[6 * t + (n - 4)];
}

__global__ __launch_bounds__(256) void k_grad_point(MeshView m, int field_order,
                                                    const double* __restrict__ nodal,
                                                    int ip_order, int q, double* __restrict__ ip)
{
  constexpr int NCIN = 9, NOUT = 27;
  const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (gid >= m.nt * q) return;
  const int64_t t = gid / q;
  const int p = (int)(gid % q);
  double xi[3];
  rule_point(ip_order, p, xi);
  double J[9];
  const int nnm = tet_nodes(m.order);
  for (int n = 0; n < nnm; ++n) {
    double g[3];
    tet_grad(m.order, xi, n, g);
    const double* x = m.xyz + 3 * node_id(m, t, n);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double v = x[j] * g[i];
        J[3 * i + j] = n == 0 ? v : J[3 * i + j] + v;
      }
  }
  double jinv[9];
  invert3(J, jinv);
  double out[NOUT];
  const int nnf = tet_nodes(field_order);
  for (int n = 0; n < nnf; ++n) {
    double lg[3], gg[3];
    tet_grad(field_order, xi, n, lg);
    for (int i = 0; i < 3; ++i) gg[i] = dot3(jinv + 3 * i, lg);
    const double* u = nodal + node_id(m, t, n) * NCIN;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int d = 0; d < 3; ++d) {
          double v = u[3 * i + j] * gg[d];
          out[i * 3 + j + d * 9] = n == 0 ? v : out[i * 3 + j + d * 9] + v;
        }
  }
  double* dst = ip + gid * NOUT;
  for (int k = 0; k < NOUT; ++k) dst[k] = out[k];
}

template <int MO, int FO, int IPO>
void launch_grad_nc(const MeshView& m, int nc_in, const double* nodal, double* ip, unsigned blocks, hipStream_t s)
{
  if (nc_in == 1)
    k_grad<MO, FO, IPO, 1><<<blocks, 256, 0, s>>>(m, nodal, ip);
  else
    k_grad<MO, FO, IPO, 3><<<blocks, 256, 0, s>>>(m, nodal, ip);
}

template <int MO, int FO>
void launch_grad_ipo(const MeshView& m, int ip_order, int nc_in, const double* nodal, double* ip,
                     unsigned blocks, hipStream_t s)
{
  if (ip_order <= 1)
    launch_grad_nc<MO, FO, 1>(m, nc_in, nodal, ip, blocks, s);
  else if (ip_order == 2)
    launch_grad_nc<MO, FO, 2>(m, nc_in, nodal, ip, blocks, s);
  else
    launch_grad_nc<MO, FO, 3>(m, nc_in, nodal, ip, blocks, s);
}

}  // namespace

// The caller has checked: nc_in 1, 3 or 9; field order 1 or 2 (2 only on a
// mesh with edges); ip_order 1..3; the mesh order is 1 or 2.
int launch_grad(const MeshView& m, int field_order, int nc_in, const double* nodal,
                int ip_order, double* ip, hipStream_t s)
{
  if (m.nt == 0) return 0;
  if (nc_in == 9) {
    const int q = rule_npoints(ip_order);
    const unsigned pblocks = (unsigned)((m.nt * q + 255) / 256);
    k_grad_point<<<pblocks, 256, 0, s>>>(m, field_order, nodal, ip_order, q, ip);
    return hipGetLastError() == hipSuccess ? 0 : sprg_fail(SPRG_EHIP, "grad launch failed");
  }
  const unsigned blocks = (unsigned)((m.nt + 255) / 256);
  if (m.order == 1) {
    if (field_order != 1) return sprg_fail(SPRG_EINVAL, "getGradIPField: an order-2 field needs mesh edges");
    launch_grad_ipo<1, 1>(m, ip_order, nc_in, nodal, ip, blocks, s);
  } else if (field_order == 1) {
    launch_grad_ipo<2, 1>(m, ip_order, nc_in, nodal, ip, blocks, s);
  } else {
    launch_grad_ipo<2, 2>(m, ip_order, nc_in, nodal, ip, blocks, s);
  }
  return hipGetLastError() == hipSuccess ? 0 : sprg_fail(SPRG_EHIP, "grad launch failed");
}

}  // namespace sprg
