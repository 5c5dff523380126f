// k_grad.hip -- spr::getGradIPField (spr/sprGetGradIPField.cc:15-65).
//
// One thread per (tet, integration point).  The Jacobian of the coordinate
// field, its cofactor inverse and the global shape gradients follow
// apf/apfVectorElement.cc:44-66, apf/apfMatrix.h:165-173 and
// apf/apfElement.cc:77-90 operation for operation; the field gradient
// follows ScalarElement::grad / VectorElement::grad / MatrixElement::grad
// (apf/apfScalarElement.cc:18-26, apf/apfVectorElement.cc:44-59,
// apf/apfMatrixElement.cc:24-46).  HBM-bound: it reads the element's nodal
// values and coordinates and writes q*3*nc_in doubles per tet.
#include "device.h"
#include "spr_math.h"

namespace sprg {
namespace {

__device__ __forceinline__ int64_t node_id(const MeshView& m, int64_t t, int n)
{
  return n < 4 ? (int64_t)m.tv[4 * t + n] : m.nv + (int64_t)m.te[6 * t + (n - 4)];
}

template <int NCIN>
__global__ __launch_bounds__(256) void k_grad(MeshView m, int field_order,
                                              const double* __restrict__ nodal,
                                              int ip_order, int q, double* __restrict__ ip)
{
  int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (gid >= m.nt * q) return;
  const int64_t t = gid / q;
  const int p = (int)(gid % q);
  double xi[3];
  rule_point(ip_order, p, xi);
  // Jacobian: J[i][j] = x_0[j]*g_0[i], then J = J + x_n[j]*g_n[i]
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
  constexpr int NOUT = 3 * NCIN;
  double out[NOUT];
  const int nnf = tet_nodes(field_order);
  for (int n = 0; n < nnf; ++n) {
    double lg[3], gg[3];
    tet_grad(field_order, xi, n, lg);
    for (int i = 0; i < 3; ++i) gg[i] = dot3(jinv + 3 * i, lg);
    const double* u = nodal + node_id(m, t, n) * NCIN;
    if (NCIN == 1) {
      for (int i = 0; i < 3; ++i) {
        double v = gg[i] * u[0];
        out[i] = n == 0 ? v : out[i] + v;
      }
    } else if (NCIN == 3) {
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double v = u[j] * gg[i];
          out[3 * i + j] = n == 0 ? v : out[3 * i + j] + v;
        }
    } else {
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          for (int d = 0; d < 3; ++d) {
            double v = u[3 * i + j] * gg[d];
            out[i * 3 + j + d * 9] = n == 0 ? v : out[i * 3 + j + d * 9] + v;
          }
    }
  }
  double* dst = ip + gid * NOUT;
  for (int k = 0; k < NOUT; ++k) dst[k] = out[k];
}

}  // namespace

int launch_grad(const MeshView& m, int field_order, int nc_in, const double* nodal,
                int ip_order, double* ip, hipStream_t s)
{
  const int q = rule_npoints(ip_order);
  const int64_t n = m.nt * q;
  if (n == 0) return 0;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  if (nc_in == 1)
    k_grad<1><<<blocks, 256, 0, s>>>(m, field_order, nodal, ip_order, q, ip);
  else if (nc_in == 3)
    k_grad<3><<<blocks, 256, 0, s>>>(m, field_order, nodal, ip_order, q, ip);
  else
    k_grad<9><<<blocks, 256, 0, s>>>(m, field_order, nodal, ip_order, q, ip);
  return hipGetLastError() == hipSuccess ? 0 : sprg_fail(SPRG_EHIP, "grad launch failed");
}

}  // namespace sprg
