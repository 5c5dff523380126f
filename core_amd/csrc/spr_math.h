// spr_math.h -- element arithmetic of the SPR path, host+device.
//
// These functions keep the operation order of the reference's apf code
// (cited per function) so that, compiled with -ffp-contract=off (every file
// of this library is), the FP64 results are bit-identical to the
// reference's x86-64 build.
#ifndef SPR_MATH_H
#define SPR_MATH_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#define SPRG_HD __host__ __device__ __forceinline__

namespace sprg {

// Tet integration rules N1, N2, N3 (apf/apfIntegrate.cc:316-361): points in
// the reference's point order, decimal constants as the reference has them.
SPRG_HD int rule_npoints(int order)
{
  return order <= 1 ? 1 : order == 2 ? 4 : order == 3 ? 5 : -1;
}

SPRG_HD void rule_point(int order, int i, double* xi)
{
  if (order <= 1) {  // N1::getPoint ignores its index
    xi[0] = 0.25; xi[1] = 0.25; xi[2] = 0.25;
    return;
  }
  if (order == 2) {
    const double a = 0.138196601125011, b = 0.585410196624969;
    xi[0] = a; xi[1] = a; xi[2] = a;
    if (i >= 1 && i <= 3) xi[i - 1] = b;
    return;
  }
  // N3: (1/4,1/4,1/4) (1/6,1/6,1/6) (1/6,1/6,1/2) (1/6,1/2,1/6) (1/2,1/6,1/6)
  if (i == 0) { xi[0] = 1. / 4.; xi[1] = 1. / 4.; xi[2] = 1. / 4.; return; }
  xi[0] = 1. / 6.; xi[1] = 1. / 6.; xi[2] = 1. / 6.;
  if (i >= 2) xi[4 - i] = 1. / 2.;
}

SPRG_HD double rule_weight(int order, int i)
{
  if (order <= 1) return 1.0 / 6.0;
  if (order == 2) return 0.25 / 6.0;
  return i == 0 ? -2. / 15. : 3. / 40.;
}

SPRG_HD int tet_nodes(int order) { return order == 1 ? 4 : 10; }

// Linear::Tetrahedron / QuadraticBase::Tetrahedron getValues
// (apf/apfShape.cc:203-230, :481-503).
SPRG_HD void tet_values(int order, const double* xi, double* v)
{
  if (order == 1) {
    v[0] = 1 - xi[0] - xi[1] - xi[2];
    v[1] = xi[0];
    v[2] = xi[1];
    v[3] = xi[2];
    return;
  }
  double xi3 = 1 - xi[0] - xi[1] - xi[2];
  v[0] = xi3 * (2 * xi3 - 1);
  v[1] = xi[0] * (2 * xi[0] - 1);
  v[2] = xi[1] * (2 * xi[1] - 1);
  v[3] = xi[2] * (2 * xi[2] - 1);
  v[4] = 4 * xi[0] * xi3;
  v[5] = 4 * xi[0] * xi[1];
  v[6] = 4 * xi[1] * xi3;
  v[7] = 4 * xi[2] * xi3;
  v[8] = 4 * xi[2] * xi[0];
  v[9] = 4 * xi[1] * xi[2];
}

// getLocalGradients of the same shapes (apf/apfShape.cc:213-222, :504-521).
SPRG_HD void tet_grad(int order, const double* xi, int n, double* g)
{
  if (order == 1) {
    g[0] = (n == 1) ? 1 : (n == 0 ? -1 : 0);
    g[1] = (n == 2) ? 1 : (n == 0 ? -1 : 0);
    g[2] = (n == 3) ? 1 : (n == 0 ? -1 : 0);
    return;
  }
  double xi3 = 1 - xi[0] - xi[1] - xi[2];
  switch (n) {
    case 0: { double d3 = 1 - 4 * xi3; g[0] = d3; g[1] = d3; g[2] = d3; return; }
    case 1: g[0] = 4 * xi[0] - 1; g[1] = 0; g[2] = 0; return;
    case 2: g[0] = 0; g[1] = 4 * xi[1] - 1; g[2] = 0; return;
    case 3: g[0] = 0; g[1] = 0; g[2] = 4 * xi[2] - 1; return;
    case 4: g[0] = 4 * xi3 - 4 * xi[0]; g[1] = -4 * xi[0]; g[2] = -4 * xi[0]; return;
    case 5: g[0] = 4 * xi[1]; g[1] = 4 * xi[0]; g[2] = 0; return;
    case 6: g[0] = -4 * xi[1]; g[1] = 4 * xi3 - 4 * xi[1]; g[2] = -4 * xi[1]; return;
    case 7: g[0] = -4 * xi[2]; g[1] = -4 * xi[2]; g[2] = 4 * xi3 - 4 * xi[2]; return;
    case 8: g[0] = 4 * xi[2]; g[1] = 0; g[2] = 4 * xi[0]; return;
    default: g[0] = 0; g[1] = 4 * xi[2]; g[2] = 4 * xi[1]; return;
  }
}

// getDeterminant<3,3>: cofactor expansion down column 0 with 2x2 minors
// expanded the same way (apf/apfMatrix.cc:86-120).
SPRG_HD double det2(double b00, double b01, double b10, double b11)
{
  double d = 0;
  d += b00 * b11;
  d += b10 * (-b01);
  return d;
}

SPRG_HD double det3(const double* A)  // row-major 3x3
{
  double c0 = det2(A[4], A[5], A[7], A[8]);
  double c1 = -det2(A[1], A[2], A[7], A[8]);
  double c2 = det2(A[1], A[2], A[4], A[5]);
  double d = 0;
  d += A[0] * c0;
  d += A[3] * c1;
  d += A[6] * c2;
  return d;
}

// invert(Matrix<3,3>) (apf/apfMatrix.h:165-173).
SPRG_HD void invert3(const double* J, double* R)
{
  // x = transpose(J): x[i] = column i of J
  const double x0[3] = {J[0], J[3], J[6]};
  const double x1[3] = {J[1], J[4], J[7]};
  const double x2[3] = {J[2], J[5], J[8]};
  R[0] = x1[1] * x2[2] - x1[2] * x2[1];
  R[1] = x1[2] * x2[0] - x1[0] * x2[2];
  R[2] = x1[0] * x2[1] - x1[1] * x2[0];
  R[3] = x2[1] * x0[2] - x2[2] * x0[1];
  R[4] = x2[2] * x0[0] - x2[0] * x0[2];
  R[5] = x2[0] * x0[1] - x2[1] * x0[0];
  R[6] = x0[1] * x1[2] - x0[2] * x1[1];
  R[7] = x0[2] * x1[0] - x0[0] * x1[2];
  R[8] = x0[0] * x1[1] - x0[1] * x1[0];
  double d = det3(J);
  for (int i = 0; i < 9; ++i) R[i] = R[i] / d;
}

SPRG_HD double dot3(const double* a, const double* b)
{
  double r = 0;
  r += a[0] * b[0];
  r += a[1] * b[1];
  r += a[2] * b[2];
  return r;
}

// evalPolynomialTerms, 3D (spr/sprRecoverField.cc:253-277).
SPRG_HD void poly_terms(int order, const double* x, double* t)
{
  t[0] = 1.0;
  t[1] = x[0];
  t[2] = x[1];
  t[3] = x[2];
  if (order == 2) {
    t[4] = x[0] * x[1];
    t[5] = x[1] * x[2];
    t[6] = x[2] * x[0];
    t[7] = x[0] * x[0];
    t[8] = x[1] * x[1];
    t[9] = x[2] * x[2];
  }
}

// apf::measure of one edge of a mesh of coordinate order 1 or 2
// (apf/apfIntegrate.cc:678-702 over the Linear::Edge / QuadraticBase::Edge
// coordinate element, rules of apf/apfIntegrate.cc:30-53).  The result does
// not depend on the edge's orientation.
SPRG_HD double edge_measure(int order, const double* x0, const double* x1,
                            const double* xm)
{
  double meas = 0;
  if (order == 1) {
    double J0[3];
    for (int j = 0; j < 3; ++j) J0[j] = x0[j] * -0.5;
    for (int j = 0; j < 3; ++j) J0[j] = J0[j] + x1[j] * 0.5;
    meas += 2 * sqrt(dot3(J0, J0));
    return meas;
  }
  for (int p = 0; p < 2; ++p) {
    double xi = p == 0 ? 0.577350269189626 : -0.577350269189626;
    double g0 = (2 * xi - 1) / 2.0, g1 = (2 * xi + 1) / 2.0, g2 = -2 * xi;
    double J0[3];
    for (int j = 0; j < 3; ++j) J0[j] = x0[j] * g0;
    for (int j = 0; j < 3; ++j) J0[j] = J0[j] + x1[j] * g1;
    for (int j = 0; j < 3; ++j) J0[j] = J0[j] + xm[j] * g2;
    meas += 1 * sqrt(dot3(J0, J0));
  }
  return meas;
}

// apf::tet_edge_verts (apf/apfMesh.cc:53-60, non-Capstone build).
SPRG_HD int tet_edge_vert(int e, int k)
{
  // (0,1) (1,2) (2,0) (0,3) (1,3) (2,3)
  const int a[6] = {0, 1, 2, 0, 1, 2};
  const int b[6] = {1, 2, 0, 3, 3, 3};
  return k == 0 ? a[e] : b[e];
}

}  // namespace sprg

#endif
