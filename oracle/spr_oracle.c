/*
 * spr_oracle.c -- FP64 CPU restatement of SCOREC/core's SPR path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  It is
 * never linked into, loaded by or called from the product library.
 *
 * Built with -O2 -ffp-contract=off: every expression below keeps the
 * reference's operation order, so that rounding matches the reference's
 * x86-64 build (SSE2 doubles, no FMA) operation for operation.  Each function
 * cites the reference file:line whose arithmetic it restates.
 */
#include "spr_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static char g_err[512];

static int fail(int code, const char* fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

const char* om_last_error(void) { return g_err; }

/* ------------------------------------------------------------------------ */
/* Integration rules.  apf/apfIntegrate.cc:316-361 (tet N1, N2, N3) and
 * :30-53 (edge N1, N2).  The decimal constants are the reference's own
 * (truncated to 15 digits); parity at 1e-12 needs them verbatim. */

static const double tet_n1_xi[1][3] = {{0.25, 0.25, 0.25}};
static const double tet_n2_xi[4][3] = {
  {0.138196601125011, 0.138196601125011, 0.138196601125011},
  {0.585410196624969, 0.138196601125011, 0.138196601125011},
  {0.138196601125011, 0.585410196624969, 0.138196601125011},
  {0.138196601125011, 0.138196601125011, 0.585410196624969}};
static const double tet_n3_xi[5][3] = {
  {1. / 4., 1. / 4., 1. / 4.},
  {1. / 6., 1. / 6., 1. / 6.},
  {1. / 6., 1. / 6., 1. / 2.},
  {1. / 6., 1. / 2., 1. / 6.},
  {1. / 2., 1. / 6., 1. / 6.}};

/* getAccurate(order) picks the first rule whose accuracy >= order
 * (apf/apfIntegrate.cc:15-25); tet rules N1..N3 have accuracy 1..3. */
int om_rule_npoints(int order)
{
  if (order <= 1) return 1;
  if (order == 2) return 4;
  if (order == 3) return 5;
  return -1;
}

static const double* rule_point(int order, int i)
{
  if (order <= 1) return tet_n1_xi[0]; /* N1::getPoint ignores its index */
  if (order == 2) return tet_n2_xi[i];
  return tet_n3_xi[i];
}

static double rule_weight(int order, int i)
{
  if (order <= 1) return 1.0 / 6.0;
  if (order == 2) return 0.25 / 6.0;
  return i == 0 ? -2. / 15. : 3. / 40.;
}

int om_rule(int order, double* xi, double* w)
{
  int np = om_rule_npoints(order);
  if (np < 0) return fail(1, "unsupported integration order %d", order);
  for (int i = 0; i < np; ++i) {
    const double* p = rule_point(order, i);
    xi[3 * i + 0] = p[0];
    xi[3 * i + 1] = p[1];
    xi[3 * i + 2] = p[2];
    w[i] = rule_weight(order, i);
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Shape functions.  apf/apfShape.cc:203-230 (Linear::Tetrahedron),
 * :481-524 (QuadraticBase::Tetrahedron), :116-139 (Linear::Edge),
 * :419-445 (QuadraticBase::Edge). */

static int tet_nodes(int order) { return order == 1 ? 4 : 10; }

static void tet_values(int order, const double* xi, double* v)
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

static void tet_grads(int order, const double* xi, double (*g)[3])
{
  if (order == 1) {
    static const double lg[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    memcpy(g, lg, sizeof lg);
    return;
  }
  double xi3 = 1 - xi[0] - xi[1] - xi[2];
  double d3 = 1 - 4 * xi3;
  g[0][0] = d3; g[0][1] = d3; g[0][2] = d3;
  g[1][0] = 4 * xi[0] - 1; g[1][1] = 0; g[1][2] = 0;
  g[2][0] = 0; g[2][1] = 4 * xi[1] - 1; g[2][2] = 0;
  g[3][0] = 0; g[3][1] = 0; g[3][2] = 4 * xi[2] - 1;
  g[4][0] = 4 * xi3 - 4 * xi[0]; g[4][1] = -4 * xi[0]; g[4][2] = -4 * xi[0];
  g[5][0] = 4 * xi[1]; g[5][1] = 4 * xi[0]; g[5][2] = 0;
  g[6][0] = -4 * xi[1]; g[6][1] = 4 * xi3 - 4 * xi[1]; g[6][2] = -4 * xi[1];
  g[7][0] = -4 * xi[2]; g[7][1] = -4 * xi[2]; g[7][2] = 4 * xi3 - 4 * xi[2];
  g[8][0] = 4 * xi[2]; g[8][1] = 0; g[8][2] = 4 * xi[0];
  g[9][0] = 0; g[9][1] = 4 * xi[2]; g[9][2] = 4 * xi[1];
}

/* Element node list: vertices in downward order, then edges in
 * apf::tet_edge_verts order (apf/apfFieldData.cc:251-300). */
static void tet_node_ids(const om_mesh* m, int order, int64_t t, int64_t* nid)
{
  for (int i = 0; i < 4; ++i) nid[i] = m->tv[4 * t + i];
  if (order == 2)
    for (int i = 0; i < 6; ++i) nid[4 + i] = m->nv + m->te[6 * t + i];
}

/* Coordinates of the mesh element's coordinate-field nodes. */
static void tet_coords(const om_mesh* m, int64_t t, double (*x)[3])
{
  int64_t nid[10];
  int nn = tet_nodes(m->order);
  tet_node_ids(m, m->order, t, nid);
  for (int i = 0; i < nn; ++i)
    for (int j = 0; j < 3; ++j) x[i][j] = m->xyz[3 * nid[i] + j];
}

/* Element::getComponents for a scalar-shape field: c = 0; c += d*N in node
 * order (apf/apfElement.cc:106-114). */
static void eval_element(int nn, int nc, const double* data, const double* N,
                         double* c)
{
  for (int ci = 0; ci < nc; ++ci) c[ci] = 0;
  for (int ni = 0; ni < nn; ++ni)
    for (int ci = 0; ci < nc; ++ci) c[ci] += data[ni * nc + ci] * N[ni];
}

/* mapLocalToGlobal (apf/apf.cc:367-370): the coordinate field's value. */
static void map_to_global(int order, double (*x)[3], const double* xi,
                          double* out)
{
  double N[10];
  tet_values(order, xi, N);
  eval_element(tet_nodes(order), 3, &x[0][0], N, out);
}

/* VectorElement::getJacobian = gradHelper(local grads):
 * g = tensorProduct(grad_0, x_0); g = g + tensorProduct(grad_i, x_i)
 * (apf/apfVectorElement.cc:44-66; tensorProduct row i = b*a[i],
 * apf/apfMatrix.h:121-128). */
static void jacobian(int order, double (*x)[3], const double* xi,
                     double J[3][3])
{
  double g[10][3];
  int nn = tet_nodes(order);
  tet_grads(order, xi, g);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) J[i][j] = x[0][j] * g[0][i];
  for (int n = 1; n < nn; ++n)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) J[i][j] = J[i][j] + x[n][j] * g[n][i];
}

/* getDeterminant<3,3> by cofactor expansion along column 0, where each 2x2
 * minor is itself expanded the same way (apf/apfMatrix.cc:86-120). */
static double det2(double b00, double b01, double b10, double b11)
{
  double d = 0;
  d += b00 * b11;
  d += b10 * (-b01);
  return d;
}

static double det3(const double A[3][3])
{
  double c0 = det2(A[1][1], A[1][2], A[2][1], A[2][2]);
  double c1 = -det2(A[0][1], A[0][2], A[2][1], A[2][2]);
  double c2 = det2(A[0][1], A[0][2], A[1][1], A[1][2]);
  double d = 0;
  d += A[0][0] * c0;
  d += A[1][0] * c1;
  d += A[2][0] * c2;
  return d;
}

static void cross3(const double* a, const double* b, double* r)
{
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}

/* invert(Matrix<3,3>): rows are cross products of the columns, divided by
 * the determinant (apf/apfMatrix.h:165-173). */
static void invert3(const double J[3][3], double R[3][3])
{
  double x[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) x[j][i] = J[i][j];
  cross3(x[1], x[2], R[0]);
  cross3(x[2], x[0], R[1]);
  cross3(x[0], x[1], R[2]);
  double d = det3(J);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) R[i][j] = R[i][j] / d;
}

/* Vector dot product r = 0; r += a[i]*b[i] (apf/apfVector.h). */
static double dot3(const double* a, const double* b)
{
  double r = 0;
  for (int i = 0; i < 3; ++i) r += a[i] * b[i];
  return r;
}

/* ------------------------------------------------------------------------ */
/* getGradIPField -- spr/sprGetGradIPField.cc:15-65. */

int om_grad_ip(const om_mesh* m, int field_order, int ncomp_in,
               const double* nodal, int ip_order, double* ip)
{
  if (ncomp_in != 1 && ncomp_in != 3 && ncomp_in != 9)
    return fail(1, "getGradIPField: value type must be scalar, vector or matrix");
  if (field_order != 1 && field_order != 2)
    return fail(1, "getGradIPField: field order must be 1 or 2");
  if (field_order == 2 && !m->te)
    return fail(1, "getGradIPField: order-2 field needs tet->edge");
  int np = om_rule_npoints(ip_order);
  if (np < 0) return fail(1, "getGradIPField: unsupported order %d", ip_order);
  int nout = 3 * ncomp_in;
  int nnf = tet_nodes(field_order);
  for (int64_t t = 0; t < m->nt; ++t) {
    double x[10][3];
    tet_coords(m, t, x);
    int64_t nid[10];
    tet_node_ids(m, field_order, t, nid);
    double u[10][9];
    for (int n = 0; n < nnf; ++n)
      for (int c = 0; c < ncomp_in; ++c) u[n][c] = nodal[nid[n] * ncomp_in + c];
    for (int p = 0; p < np; ++p) {
      const double* xi = rule_point(ip_order, p);
      double J[3][3], jinv[3][3];
      jacobian(m->order, x, xi, J);
      invert3(J, jinv);
      double lg[10][3], gg[10][3];
      tet_grads(field_order, xi, lg);
      /* Element::getGlobalGradients: jinv * localGradients[i]
       * (apf/apfElement.cc:77-90, Matrix*Vector = row dot products). */
      for (int n = 0; n < nnf; ++n)
        for (int i = 0; i < 3; ++i) gg[n][i] = dot3(jinv[i], lg[n]);
      double* out = ip + ((size_t)t * np + p) * nout;
      if (ncomp_in == 1) {
        /* ScalarElement::grad (apf/apfScalarElement.cc:18-26) */
        double g[3];
        for (int i = 0; i < 3; ++i) g[i] = gg[0][i] * u[0][0];
        for (int n = 1; n < nnf; ++n)
          for (int i = 0; i < 3; ++i) g[i] = g[i] + gg[n][i] * u[n][0];
        for (int i = 0; i < 3; ++i) out[i] = g[i];
      } else if (ncomp_in == 3) {
        /* VectorElement::grad: g[i][j] = sum_n u_n[j]*grad_n[i]
         * (apf/apfVectorElement.cc:44-59) */
        double g[3][3];
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) g[i][j] = u[0][j] * gg[0][i];
        for (int n = 1; n < nnf; ++n)
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) g[i][j] = g[i][j] + u[n][j] * gg[n][i];
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) out[3 * i + j] = g[i][j];
      } else {
        /* MatrixElement::grad: g[i*3+j+9d] (apf/apfMatrixElement.cc:24-46) */
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j)
            for (int d = 0; d < 3; ++d)
              out[i * 3 + j + d * 9] = u[0][3 * i + j] * gg[0][d];
        for (int n = 1; n < nnf; ++n)
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
              for (int d = 0; d < 3; ++d)
                out[i * 3 + j + d * 9] += u[n][3 * i + j] * gg[n][d];
      }
    }
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Householder QR with explicit Q -- mth/mthQR.cc:19-163. */

static double sgn(double x) { return (x < 0) ? -1 : 1; }

/* get_reflector + reflect_columns on R (row-major m x n), returning 0 when
 * the column norm is below the reference's absolute threshold 1e-10
 * (mth/mthQR.cc:19-63).  v has m entries. */
static int qr_r_step(int m, int n, double* R, double* v, int k)
{
  double cnorm = 0;
  for (int i = k; i < m; ++i) cnorm += R[i * n + k] * R[i * n + k];
  cnorm = sqrt(cnorm);
  if (cnorm < 1e-10) return 0;
  for (int i = 0; i < k; ++i) v[i] = 0;
  for (int i = k; i < m; ++i) v[i] = R[i * n + k];
  v[k] += sgn(R[k * n + k]) * cnorm;
  double rnorm = 0;
  for (int i = k; i < m; ++i) rnorm += v[i] * v[i];
  rnorm = sqrt(rnorm);
  for (int i = k; i < m; ++i) v[i] /= rnorm;
  for (int j = 0; j < n; ++j) {
    double dot = 0;
    for (int i = k; i < m; ++i) dot += R[i * n + j] * v[i];
    for (int i = k; i < m; ++i) R[i * n + j] -= 2 * dot * v[i];
  }
  return 1;
}

/* reflect_rows on Q (m x m) -- mth/mthQR.cc:65-80. */
static void qr_q_step(int m, double* Q, const double* v, int k)
{
  for (int i = 0; i < m; ++i) {
    double* q = Q + (size_t)i * m;
    double dot = 0;
    for (int j = k; j < m; ++j) dot += q[j] * v[j];
    for (int j = k; j < m; ++j) q[j] -= 2 * dot * v[j];
  }
}

/* The reflectors depend on R only, so the rank test runs the R recurrence
 * alone and Q is formed afterwards from the saved reflectors; the values are
 * the same as mth::decomposeQR's interleaved loop. */
static int qr_rank(int m, int n, double* R, double* V, unsigned char* used)
{
  int rank = 0;
  for (int k = 0; k < n; ++k) {
    used[k] = (unsigned char)qr_r_step(m, n, R, V + (size_t)k * m, k);
    rank += used[k];
  }
  return rank;
}

static void qr_form_q(int m, int n, double* Q, const double* V,
                      const unsigned char* used)
{
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) Q[(size_t)i * m + j] = (double)(i == j);
  for (int k = 0; k < n; ++k)
    if (used[k]) qr_q_step(m, Q, V + (size_t)k * m, k);
}

int om_decompose_qr(int m, int n, const double* a, double* q, double* r)
{
  if (m < n) return fail(1, "decomposeQR: m < n");
  double* V = (double*)malloc(sizeof(double) * (size_t)m * n);
  unsigned char used[64];
  if (n > 64) { free(V); return fail(1, "decomposeQR: n > 64"); }
  memcpy(r, a, sizeof(double) * (size_t)m * n);
  int rank = qr_rank(m, n, r, V, used);
  qr_form_q(m, n, q, V, used);
  free(V);
  return rank;
}

/* solveFromQR: y = transpose(Q)*b (mth/mth_def.h:221-232), then backsubUT
 * (mth/mthQR.cc:131-163).  Only y[0..n) enters the back substitution. */
void om_solve_from_qr(int m, int n, const double* q, const double* r,
                      const double* b, double* x)
{
  double y[64];
  for (int i = 0; i < n; ++i) {
    y[i] = 0;
    for (int j = 0; j < m; ++j) y[i] += q[(size_t)j * m + i] * b[j];
  }
  for (int ii = 0; ii < n; ++ii) {
    int i = n - ii - 1;
    x[i] = y[i];
    for (int j = i + 1; j < n; ++j) x[i] -= r[i * n + j] * x[j];
    x[i] /= r[i * n + i];
  }
}

/* ------------------------------------------------------------------------ */
/* Upward adjacency.  A patch is a std::set<MeshEntity*> ordered by handle,
 * and MDS handles of one type ascend with the entity index
 * (mds/apfMDS.cc:98-107, mds/mds.c:291-293); so rows are ascending tet
 * indices. */

static int build_csr(int64_t nrows, int64_t nt, int per, const int32_t* conn,
                     int32_t* off, int32_t* idx)
{
  memset(off, 0, sizeof(int32_t) * (size_t)(nrows + 1));
  for (int64_t i = 0; i < nt * per; ++i) {
    int32_t r = conn[i];
    if (r < 0 || r >= nrows) return fail(1, "connectivity index out of range");
    off[r + 1]++;
  }
  for (int64_t r = 0; r < nrows; ++r) off[r + 1] += off[r];
  int32_t* fill = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nrows + 1));
  memcpy(fill, off, sizeof(int32_t) * (size_t)(nrows + 1));
  for (int64_t t = 0; t < nt; ++t) /* ascending t keeps rows sorted */
    for (int j = 0; j < per; ++j) idx[fill[conn[t * per + j]]++] = (int32_t)t;
  free(fill);
  return 0;
}

int om_vertex_tets(const om_mesh* m, int32_t* off, int32_t* idx)
{
  return build_csr(m->nv, m->nt, 4, m->tv, off, idx);
}

int om_edge_tets(const om_mesh* m, int32_t* off, int32_t* idx)
{
  if (!m->te) return fail(1, "edge adjacency needs tet->edge");
  return build_csr(m->ne, m->nt, 6, m->te, off, idx);
}

/* ------------------------------------------------------------------------ */
/* Patch recovery -- spr/sprRecoverField.cc. */

typedef struct om_patcher {
  const om_mesh* m;
  int q, nterms, nc;
  int32_t *v2t_off, *v2t;
  int32_t *e2t_off, *e2t;
  /* scratch */
  int32_t* set;
  int64_t set_n, set_cap;
  int32_t* tmp;
  int64_t tmp_cap;
  double *pts, *A, *R, *Q, *V;
  int64_t rows_cap;
} patcher;

static int cmp_i32(const void* a, const void* b)
{
  int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

static void ensure_i32(int32_t** p, int64_t* cap, int64_t need)
{
  if (need <= *cap) return;
  int64_t c = *cap ? *cap : 64;
  while (c < need) c *= 2;
  *p = (int32_t*)realloc(*p, sizeof(int32_t) * (size_t)c);
  *cap = c;
}

static void ensure_rows(patcher* P, int64_t rows)
{
  if (rows <= P->rows_cap) return;
  int64_t c = P->rows_cap ? P->rows_cap : 64;
  while (c < rows) c *= 2;
  int n = P->nterms;
  P->pts = (double*)realloc(P->pts, sizeof(double) * (size_t)c * 3);
  P->A = (double*)realloc(P->A, sizeof(double) * (size_t)c * n);
  P->R = (double*)realloc(P->R, sizeof(double) * (size_t)c * n);
  P->V = (double*)realloc(P->V, sizeof(double) * (size_t)c * n);
  P->Q = (double*)realloc(P->Q, sizeof(double) * (size_t)c * c);
  P->rows_cap = c;
}

/* evalPolynomialTerms, 3D (spr/sprRecoverField.cc:253-277). */
static void poly_terms(int order, const double* x, double* t)
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

/* getSamplePoints (:193-211): element order x points_per_element, sample
 * location = getIntPoint at the MESH order (:205); the N1 rule ignores the
 * index, so an order-1 mesh puts every sample at the centroid. */
static int sample_points(patcher* P)
{
  const om_mesh* m = P->m;
  int q = P->q;
  if (m->order == 2 && q > om_rule_npoints(2))
    return fail(1, "SPR: %d points per element exceed the order-2 rule", q);
  ensure_rows(P, P->set_n * q);
  for (int64_t e = 0; e < P->set_n; ++e) {
    double x[10][3];
    tet_coords(m, P->set[e], x);
    for (int l = 0; l < q; ++l)
      map_to_global(m->order, x, rule_point(m->order, l),
                    P->pts + 3 * (e * q + l));
  }
  return 0;
}

/* hasEnoughPoints + prepareSpr + preparePolynomialFit (:283-363): rows A_i =
 * terms(x_i), R recurrence for the rank.  Leaves R and the reflectors in P. */
static int has_enough_points(patcher* P, unsigned char* used, int* err)
{
  *err = 0;
  int64_t mrows = P->set_n * P->q;
  if (mrows < P->nterms) return 0;
  if ((*err = sample_points(P))) return 0;
  int n = P->nterms;
  for (int64_t i = 0; i < mrows; ++i)
    poly_terms(P->m->order, P->pts + 3 * i, P->A + i * n);
  memcpy(P->R, P->A, sizeof(double) * (size_t)mrows * n);
  return qr_rank((int)mrows, n, P->R, P->V, used) == n;
}

/* Tets that share at least `shared` vertices with tet t (a face: 3, an
 * edge: 2, a vertex: 1); this is getAdjacent(downward(t, shared-1), 3)
 * (:165-188) on a conforming tet mesh. */
static void add_neighbors(patcher* P, int32_t t, int shared, int64_t* n)
{
  const om_mesh* m = P->m;
  const int32_t* tv = m->tv + 4 * (int64_t)t;
  for (int a = 0; a < 4; ++a) {
    int32_t v = tv[a];
    for (int32_t k = P->v2t_off[v]; k < P->v2t_off[v + 1]; ++k) {
      int32_t u = P->v2t[k];
      const int32_t* uv = m->tv + 4 * (int64_t)u;
      int s = 0;
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) s += (uv[i] == tv[j]);
      if (s >= shared) {
        ensure_i32(&P->tmp, &P->tmp_cap, *n + 1);
        P->tmp[(*n)++] = u;
      }
    }
  }
}

static void unite_sorted(patcher* P, int64_t n_new)
{
  /* merge current set with tmp[0..n_new) into a sorted unique set */
  ensure_i32(&P->tmp, &P->tmp_cap, n_new + P->set_n);
  memcpy(P->tmp + n_new, P->set, sizeof(int32_t) * (size_t)P->set_n);
  int64_t n = n_new + P->set_n;
  qsort(P->tmp, (size_t)n, sizeof(int32_t), cmp_i32);
  int64_t k = 0;
  for (int64_t i = 0; i < n; ++i)
    if (k == 0 || P->tmp[i] != P->tmp[k - 1]) P->tmp[k++] = P->tmp[i];
  ensure_i32(&P->set, &P->set_cap, k);
  memcpy(P->set, P->tmp, sizeof(int32_t) * (size_t)k);
  P->set_n = k;
}

/* buildPatch / expandAsNecessary (:365-393). */
static int build_patch(patcher* P, int64_t ent, unsigned char* used)
{
  const om_mesh* m = P->m;
  const int32_t *off, *idx;
  int64_t row;
  if (ent < m->nv) { off = P->v2t_off; idx = P->v2t; row = ent; }
  else { off = P->e2t_off; idx = P->e2t; row = ent - m->nv; }
  P->set_n = off[row + 1] - off[row];
  ensure_i32(&P->set, &P->set_cap, P->set_n);
  memcpy(P->set, idx + off[row], sizeof(int32_t) * (size_t)P->set_n);
  int err;
  if (has_enough_points(P, used, &err)) return 0;
  if (err) return err;
  for (;;) {
    int64_t old_n = P->set_n;
    int32_t* old = (int32_t*)malloc(sizeof(int32_t) * (size_t)(old_n ? old_n : 1));
    memcpy(old, P->set, sizeof(int32_t) * (size_t)old_n);
    for (int shared_dim = 2; shared_dim >= 0; --shared_dim) {
      int64_t n_new = 0;
      for (int64_t i = 0; i < old_n; ++i)
        add_neighbors(P, old[i], shared_dim + 1, &n_new);
      unite_sorted(P, n_new);
      if (has_enough_points(P, used, &err)) { free(old); return 0; }
      if (err) { free(old); return err; }
    }
    free(old);
    if (P->set_n <= old_n)
      return fail(3, "SPR: patch construction: all hope is lost.");
  }
}

static void patcher_free(patcher* P)
{
  free(P->v2t_off); free(P->v2t); free(P->e2t_off); free(P->e2t);
  free(P->set); free(P->tmp); free(P->pts); free(P->A); free(P->R);
  free(P->Q); free(P->V);
}

static int patcher_init(patcher* P, const om_mesh* m, int q, int nc)
{
  memset(P, 0, sizeof *P);
  P->m = m;
  P->q = q;
  P->nc = nc;
  if (m->order != 1 && m->order != 2)
    return fail(1, "SPR: mesh order must be 1 or 2");
  if (m->order == 2 && !m->te) return fail(1, "SPR: order-2 mesh needs tet->edge");
  if (q < 1) return fail(1, "SPR: no integration points");
  P->nterms = m->order == 1 ? 4 : 10; /* countPolynomialTerms (:66-77) */
  P->v2t_off = (int32_t*)malloc(sizeof(int32_t) * (size_t)(m->nv + 1));
  P->v2t = (int32_t*)malloc(sizeof(int32_t) * (size_t)(4 * m->nt + 1));
  int err = om_vertex_tets(m, P->v2t_off, P->v2t);
  if (err) return err;
  if (m->order == 2) {
    P->e2t_off = (int32_t*)malloc(sizeof(int32_t) * (size_t)(m->ne + 1));
    P->e2t = (int32_t*)malloc(sizeof(int32_t) * (size_t)(6 * m->nt + 1));
    err = om_edge_tets(m, P->e2t_off, P->e2t);
    if (err) return err;
  }
  return 0;
}

static int64_t num_patches(const om_mesh* m)
{
  return m->nv + (m->order == 2 ? m->ne : 0);
}

int om_patch_sets(const om_mesh* m, int q, int32_t* set_off, int32_t* set_idx,
                  int64_t cap, int64_t* used_out)
{
  patcher P;
  int err = patcher_init(&P, m, q, 1);
  if (err) { patcher_free(&P); return err; }
  unsigned char used[16];
  int64_t np = num_patches(m), pos = 0;
  set_off[0] = 0;
  for (int64_t e = 0; e < np; ++e) {
    if ((err = build_patch(&P, e, used))) break;
    if (pos + P.set_n <= cap)
      memcpy(set_idx + pos, P.set, sizeof(int32_t) * (size_t)P.set_n);
    pos += P.set_n;
    set_off[e + 1] = (int32_t)pos;
  }
  patcher_free(&P);
  *used_out = pos;
  if (!err && pos > cap) return fail(2, "patch set capacity too small");
  return err;
}

/* runSpr (:327-353): per component, solveFromQR then evalPolynomial at the
 * entity's node (m->getPoint). */
/* recoverField's loop over the patches [p0, p1) (spr/sprRecoverField.cc:
 * 283-302) with a patcher whose adjacency is built. */
static int recover_with(patcher* P, const double* ip, double* nodal,
                        int64_t p0, int64_t p1)
{
  const om_mesh* m = P->m;
  const int q = P->q, nc = P->nc;
  int err = 0;
  unsigned char used[16];
  int n = P->nterms;
  double* b = NULL;
  int64_t bcap = 0;
  int64_t np = num_patches(m);
  if (p1 > np) p1 = np;
  for (int64_t e = p0; e < p1 && !err; ++e) {
    if ((err = build_patch(P, e, used))) break;
    int64_t mrows = P->set_n * q;
    ensure_rows(P, mrows);
    qr_form_q((int)mrows, n, P->Q, P->V, used);
    if (mrows > bcap) { bcap = mrows; b = (double*)realloc(b, sizeof(double) * (size_t)bcap); }
    double node[3], terms[10], coeffs[10];
    for (int j = 0; j < 3; ++j) node[j] = m->xyz[3 * e + j];
    poly_terms(m->order, node, terms);
    for (int c = 0; c < nc; ++c) {
      for (int64_t s = 0; s < P->set_n; ++s)
        for (int l = 0; l < q; ++l)
          b[s * q + l] = ip[((size_t)P->set[s] * q + l) * nc + c];
      om_solve_from_qr((int)mrows, n, P->Q, P->R, b, coeffs);
      double v = 0; /* coeffs * terms (mth/mthVector.h:104-110) */
      for (int i = 0; i < n; ++i) v += coeffs[i] * terms[i];
      nodal[e * nc + c] = v;
    }
  }
  free(b);
  return err;
}

om_patcher* om_patcher_new(const om_mesh* m, int q, int nc)
{
  patcher* P = (patcher*)malloc(sizeof(patcher));
  if (!P) return NULL;
  if (patcher_init(P, m, q, nc)) {
    patcher_free(P);
    free(P);
    return NULL;
  }
  return P;
}

int om_patcher_recover(om_patcher* P, const double* ip, double* nodal,
                       int64_t p0, int64_t p1)
{
  return recover_with(P, ip, nodal, p0, p1);
}

void om_patcher_free(om_patcher* P)
{
  if (!P) return;
  patcher_free(P);
  free(P);
}

int om_recover_range(const om_mesh* m, int q, int nc, const double* ip,
                     double* nodal, int64_t p0, int64_t p1)
{
  patcher P;
  int err = patcher_init(&P, m, q, nc);
  if (!err) err = recover_with(&P, ip, nodal, p0, p1);
  patcher_free(&P);
  return err;
}

int om_recover(const om_mesh* m, int q, int nc, const double* ip, double* nodal)
{
  return om_recover_range(m, q, nc, ip, nodal, 0, num_patches(m));
}

/* ------------------------------------------------------------------------ */
/* Size-field estimators -- spr/sprEstimateError.cc,
 * spr/sprEstimateTargetError.cc. */

/* Gather of the recovered Lagrange field on one element (vertices then
 * edges), nodal data layout [node][nc]. */
static void gather_rec(const om_mesh* m, int64_t t, int nc, const double* rec,
                       double* data)
{
  int64_t nid[10];
  int nn = tet_nodes(m->order);
  tet_node_ids(m, m->order, t, nid);
  for (int n = 0; n < nn; ++n)
    for (int c = 0; c < nc; ++c) data[n * nc + c] = rec[nid[n] * nc + c];
}

/* ElementError::atPoint summed over the rule (spr/sprEstimateError.cc:153-160,
 * spr/sprEstimateTargetError.cc:99-113, via
 * Integrator::process(MeshElement*), apf/apfIntegrate.cc:662-676). */
static double element_error(const om_mesh* m, int ip_order, int nc,
                            const double* ip, const double* rec, int64_t t)
{
  double x[10][3], data[10 * 9], N[10], v2[9], v1[9];
  int np = om_rule_npoints(ip_order);
  int nn = tet_nodes(m->order);
  tet_coords(m, t, x);
  gather_rec(m, t, nc, rec, data);
  double sum = 0;
  for (int p = 0; p < np; ++p) {
    const double* xi = rule_point(ip_order, p);
    double w = rule_weight(ip_order, p);
    double J[3][3];
    jacobian(m->order, x, xi, J);
    double dV = det3(J);
    for (int c = 0; c < nc; ++c) v1[c] = ip[((size_t)t * np + p) * nc + c];
    tet_values(m->order, xi, N);
    eval_element(nn, nc, data, N, v2);
    for (int c = 0; c < nc; ++c) v1[c] -= v2[c];
    double dd = 0;
    for (int c = 0; c < nc; ++c) dd += v1[c] * v1[c];
    sum += dd * w * dV;
  }
  return sum;
}

int om_element_error(const om_mesh* m, int ip_order, int nc, const double* ip,
                     const double* rec, double* s)
{
  return om_element_error_range(m, ip_order, nc, ip, rec, s, 0, m->nt);
}

int om_element_error_range(const om_mesh* m, int ip_order, int nc,
                           const double* ip, const double* rec, double* s,
                           int64_t t0, int64_t t1)
{
  if (t1 > m->nt) t1 = m->nt;
  for (int64_t t = t0; t < t1; ++t)
    s[t] = element_error(m, ip_order, nc, ip, rec, t);
  return 0;
}

/* apf::measure(edge) (apf/apfIntegrate.cc:678-702): integrator of the mesh
 * order over the edge's coordinate element; P1 N1 rule, P2 N2 rule. */
static double edge_measure(int order, const double* x0, const double* x1,
                           const double* xm)
{
  double meas = 0;
  if (order == 1) {
    double J0[3];
    for (int j = 0; j < 3; ++j) J0[j] = x0[j] * -0.5;
    for (int j = 0; j < 3; ++j) J0[j] = J0[j] + x1[j] * 0.5;
    double dV = sqrt(dot3(J0, J0));
    meas += 2 * dV;
    return meas;
  }
  static const double pts[2] = {0.577350269189626, -0.577350269189626};
  for (int p = 0; p < 2; ++p) {
    double xi = pts[p];
    double g0 = (2 * xi - 1) / 2.0, g1 = (2 * xi + 1) / 2.0, g2 = -2 * xi;
    double J0[3];
    for (int j = 0; j < 3; ++j) J0[j] = x0[j] * g0;
    for (int j = 0; j < 3; ++j) J0[j] = J0[j] + x1[j] * g1;
    for (int j = 0; j < 3; ++j) J0[j] = J0[j] + xm[j] * g2;
    double dV = sqrt(dot3(J0, J0));
    meas += 1 * dV;
  }
  return meas;
}

/* getCurrentSize (spr/sprEstimateError.cc:219-228,
 * spr/sprEstimateTargetError.cc:139-147): max edge measure over the six
 * edges. */
static const int tev[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};

static double current_size(const om_mesh* m, int64_t t)
{
  double h = 0;
  for (int k = 0; k < 6; ++k) {
    const double* x0 = m->xyz + 3 * (int64_t)m->tv[4 * t + tev[k][0]];
    const double* x1 = m->xyz + 3 * (int64_t)m->tv[4 * t + tev[k][1]];
    const double* xm = m->order == 2 ? m->xyz + 3 * (m->nv + m->te[6 * t + k]) : NULL;
    double e = edge_measure(m->order, x0, x1, xm);
    h = (h < e) ? e : h; /* std::max */
  }
  return h;
}

int om_current_size(const om_mesh* m, double* h)
{
  for (int64_t t = 0; t < m->nt; ++t) h[t] = current_size(m, t);
  return 0;
}

void om_csr_mean(int64_t n, const int32_t* off, const int32_t* idx,
                 const double* val, double* out)
{
  for (int64_t v = 0; v < n; ++v) {
    double s = 0;
    for (int32_t k = off[v]; k < off[v + 1]; ++k) s += val[idx[k]];
    s /= (double)(off[v + 1] - off[v]);
    out[v] = s;
  }
}

/* averageToVertex (spr/sprEstimateError.cc:255-265; avgToVtx,
 * spr/sprEstimateTargetError.cc:180-190): mean of the element sizes around a
 * vertex.
 * The reference sums in mds get_up order (mds/mds.c:695-713); this
 * restatement sums in ascending tet order, which can differ from it by a few
 * ulp per vertex. */
static void average_to_vertex(const om_mesh* m, const double* esize,
                              const int32_t* off, const int32_t* idx,
                              double* size)
{
  om_csr_mean(m->nv, off, idx, esize, size);
}

static int check_ip(const om_mesh* m, int ip_order)
{
  if (om_rule_npoints(ip_order) < 0)
    return fail(1, "SPR: unsupported integration order %d", ip_order);
  return 0;
}

/* SelfProduct (spr/sprEstimateError.cc:98-123): the integral of the
 * recovered field's square over the tets [t0, t1), summed in tet order. */
static double self_product(const om_mesh* m, int ip_order, int nc,
                           const double* rec, int64_t t0, int64_t t1)
{
  const int q = om_rule_npoints(ip_order);
  double self = 0;
  int nn = tet_nodes(m->order);
  double* data = (double*)malloc(sizeof(double) * 10 * 9);
  for (int64_t t = t0; t < t1; ++t) {
    double x[10][3], N10[10], v[9];
    tet_coords(m, t, x);
    gather_rec(m, t, nc, rec, data);
    for (int pt = 0; pt < q; ++pt) {
      const double* xi = rule_point(ip_order, pt);
      double w = rule_weight(ip_order, pt);
      double J[3][3];
      jacobian(m->order, x, xi, J);
      double dV = det3(J);
      tet_values(m->order, xi, N10);
      eval_element(nn, nc, data, N10, v);
      double vv = 0;
      for (int c = 0; c < nc; ++c) vv += v[c] * v[c];
      self += vv * w * dV;
    }
  }
  free(data);
  return self;
}

int om_self_product_range(const om_mesh* m, int ip_order, int nc,
                          const double* rec, int64_t t0, int64_t t1,
                          double* out)
{
  int err = check_ip(m, ip_order);
  if (err) return err;
  if (t1 > m->nt) t1 = m->nt;
  *out = self_product(m, ip_order, nc, rec, t0, t1);
  return 0;
}

int om_current_size_range(const om_mesh* m, double* h, int64_t t0, int64_t t1)
{
  if (t1 > m->nt) t1 = m->nt;
  for (int64_t t = t0; t < t1; ++t) h[t] = current_size(m, t);
  return 0;
}

/* getSPRSizeField (spr/sprEstimateError.cc:301-321): recover, size factor,
 * element size, vertex average. */
int om_spr_size_rec(const om_mesh* m, int ip_order, int nc, const double* ip,
                    const double* rec, double adapt_ratio, double* size,
                    double* diag)
{
  int err = check_ip(m, ip_order);
  if (err) return err;
  double d = 3, p = m->order;
  /* SelfProduct (spr/sprEstimateError.cc:98-123) */
  double self = self_product(m, ip_order, nc, rec, 0, m->nt);
  /* Error (spr/sprEstimateError.cc:171-185) and computeSizeFactor (:205-217) */
  double* s = (double*)malloc(sizeof(double) * (size_t)m->nt);
  double errsum = 0;
  for (int64_t t = 0; t < m->nt; ++t) {
    s[t] = element_error(m, ip_order, nc, ip, rec, t);
    errsum += pow(sqrt(s[t]), ((2 * d) / (2 * p + d)));
  }
  double epsStarNorm = sqrt(self);
  double a = adapt_ratio * adapt_ratio * epsStarNorm * epsStarNorm;
  double b = a / errsum;
  double size_factor = pow(b, 1.0 / (2.0 * p));
  /* getDesiredSize / getElementSizeField (spr/sprEstimateError.cc:230-253):
   * Error2 (:189-203) recomputes the same s_e */
  double* esize = (double*)malloc(sizeof(double) * (size_t)m->nt);
  for (int64_t t = 0; t < m->nt; ++t) {
    double errorNorm = pow(sqrt(s[t]), -(2 / (2 * p + d)));
    double h = current_size(m, t);
    esize[t] = h * errorNorm * size_factor;
  }
  int32_t* off = (int32_t*)malloc(sizeof(int32_t) * (size_t)(m->nv + 1));
  int32_t* idx = (int32_t*)malloc(sizeof(int32_t) * (size_t)(4 * m->nt + 1));
  om_vertex_tets(m, off, idx);
  average_to_vertex(m, esize, off, idx, size);
  if (diag) { diag[0] = self; diag[1] = errsum; diag[2] = size_factor; diag[3] = 0; }
  free(off); free(idx); free(esize); free(s);
  return 0;
}

int om_spr_size(const om_mesh* m, int ip_order, int nc, const double* ip,
                double adapt_ratio, double* size, double* diag)
{
  int err = check_ip(m, ip_order);
  if (err) return err;
  double* rec = (double*)malloc(sizeof(double) * (size_t)num_patches(m) * nc);
  if (!(err = om_recover(m, om_rule_npoints(ip_order), nc, ip, rec)))
    err = om_spr_size_rec(m, ip_order, nc, ip, rec, adapt_ratio, size, diag);
  free(rec);
  return err;
}

/* getTargetSPRSizeField (spr/sprEstimateTargetError.cc:241-258): GlobalErrorTerm
 * (:116-127), computeSizeFactor (:129-137), getNewSize (:149-164), vertex
 * average (:180-227). */
int om_target_size_rec(const om_mesh* m, int ip_order, int nc,
                       const double* ip, const double* rec, double target,
                       double alpha, double beta, double* size, double* diag)
{
  int err = check_ip(m, ip_order);
  if (err) return err;
  if (!(target > 0)) return fail(1, "SPR target: target must be > 0");
  if (!(alpha < beta)) return fail(1, "SPR target: alpha must be < beta");
  double dd = 3, pp = m->order;
  int p = m->order, d = 3;
  double* s = (double*)malloc(sizeof(double) * (size_t)m->nt);
  double G = 0;
  for (int64_t t = 0; t < m->nt; ++t) {
    s[t] = element_error(m, ip_order, nc, ip, rec, t);
    G += pow(sqrt(s[t]), ((2 * dd) / (2 * pp + dd)));
  }
  double Nt = target;
  double size_factor = pow((G / Nt), (1.0 / d));
  double* esize = (double*)malloc(sizeof(double) * (size_t)m->nt);
  for (int64_t t = 0; t < m->nt; ++t) {
    double h = current_size(m, t);
    double theta_e = sqrt(s[t]);
    double r = pow(theta_e, ((-2.0) / (2.0 * p + d)));
    double h_new = size_factor * r * h;
    if (h_new < alpha * h) h_new = alpha * h;
    if (h_new > beta * h) h_new = beta * h;
    esize[t] = h_new;
  }
  int32_t* off = (int32_t*)malloc(sizeof(int32_t) * (size_t)(m->nv + 1));
  int32_t* idx = (int32_t*)malloc(sizeof(int32_t) * (size_t)(4 * m->nt + 1));
  om_vertex_tets(m, off, idx);
  average_to_vertex(m, esize, off, idx, size);
  if (diag) { diag[0] = G; diag[1] = size_factor; }
  free(off); free(idx); free(esize); free(s);
  return 0;
}

int om_target_size(const om_mesh* m, int ip_order, int nc, const double* ip,
                   double target, double alpha, double beta, double* size,
                   double* diag)
{
  int err = check_ip(m, ip_order);
  if (err) return err;
  if (!(target > 0)) return fail(1, "SPR target: target must be > 0");
  if (!(alpha < beta)) return fail(1, "SPR target: alpha must be < beta");
  double* rec = (double*)malloc(sizeof(double) * (size_t)num_patches(m) * nc);
  if (!(err = om_recover(m, om_rule_npoints(ip_order), nc, ip, rec)))
    err = om_target_size_rec(m, ip_order, nc, ip, rec, target, alpha, beta, size, diag);
  free(rec);
  return err;
}
