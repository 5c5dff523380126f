/*
 * spr_oracle.h -- FP64 CPU restatement of SCOREC/core's Superconvergent Patch
 * Recovery path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP product
 * path is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product library never links,
 * loads or calls it.
 *
 * Every function restates the reference's arithmetic in the reference's
 * operation order (same summation order, no fused multiply-add: build with
 * -ffp-contract=off) so that its results can be compared bit-for-bit with the
 * reference wherever the reference is available, and with the HIP "exact"
 * form everywhere.  Reference file:line citations are on each function in
 * spr_oracle.c.
 *
 * Mesh convention (one apf::Mesh2 part of tetrahedra):
 *   tv [nt][4]   tet -> vertex, in the element's downward vertex order
 *   te [nt][6]   tet -> edge in apf::tet_edge_verts order (0,1)(1,2)(2,0)
 *                (0,3)(1,3)(2,3); required when order == 2
 *   ev [ne][2]   edge -> vertex
 *   xyz [(nv + (order==2 ? ne : 0))][3]  coordinate-field nodes: vertex
 *                nodes first, then one node per edge (Lagrange-2)
 * Nodal fields use the same node numbering (vertices, then edges).
 * Integration-point fields are [nt][q][ncomp], q = points of the tet rule.
 */
#ifndef SPR_ORACLE_H
#define SPR_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int64_t nv, nt, ne;
  const int32_t* tv;
  const int32_t* te;
  const int32_t* ev;
  const double* xyz;
  int order;
} om_mesh;

/* 0 on success; nonzero = error code, message in om_last_error(). */
const char* om_last_error(void);

/* Tet integration rule: number of points for an integration order (1..7). */
int om_rule_npoints(int order);
/* Copies the rule's points (xi[np][3]) and weights (w[np]). */
int om_rule(int order, double* xi, double* w);

/* mth::decomposeQR on a row-major m x n matrix; Q is m x m, R is m x n. */
int om_decompose_qr(int m, int n, const double* a, double* q, double* r);
/* mth::solveFromQR: x[n] from Q (m x m), R (m x n), b[m]. */
void om_solve_from_qr(int m, int n, const double* q, const double* r,
                      const double* b, double* x);

/* Upward adjacency as CSR with rows sorted by ascending tet index.
 * v2t_off[nv+1], v2t[4*nt]; e2t_off[ne+1], e2t[6*nt] (te required). */
int om_vertex_tets(const om_mesh* m, int32_t* off, int32_t* idx);
int om_edge_tets(const om_mesh* m, int32_t* off, int32_t* idx);

/* spr::getGradIPField.  ncomp_in in {1,3,9}, field_order in {1,2}.
 * nodal[N][ncomp_in] -> ip[nt][q][3*ncomp_in]. */
int om_grad_ip(const om_mesh* m, int field_order, int ncomp_in,
               const double* nodal, int ip_order, double* ip);

/* Patch sets of spr::recoverField.  Entities are the vertices, then (order 2)
 * the edges.  set_off[npatch+1]; set_idx has room for cap entries.  Returns
 * the number of entries written through *used (or the required count when
 * cap is too small: error 2). */
int om_patch_sets(const om_mesh* m, int q, int32_t* set_off, int32_t* set_idx,
                  int64_t cap, int64_t* used);

/* spr::recoverField: ip[nt][q][nc] -> nodal[N][nc]. */
int om_recover(const om_mesh* m, int q, int nc, const double* ip,
               double* nodal);

/* Patch range variant used for bounded CPU timing samples: recovers only the
 * patches [p0, p1) of the vertex-then-edge patch list. */
int om_recover_range(const om_mesh* m, int q, int nc, const double* ip,
                     double* nodal, int64_t p0, int64_t p1);

/* A patch fitter whose upward adjacency is built once, for repeated range
 * recoveries on one part (om_recover_range builds it on every call).
 * om_patcher_new returns NULL on error (message in om_last_error()). */
typedef struct om_patcher om_patcher;
om_patcher* om_patcher_new(const om_mesh* m, int q, int nc);
int om_patcher_recover(om_patcher* p, const double* ip, double* nodal,
                       int64_t p0, int64_t p1);
void om_patcher_free(om_patcher* p);

/* spr::getSPRSizeField: ip field of integration order ip_order -> size[nv].
 * diag (optional, 4 doubles): ||eps*||^2, sum of errors, size_factor, 0. */
int om_spr_size(const om_mesh* m, int ip_order, int nc, const double* ip,
                double adapt_ratio, double* size, double* diag);

/* spr::getTargetSPRSizeField -> size[nv]; diag: G, size_factor. */
int om_target_size(const om_mesh* m, int ip_order, int nc, const double* ip,
                   double target, double alpha, double beta, double* size,
                   double* diag);

/* The estimators from a given recovered field rec[N][nc] (what the two
 * functions above do after om_recover). */
int om_spr_size_rec(const om_mesh* m, int ip_order, int nc, const double* ip,
                    const double* rec, double adapt_ratio, double* size,
                    double* diag);
int om_target_size_rec(const om_mesh* m, int ip_order, int nc,
                       const double* ip, const double* rec, double target,
                       double alpha, double beta, double* size, double* diag);

/* Pieces of the estimators, exposed for unit checks.
 * Element error s_e for every tet, given the recovered field. */
int om_element_error(const om_mesh* m, int ip_order, int nc, const double* ip,
                     const double* rec, double* s);
/* The same for tets [t0, t1) only (bounded CPU timing samples). */
int om_element_error_range(const om_mesh* m, int ip_order, int nc,
                           const double* ip, const double* rec, double* s,
                           int64_t t0, int64_t t1);
/* apf::measure of the six edges, max per tet (getCurrentSize). */
int om_current_size(const om_mesh* m, double* h);
/* The same for tets [t0, t1) only. */
int om_current_size_range(const om_mesh* m, double* h, int64_t t0, int64_t t1);
/* out[v] = mean of val[idx[k]] for k in [off[v], off[v+1]), summed in list
 * order (the vertex average of the size estimators). */
void om_csr_mean(int64_t n, const int32_t* off, const int32_t* idx,
                 const double* val, double* out);
/* SelfProduct of getSPRSizeField (the recovered field's square integrated)
 * over tets [t0, t1), summed in tet order into *out. */
int om_self_product_range(const om_mesh* m, int ip_order, int nc,
                          const double* rec, int64_t t0, int64_t t1,
                          double* out);

#ifdef __cplusplus
}
#endif
#endif
