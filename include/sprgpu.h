/*
 * sprgpu.h -- C ABI of the MI355X (gfx950) Superconvergent Patch Recovery
 * library.  It replaces the error-estimation path of SCOREC/core's `spr`
 * library (spr/spr.h:33-63) for one mesh part of tetrahedra; an apf
 * adapter maps apf::Field* in/out onto these calls (the four spr::
 * functions over one device mirror per mesh).
 *
 * Conventions
 *   - Every function returns 0 on success or an SPRG_E* code; the message
 *     of the calling thread's last failure is sprg_last_error().  Nothing
 *     aborts: the reference's PCU_ALWAYS_ASSERT / apf::fail conditions come
 *     back as codes (SPRG_EPATCH is "SPR: patch construction: all hope is
 *     lost.", spr/sprRecoverField.cc:383).
 *   - Field arrays passed to compute calls are DEVICE pointers (allocate
 *     them with hipMalloc, torch, or sprg_malloc).  Mesh arrays passed to
 *     sprg_mesh_create are HOST pointers; the library keeps its own device
 *     mirror.
 *   - Layouts are the apf ones:
 *       nodal field  [N][nc], N = nv (order 1) or nv + ne (order 2):
 *                    vertex nodes first, then one node per edge
 *       IP field     [nt][q][nc], q = points of the tet rule of the IP
 *                    order (1 -> 1, 2 -> 4, 3 -> 5), point order of
 *                    apf/apfIntegrate.cc:316-361
 *       value types  nc = 1 scalar, 3 vector, 9 matrix (row-major 3x3)
 *   - A context owns one HIP stream; calls on one context are serialised
 *     and are not re-entrant; distinct contexts are independent.  Every
 *     compute call is stream-ordered; sprg_ctx_sync waits for it.
 */
#ifndef SPRGPU_H
#define SPRGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SPRG_OK = 0,
  SPRG_EINVAL = 1,       /* bad argument (PCU_ALWAYS_ASSERT in the reference) */
  SPRG_ENOMEM = 2,       /* device allocation failed */
  SPRG_EHIP = 3,         /* HIP runtime error */
  SPRG_EPATCH = 4,       /* patch cannot reach full rank */
  SPRG_EUNSUPPORTED = 5, /* input the reference handles that this build does not */
  SPRG_EINTERNAL = 6
};

/* Numeric form of the recovery and estimation kernels. */
enum {
  SPRG_FP64 = 0,   /* plain FP64 form: the reference's Householder QR with
                      explicit Q, operation for operation (bit-identical
                      recovered field) */
  SPRG_FP32R = 1,  /* FP32 normal equations in a node-centred basis + FP64
                      residual refinement (3 steps); FP64 fallback for
                      ill-conditioned patches */
  SPRG_FP64F = 2   /* the same fast algorithm with FP64 normal equations */
};

typedef struct sprg_ctx sprg_ctx;
typedef struct sprg_mesh sprg_mesh;

const char* sprg_last_error(void);
/* Library version, e.g. "sprgpu 0.1 gfx950". */
const char* sprg_version(void);

/* ---- context and device memory --------------------------------------- */
int sprg_ctx_create(int device, sprg_ctx** out);
int sprg_ctx_destroy(sprg_ctx* ctx);
int sprg_ctx_sync(sprg_ctx* ctx);
/* The context's HIP stream (hipStream_t), for callers that order their own
 * work against it. */
void* sprg_ctx_stream(sprg_ctx* ctx);
int sprg_malloc(sprg_ctx* ctx, size_t bytes, void** dptr);
int sprg_free(sprg_ctx* ctx, void* dptr);
int sprg_upload(sprg_ctx* ctx, void* dst, const void* src, size_t bytes);
int sprg_download(sprg_ctx* ctx, void* dst, const void* src, size_t bytes);

/* ---- mesh mirror ------------------------------------------------------
 * Replaces the apf::Mesh2 queries the reference path makes on one part:
 * m->begin(3) element order, getDownward(tet, 0|1) (tet_verts [nt][4],
 * tet_edges [nt][6] in apf::tet_edge_verts order), the coordinate field
 * (node_xyz [N][3], vertices then edge nodes for a Lagrange-2 mesh) and
 * getAdjacent(vertex|edge, 3) (built on the device as CSR sorted by tet
 * index, mds/mds.c:726-743 with the handle order of mds/apfMDS.cc:98-107).
 * tet_edges may be NULL for order 1; ne is then 0. */
int sprg_mesh_create(sprg_ctx* ctx, int64_t nv, int64_t nt, int64_t ne,
                     const int32_t* tet_verts, const int32_t* tet_edges,
                     const double* node_xyz, int order, sprg_mesh** out);
int sprg_mesh_destroy(sprg_mesh* mesh);
/* counts[0..3] = nv, nt, ne, nodes */
int sprg_mesh_counts(sprg_mesh* mesh, int64_t* counts);
/* GPU time (ms, HIP events on the context stream) of the phases of the last
 * recover / size-field call on this mesh: [0] fit of the initial patch sets,
 * [1] patch expansion, [2] fit of the expanded sets, [3] element pass,
 * [4] size factor + vertex average (0 for phases the call did not run);
 * after a gradient call, [0] is the gradient kernel and the rest 0.
 * Writes min(n, 5) values and returns 5.  No reference counterpart (the
 * reference times the whole call, spr/sprEstimateError.cc:313-319). */
int sprg_mesh_phase_times(sprg_mesh* mesh, double* ms, int n);
/* Patch routes of the last recovery on the mesh: [0] patches, [1] patches
 * queued for the exact rank test / expansion, [2] expanded sets fitted,
 * [3] patches refitted by the fast forms' FP64 fallback, [4] status bits,
 * [5] queued patches whose expansion needed the reference's rank test,
 * [6] sets above 192 sample rows fitted by the large-set kernel.
 * Writes min(n, 7) values and returns 7.  No reference counterpart. */
int sprg_mesh_last_stats(sprg_mesh* mesh, int64_t* stats, int n);
/* Chooses the launch shapes of the patch kernels for this mesh by timing
 * recoveries of `ip` in `form` (coordinate descent over block sizes and
 * persistent grid sizes) and keeps the fastest for the mesh's later calls.
 * Every shape it tries computes every patch with the same arithmetic, so
 * results do not change.  Writes min(n, 5) values (fast-fit threads per
 * block, fast-fit grid cap, exact-fit grid, expansion grid, tiled-pass grid
 * cap) to `shape`, which may be NULL.  No reference counterpart. */
int sprg_mesh_autotune(sprg_mesh* mesh, int ip_order, int nc, const double* ip,
                       int form, int64_t* shape, int n);
/* Sets the launch-shape values sprg_mesh_autotune reports (n = 4 or 5:
 * block size 64/128/256; grids in [1, 2^20]; the fifth, the tiled regular
 * pass's grid cap, in [0, 2^20], where 0 runs the fast forms' regular pass
 * thread per patch instead -- the same fit, its rows summed in another
 * order), e.g. from a cache of tuned shapes.  No reference counterpart. */
int sprg_mesh_set_shape(sprg_mesh* mesh, const int64_t* shape, int n);
/* Device CSR of getAdjacent(entity, 3): dim 0 (vertices) or 1 (edges).
 * off has n+1 entries, idx has off[n]; both stay owned by the mesh. */
int sprg_mesh_adjacency(sprg_mesh* mesh, int dim, const int32_t** off,
                        const int32_t** idx);

/* ---- spr::getGradIPField (spr/sprGetGradIPField.cc:15-65) -------------
 * nodal [N][nc_in] of a Lagrange field of field_order (1|2) -> ip
 * [nt][q][3*nc_in]; nc_in 1 -> 3 (gradient), 3 -> 9 (g[i][j] = du_j/dx_i),
 * 9 -> 27 (apf PACKED layout i*3+j+9*d). */
int sprg_grad_ip(sprg_mesh* mesh, int field_order, int nc_in,
                 const double* nodal, int ip_order, double* ip);

/* ---- patch sets of spr::recoverField (spr/sprRecoverField.cc:155-393) --
 * For every recovered entity (vertices, then edges of an order-2 mesh) the
 * tet set its polynomial fit uses, after expansion.  Writes
 * set_off[npatch+1] (device, int64) and, if set_idx is not NULL and cap is
 * large enough, set_idx[set_off[npatch]] (device, int32).  *total gets the
 * required size.  Rank decisions follow mth/mthQR.cc:31 (absolute column
 * norm threshold 1e-10 on raw-coordinate monomials). */
int sprg_patch_sets(sprg_mesh* mesh, int ip_order, int64_t* set_off,
                    int32_t* set_idx, int64_t cap, int64_t* total);

/* ---- spr::recoverField (spr/sprRecoverField.cc:419-428) ----------------
 * ip [nt][q][nc] -> nodal [N][nc], N = nodes of the mesh's coordinate
 * order (the reference's "spr_<name>" Lagrange field). */
int sprg_recover(sprg_mesh* mesh, int ip_order, int nc, const double* ip,
                 double* nodal, int form);

/* ---- spr::recoverField without allocation ------------------------------
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The same recovery as sprg_recover, bit for bit, with
 * its work lists and expanded patch sets in a device array of
 * SPRG_NOALLOC_SCRATCH_BYTES that every GPU the library runs on holds from
 * its first context to the end of the process; it allocates nothing.  The
 * mesh (sprg_mesh_create) and the fields are the caller's.  Meshes whose
 * lists do not fit (about 500,000 patches) get SPRG_EINVAL, expanded sets
 * that overflow the rest SPRG_ENOMEM; use sprg_recover for those.  Every
 * call on one GPU shares the array: calls that overlap on one GPU (two
 * contexts' streams, or two host threads) race. */
#define SPRG_NOALLOC_SCRATCH_BYTES (64u << 20)
int sprg_recover_noalloc(sprg_mesh* mesh, int ip_order, int nc, const double* ip,
                         double* nodal, int form);

/* ---- spr::getSPRSizeField (spr/sprEstimateError.cc:311-321) ------------
 * ip [nt][q][nc] -> size [nv].  diag (host, optional, 4 doubles):
 * ||eps*||^2, sum of element error terms, size_factor, wall ms. */
int sprg_spr_size(sprg_mesh* mesh, int ip_order, int nc, const double* ip,
                  double adapt_ratio, double* size, int form, double* diag);

/* ---- spr::getTargetSPRSizeField (spr/sprEstimateTargetError.cc:241-258)
 * target: the element count aimed at (size_t t of the reference; 0 is
 * rejected with SPRG_EINVAL, PCU_ALWAYS_ASSERT(target > 0) at :249), as is
 * alpha >= beta (:250).  diag (host, optional, 4 doubles): G, size_factor,
 * 0, wall ms. */
int sprg_target_spr_size(sprg_mesh* mesh, int ip_order, int nc,
                         const double* ip, uint64_t target, double alpha,
                         double beta, double* size, int form, double* diag);

/* ---- spr::averageToVertex (spr/sprEstimateError.cc:255-265) -----------
 * elem [nt] (an element scalar field) -> vert [nv]: the mean over each
 * vertex's tets (getAdjacent(vertex, 3), summed in ascending tet order). */
int sprg_average_to_vertex(sprg_mesh* mesh, const double* elem, double* vert);

/* ---- many meshes per launch sequence ------------------------------------
 * A batch holds meshes of one context (each at most once).  sprg_batch_size
 * runs, for every mesh i, spr::getSPRSizeField(ip[i], adapt_ratio) when
 * target is NULL, else spr::getTargetSPRSizeField(ip[i], target[i], alpha,
 * beta) (every target[i] > 0), writing size[i] [nv_i] (device pointers).
 * The first call records all pipelines into one HIP graph; later calls
 * replay it when the fields, parameters, launch shapes and every mesh's
 * workspace are unchanged (otherwise they record again: an eager call that
 * grows a mesh's workspace invalidates the recorded graph).  ms (optional):
 * wall time of the replay.  Results equal the per-mesh calls bit for bit.
 * No reference counterpart: the reference estimates one mesh per call. */
typedef struct sprg_batch sprg_batch;
int sprg_batch_create(sprg_ctx* ctx, int nmesh, sprg_mesh* const* meshes,
                      sprg_batch** out);
int sprg_batch_destroy(sprg_batch* batch);
int sprg_batch_size(sprg_batch* batch, int ip_order, int nc,
                    const double* const* ip, const uint64_t* target,
                    double adapt_ratio, double alpha, double beta,
                    double* const* size, int form, double* ms);

/* ---- MDS .smb files (host; mds/mds_smb.c:562-606, apf::loadMdsPart) -----
 * One part file ("name0.smb"), versions 0..6, uncompressed.  sprg_smb_counts
 * reads the header: counts[8] = vertices, edges, triangles, quads, hexes,
 * prisms, pyramids, tets (SMB type order); version may be NULL.
 * sprg_smb_read fills, for a part of tetrahedra with a linear coordinate
 * field, tet_verts [nt][4] and tet_edges [nt][6] in the order
 * apf::Mesh::getDownward gives (MDS derives both from the tet's stored
 * triangles, mds/mds.c:652-686), edge_verts [ne][2] and xyz [nv][3], with
 * the file's entity indices; tet_edges and edge_verts may be NULL.  Parts
 * holding other element types get SPRG_EUNSUPPORTED (the message lists
 * them).  sprg_mesh_load_smb reads a part into a device mirror of
 * coordinate order 1, or 2 with edge nodes at the edge midpoints (the
 * projection of apf::changeShape to Lagrange-2). */
int sprg_smb_counts(const char* path, int64_t* counts, int* version);
int sprg_smb_read(const char* path, int32_t* tet_verts, int32_t* tet_edges,
                  int32_t* edge_verts, double* xyz);
int sprg_mesh_load_smb(sprg_ctx* ctx, const char* path, int order, sprg_mesh** out);

/* ---- synthetic box meshes (host; mds/apfBox.cc:178-336 numbering) ------ */
int sprg_box_counts(int n, int64_t* nv, int64_t* nt, int64_t* ne);
/* Fills tv[nt][4], te[nt][6] (may be NULL for order 1), ev[ne][2] (may be
 * NULL) and xyz[nv (+ ne for order 2)][3].  jitter is the interior-vertex
 * displacement bound in lattice units (0 = none), grade != 0 maps x -> x^2. */
int sprg_box_build(int n, int order, double jitter, uint64_t seed, int grade,
                   int32_t* tv, int32_t* te, int32_t* ev, double* xyz);

#ifdef __cplusplus
}
#endif
#endif
