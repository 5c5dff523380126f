// device.h -- internal device-side structures and kernel launchers.
#ifndef SPRGPU_DEVICE_H
#define SPRGPU_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <functional>

#include "sprgpu_internal.h"

namespace sprg {

// Device view of one mesh part.
struct MeshView {
  int64_t nv, nt, ne;
  int order;         // coordinate-field order 1|2
  int straight;      // every edge node is the midpoint of its edge
  const int32_t* tv;   // [nt][4]
  const int32_t* te;   // [nt][6] or null
  const double* xyz;   // [nodes][3]
  const int32_t* v2t_off;  // [nv+1]
  const int32_t* v2t;      // [4nt]
  const int32_t* e2t_off;  // [ne+1]
  const int32_t* e2t;      // [6nt]
};

// Patch work lists.  Bucket b holds patches whose row count m = q*|set|
// fits rows_of_bucket(b); larger sets, up to BIG_MMAX rows, go to the big
// list.
enum { NBUCKET = 6 };
constexpr int BIG_MMAX = 512;
__host__ __device__ inline int bucket_rows(int b)
{
  const int r[NBUCKET] = {16, 32, 64, 96, 128, 192};
  return r[b];
}
__host__ __device__ inline int bucket_of(int m)
{
  for (int b = 0; b < NBUCKET; ++b)
    if (m <= bucket_rows(b)) return b;
  return NBUCKET;  // too large
}

// Status bits written by kernels (read back by the host after a call).
enum {
  ST_PATCH_FAIL = 1,     // "all hope is lost"
  ST_TOO_LARGE = 2,      // a set exceeded BIG_MMAX rows, or an expansion 192
  ST_XSET_OVERFLOW = 4,  // expanded-set storage too small
  ST_INTERNAL = 8,
  ST_ILLCOND = 16        // a patch too ill-conditioned for the FP64 fallback
};

struct Lists {
  int32_t* items[NBUCKET];   // regular: patch ids; expanded: item ids
  int32_t* count;            // [NBUCKET] regular, [NBUCKET] expanded, [1] expand-queue
  int32_t* xitems[NBUCKET];
  int32_t* xq;               // patches queued for the exact rank test /
                             // expansion; bit 31: test the initial set first
  int32_t* xq_exact;         // queued patches whose expansion needs the exact
                             // rank test (same encoding as xq); its count
                             // follows the status word
  int32_t* fbq;              // fast forms: FP64 fallback queue; bit 31: an
                             // expanded item rather than a patch id
  int32_t* big;              // sets of more than 192 (at most BIG_MMAX) rows
                             // for the exact kernel k_exact_big; bit 31: an
                             // expanded item, bit 30: rank not yet decided.
                             // Its count is status[3].
  // expanded sets
  int32_t* x_patch;
  int64_t* x_off;
  int32_t* x_len;
  int32_t* x_set;
  int64_t* x_cursor;         // [0] set entries reserved, [1] item ids (one
                             // per queued patch; x_len 0: no set stored)
  int64_t x_cap;             // capacity of x_set
  int64_t cap;               // capacity of each list (>= number of patches)
  int32_t* status;
};

// Launch shapes of the patch kernels (chosen per mesh by sprg_mesh_autotune;
// every shape computes every patch with the same arithmetic).
struct LaunchShape {
  int fast_threads = 256;   // k_fast block size (64, 128 or 256)
  int fast_blocks = 8192;   // k_fast persistent grid cap
  int exact_blocks = 8192;  // k_exact_fit grid per bucket (a quarter for 192 rows)
  int expand_blocks = 4096; // k_expand grid
  int tile_blocks = 8192;   // k_tile persistent grid cap; 0: the regular pass
                            // runs thread per patch (k_fast) instead
};

// Tiles of the regular fit pass (built once per mesh, k_tile.hip).  A tile
// takes patches in vertex order -- a vertex's patch, then the edges it owns
// (an edge belongs to its lower vertex; its tets are among that vertex's
// tets) -- and may end inside a vertex's edges, so the union of the tets of
// the vertices it touches holds every initial set of the tile.  Each
// patch is fitted by a group of g = 1, 2, 4 or 8 lanes (by set size), and
// the groups of a tile are laid out in descending size from lane 0, so a
// group never straddles a wave and starts at a multiple of its size.
struct TileView {
  int64_t ntile = 0;
  int max_union = 0;              // tets in the largest tile union
  const int32_t* slot_off = nullptr;  // [ntile+1] lanes of each tile
  // per lane: {patch | log2(g) << 27 | rank << 29, first entry of the
  // patch's set in loc, set size, 0}
  const int4* slot = nullptr;
  const int32_t* tet_off = nullptr;   // [ntile+1]
  const int32_t* tets = nullptr;      // tile unions, ascending tet ids
  // [4nt + 6nt]: the position of each v2t entry, then of each e2t entry,
  // in the union of the tile that fits its patch
  const uint16_t* loc = nullptr;
  const int32_t* leftover = nullptr;  // patches outside every tile
  int64_t nleftover = 0;
};
constexpr int TILE_DESC = 6;         // ints describing a tile to its build
constexpr int TILE_SORT_CAP = 2048;  // candidate tets a tile build sorts
constexpr int TILE_B = 64;           // lanes per tile: one wave, so a tile's
                                     // barriers never wait on other waves
// tets a tile union may hold: two per lane (the staging and its prefetch
// take a lane's two tets), 12.8 KB of LDS at 4 points per tet
constexpr int TILE_MAX_UNION = 2 * TILE_B;

struct FitArgs {
  LaunchShape shape;
  MeshView m;
  TileView tiles;
  int fast;        // fast forms: one regular list, one expanded list
  int q;           // points per element
  int nc;          // components
  int64_t npatch;  // nv (+ ne)
  const double* ip;
  double* out;
  Lists L;
};

// The dispatcher deals the workgroups of a launch round-robin over the
// eight XCDs, each with its own L2: physical block b runs on XCD b % 8, so
// neighbouring blocks -- which read the same mesh nodes -- land on different
// L2s.  xcd_block(b) renumbers the blocks so that each XCD walks chunks of
// K consecutive logical blocks, the chunks of the eight XCDs interleaved
// over the grid (every XCD still sees work from the whole index range).  A
// bijection on [0, nblocks); the tail past a multiple of 8K keeps its
// number; K = 1 is the identity.
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t nblocks, int K)
{
  const int64_t span = 8 * (int64_t)K;
  const int64_t full = nblocks / span * span;
  if (b >= full) return b;
  const int64_t i = b / 8, x = b % 8;
  return (i / K) * span + x * K + (i % K);
}

// Appends v to list[*counter] with one atomic per wave and target list.
// Every lane of the wave calls it (convergent), so the ballot sees the whole
// wave.
__device__ __forceinline__ void wave_append(int target, int my_target, int32_t* count,
                                            int32_t* list, int32_t v)
{
  const uint64_t mask = __ballot(my_target == target);
  if (!mask) return;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((unsigned long long)mask) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(count, __popcll(mask));
  base = __builtin_amdgcn_readlane(base, leader);  // leader: the same lane for the whole wave
  if (my_target == target) {
    const uint64_t below = mask & ((1ull << lane) - 1);
    list[base + __popcll(below)] = v;
  }
}

// Launchers (return 0 or an SPRG_E* code).
int launch_csr(const MeshView& m, int dim, int32_t* off, int32_t* idx,
               int32_t* tmp, hipStream_t s);
int launch_scan_i32(int32_t* a, int64_t n, int32_t* work, hipStream_t s);
int launch_copy_i32(int32_t* dst, const int32_t* src, int64_t n, hipStream_t s);
int launch_classify(const FitArgs& a, int nterms, hipStream_t s);
// mode 0: regular patches, 1: expanded sets, 2: the fast forms' refits
// (items[] lists holding patch ids or, with bit 31, expanded-set ids).
int launch_exact_fit(const FitArgs& a, int nterms, int mode, hipStream_t s);
int launch_expand(const FitArgs& a, int nterms, hipStream_t s);
// The exact fits of the big list (one wave per set, workspace ws of
// big_ws_bytes(grid) bytes); grid 0 launches nothing.
size_t big_ws_bytes(int grid);
int launch_exact_big(const FitArgs& a, int nterms, double* ws, int grid, hipStream_t s);
int launch_fast_fit(const FitArgs& a, int nterms, int form, int src, hipStream_t s);
// The regular pass of the fast forms over the mesh's tiles (k_tile.hip).
int launch_tile_fit(const FitArgs& a, int nterms, int form, hipStream_t s);
// Tile construction: desc [TILE_DESC ntile] (each tile's candidate vertex
// rows [cv0, cv1), the vertices whose patches it fits [pv0, pv1) and the
// owned edges it fits own_e[pe0, pe1)), own_e (the edges grouped by their
// owning vertex, in vertex order; null for an order-1 mesh); fills tet_off
// (device, [ntile+1]), tets and the loc arrays.  pass 0 writes each tile's
// union size to tet_off[tile + 1], pass 1 (after the scan) the unions and
// the positions of the tile's patches' tets.
int launch_tile_build(const MeshView& m, int64_t ntile, const int32_t* desc, const int32_t* own_e, int pass,
                      int32_t* tet_off, int32_t* tets, uint16_t* v2t_loc, uint16_t* e2t_loc, int32_t* status,
                      hipStream_t s);
int launch_grad(const MeshView& m, int field_order, int nc_in,
                const double* nodal, int ip_order, double* ip, hipStream_t s);

struct EstArgs {
  MeshView m;
  int ip_order, nc;
  const double* ip;
  const double* rec;   // recovered field [nodes][nc]
  int target;          // 0: getSPRSizeField, 1: getTargetSPRSizeField
  double adapt_ratio;
  double target_n, alpha, beta;
  double* elem;        // per tet: [nt] (h*errNorm) or [2nt] (r, h)
  double* partial;     // per block partial sums [nblocks][2]
  double* scalars;     // [4]: self, errsum|G, size_factor, spare
  double* size;        // [nv]
  int nblocks;
};
// mark(i) records phase event i on the stream (3: before the element pass,
// 4: after it, 5: after the vertex average).
int launch_estimate(EstArgs& a, hipStream_t s, const std::function<void(int)>& mark);
int launch_vertex_mean(const MeshView& m, const double* elem, double* out, hipStream_t s);

}  // namespace sprg

#endif
