// runtime.hip -- the C ABI: contexts, device memory, the mesh mirror and the
// SPR entry points (include/sprgpu.h).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <iterator>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "device.h"
#include "spr_math.h"

using namespace sprg;

// Phases timed with events: 0 classify + fit of the initial patch sets,
// 1 expansion, 2 fit of the expanded sets, 3 element pass, 4 size factor and
// vertex average.
constexpr int NPHASE = 5;

struct sprg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  void* scratch = nullptr;  // this device's address of `scratch` (below)
};

namespace {

// Grow-only device buffer.
struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  int ensure(size_t n)
  {
    if (n <= bytes) return 0;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    if (hipMalloc(&p, n) != hipSuccess) {
      (void)hipGetLastError();
      return sprg_fail(SPRG_ENOMEM, "device allocation of %zu bytes failed", n);
    }
    bytes = n;
    return 0;
  }
  void release()
  {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T>
  T* as() const { return static_cast<T*>(p); }
};

}  // namespace

struct sprg_mesh {
  sprg_ctx* ctx = nullptr;
  int64_t nv = 0, nt = 0, ne = 0, nnodes = 0;
  int order = 1;
  int straight = 1;
  DBuf tv, te, xyz, v2t_off, v2t, e2t_off, e2t;
  // tiles of the fast forms' regular pass (k_tile.hip)
  DBuf t_slot_off, t_slot, t_tet_off, t_tets, t_loc, t_leftover;
  int64_t ntile = 0, nleftover = 0;
  int max_union = 0;
  // per-call workspace
  DBuf lists, xset, rec, elem, partial, scalars, status, setlen, big_ws;
  int max_set = 0;    // tets in the largest initial patch set
  int64_t list_cap = 0, xset_cap = 0;
  // phase timing of the last call: events on the context stream
  hipEvent_t ev[NPHASE + 1] = {};
  bool ev_ok = false;
  double phase_ms[NPHASE] = {};
  int64_t last_stats[7] = {};
  LaunchShape shape;  // launch shapes of the patch kernels (sprg_mesh_autotune)
  int phase_last = 0;  // highest phase recorded by the last call
  void mark(int i)
  {
    if (ev_ok) (void)hipEventRecord(ev[i], ctx->stream);
    if (i > phase_last) phase_last = i;
  }

  MeshView view() const
  {
    MeshView m;
    m.nv = nv; m.nt = nt; m.ne = ne;
    m.order = order;
    m.straight = straight;
    m.tv = tv.as<int32_t>();
    m.te = te.as<int32_t>();
    m.xyz = xyz.as<double>();
    m.v2t_off = v2t_off.as<int32_t>();
    m.v2t = v2t.as<int32_t>();
    m.e2t_off = e2t_off.as<int32_t>();
    m.e2t = e2t.as<int32_t>();
    return m;
  }
  int64_t npatch() const { return nv + (order == 2 ? ne : 0); }
  TileView tiles() const
  {
    TileView t;
    t.ntile = ntile;
    t.max_union = max_union;
    t.slot_off = t_slot_off.as<int32_t>();
    t.slot = t_slot.as<int4>();
    t.tet_off = t_tet_off.as<int32_t>();
    t.tets = t_tets.as<int32_t>();
    t.loc = t_loc.as<uint16_t>();
    t.leftover = t_leftover.as<int32_t>();
    t.nleftover = nleftover;
    return t;
  }
};

#define HIP_OK(call, what)                                                    \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess)                                                     \
      return sprg_fail(SPRG_EHIP, "%s: %s", what, hipGetErrorString(e_));     \
  } while (0)

// Scratch of the allocation-free entry point (sprg_recover_noalloc): the
// work lists and expanded-set storage of one recovery.  A file-scope device
// array, so that entry point never allocates; every GPU the library runs on
// holds its own copy (SPRG_NOALLOC_SCRATCH_BYTES) until the process ends.
// Host code uses only the per-device address from hipGetSymbolAddress, never
// the array's name.
static __device__ __align__(256) char scratch[SPRG_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[MAX_DEVICES] = {};

// The device address of `scratch` on the current device, looked up once
// per device.
int scratch_address(int device, void** out)
{
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < 0 || device >= MAX_DEVICES) return sprg_fail(SPRG_EUNSUPPORTED, "device %d: at most %d devices", device, MAX_DEVICES);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return sprg_fail(SPRG_EHIP, "scratch address on device %d: %s", device, hipGetErrorString(e));
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return 0;
}

}  // namespace

extern "C" const char* sprg_version(void) { return "sprgpu 0.1 gfx950"; }

extern "C" int sprg_ctx_create(int device, sprg_ctx** out)
{
  if (!out) return sprg_fail(SPRG_EINVAL, "ctx_create: null out");
  *out = nullptr;
  int n = 0;
  HIP_OK(hipGetDeviceCount(&n), "hipGetDeviceCount");
  if (device < 0 || device >= n)
    return sprg_fail(SPRG_EINVAL, "ctx_create: device %d of %d", device, n);
  HIP_OK(hipSetDevice(device), "hipSetDevice");
  sprg_ctx* c = new (std::nothrow) sprg_ctx;
  if (!c) return sprg_fail(SPRG_ENOMEM, "ctx_create: out of host memory");
  c->device = device;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return sprg_fail(SPRG_EHIP, "ctx_create: stream creation failed");
  }
  if (int rc = scratch_address(device, &c->scratch)) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return rc;
  }
  *out = c;
  return 0;
}

extern "C" int sprg_ctx_destroy(sprg_ctx* ctx)
{
  if (!ctx) return 0;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return 0;
}

extern "C" int sprg_ctx_sync(sprg_ctx* ctx)
{
  if (!ctx) return sprg_fail(SPRG_EINVAL, "ctx_sync: null ctx");
  HIP_OK(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  return 0;
}

extern "C" void* sprg_ctx_stream(sprg_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

extern "C" int sprg_malloc(sprg_ctx* ctx, size_t bytes, void** dptr)
{
  if (!ctx || !dptr) return sprg_fail(SPRG_EINVAL, "malloc: null argument");
  HIP_OK(hipSetDevice(ctx->device), "hipSetDevice");
  if (hipMalloc(dptr, bytes ? bytes : 1) != hipSuccess) {
    (void)hipGetLastError();
    *dptr = nullptr;
    return sprg_fail(SPRG_ENOMEM, "malloc: %zu bytes", bytes);
  }
  return 0;
}

extern "C" int sprg_free(sprg_ctx* ctx, void* dptr)
{
  if (!ctx) return sprg_fail(SPRG_EINVAL, "free: null ctx");
  if (dptr) HIP_OK(hipFree(dptr), "hipFree");
  return 0;
}

extern "C" int sprg_upload(sprg_ctx* ctx, void* dst, const void* src, size_t bytes)
{
  if (!ctx || (!dst && bytes) || (!src && bytes)) return sprg_fail(SPRG_EINVAL, "upload: null argument");
  HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream), "upload");
  HIP_OK(hipStreamSynchronize(ctx->stream), "upload sync");
  return 0;
}

extern "C" int sprg_download(sprg_ctx* ctx, void* dst, const void* src, size_t bytes)
{
  if (!ctx || (!dst && bytes) || (!src && bytes)) return sprg_fail(SPRG_EINVAL, "download: null argument");
  HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream), "download");
  HIP_OK(hipStreamSynchronize(ctx->stream), "download sync");
  return 0;
}

// ---------------------------------------------------------------------------
// mesh mirror

namespace {

// log2 of the lanes that fit a patch of ns tets in the tiled pass: about
// six tets per lane, at most eight lanes, and never more lanes than tets
// (k_tile's weighted sum gives every lane of a solved set a first tet)
int group_log2(int64_t ns) { return ns <= 6 ? 0 : ns <= 12 ? 1 : ns <= 24 ? 2 : 3; }

// Builds the tiles of the regular fast pass (k_tile.hip) from the host
// connectivity and the device CSR: the host takes the patches in vertex
// order (each vertex's patch, then the edges it owns, the lower vertex
// owning an edge) into tiles of at most TILE_B lanes whose tet unions (of
// the v2t rows of the vertices whose patches or edges they take) hold at
// most TILE_MAX_UNION tets -- a tile may end inside a vertex's edges, so
// tiles fill their lanes -- and the device sorts each tile's union and
// records every patch's positions in it.  A vertex with more than
// TILE_MAX_UNION tets stays outside, with its owned edges: the
// thread-per-patch pass fits those.
int build_tiles(sprg_mesh* m, const int32_t* tet_verts, const int32_t* tet_edges)
{
  const int64_t nv = m->nv, nt = m->nt;
  const bool edges = m->order == 2;
  const int64_t ne = edges ? m->ne : 0;
  std::vector<int32_t> vcount(nv, 0), ecount(ne, 0), owner(ne, -1);
  for (int64_t i = 0; i < 4 * nt; ++i) ++vcount[tet_verts[i]];
  if (edges)
    for (int64_t t = 0; t < nt; ++t)
      for (int k = 0; k < 6; ++k) {
        const int32_t e = tet_edges[6 * t + k];
        ++ecount[e];
        if (owner[e] < 0) {
          const int32_t a = tet_verts[4 * t + tet_edge_vert(k, 0)], b = tet_verts[4 * t + tet_edge_vert(k, 1)];
          owner[e] = a < b ? a : b;
        }
      }
  m->max_set = 0;
  for (int32_t c : vcount) m->max_set = std::max(m->max_set, (int)c);
  for (int32_t c : ecount) m->max_set = std::max(m->max_set, (int)c);
  if (nv + ne >= (1 << 27)) return 0;  // the slot encoding holds 27-bit patch ids
  if (10 * nt >= INT32_MAX) return 0;  // loc offsets are int32
  std::vector<int32_t> own_off(nv + 1, 0), own_e(ne);
  for (int64_t e = 0; e < ne; ++e) ++own_off[owner[e] + 1];
  for (int64_t v = 0; v < nv; ++v) own_off[v + 1] += own_off[v];
  {
    std::vector<int32_t> cur(own_off.begin(), own_off.end() - 1);
    for (int64_t e = 0; e < ne; ++e) own_e[cur[owner[e]]++] = (int32_t)e;
  }
  hipStream_t s = m->ctx->stream;
  DBuf d_desc, d_own_e, d_status;
  int rc = 0;
  if ((rc = d_own_e.ensure(sizeof(int32_t) * (ne + 1))) || (rc = d_status.ensure(sizeof(int32_t)))) return rc;
  if (ne) HIP_OK(hipMemcpyAsync(d_own_e.p, own_e.data(), sizeof(int32_t) * ne, hipMemcpyHostToDevice, s), "tiles");
  HIP_OK(hipMemsetAsync(d_status.p, 0, sizeof(int32_t), s), "tiles");
  // A tile takes patches in vertex order (a vertex's patch, then the edges
  // it owns) while they fit TILE_B lanes, the union of the v2t rows of the
  // vertices they need holds at most TILE_MAX_UNION tets and those rows at
  // most TILE_SORT_CAP entries; it may end inside a vertex's edges, the next
  // tile taking the rest.  The unions are followed on the host (the rows
  // are ascending tet ids, as the device CSR's).  Per tile (TILE_DESC
  // ints): candidate rows [cv0, cv1), vertex patches [pv0, pv1), owned
  // edges own_e[pe0, pe1).
  std::vector<int32_t> hoff(nv + 1, 0), hrow((size_t)4 * nt);
  for (int64_t v = 0; v < nv; ++v) hoff[v + 1] = hoff[v] + vcount[v];
  {
    std::vector<int32_t> cur(hoff.begin(), hoff.end() - 1);
    for (int64_t t = 0; t < nt; ++t)
      for (int k = 0; k < 4; ++k) hrow[cur[tet_verts[4 * t + k]]++] = (int32_t)t;
  }
  std::vector<int32_t> desc, counts, leftover, uni, merged;
  int32_t d[TILE_DESC] = {0, 0, 0, 0, 0, 0};
  bool open = false;
  int sl = 0, cand = 0;
  auto close = [&]() {
    if (open) desc.insert(desc.end(), d, d + TILE_DESC);
    open = false;
  };
  auto start = [&](int32_t v) {
    d[0] = d[1] = d[2] = d[3] = v;
    d[4] = d[5] = own_off[v];
    sl = cand = 0;
    uni.clear();
    open = true;
  };
  auto row = [&](int32_t v) { return hrow.begin() + hoff[v]; };
  // an item of vertex v taking `lanes` lanes: the tile in progress takes it
  // if it fits, else a new tile starts with it
  auto take = [&](int32_t v, int lanes) {
    if (!open) start(v);
    bool fits = sl + lanes <= TILE_B;
    if (fits && d[1] <= v) {  // v's row would join the candidates
      merged.clear();
      std::set_union(uni.begin(), uni.end(), row(v), row(v + 1), std::back_inserter(merged));
      fits = (int)merged.size() <= TILE_MAX_UNION && cand + vcount[v] <= TILE_SORT_CAP;
    }
    if (!fits) {
      close();
      start(v);
      merged.assign(row(v), row(v + 1));
    }
    if (d[1] <= v) {
      uni.swap(merged);
      d[1] = v + 1;
      cand += vcount[v];
    }
    sl += lanes;
  };
  for (int32_t v = 0; v < (int32_t)nv; ++v) {
    if (vcount[v] > TILE_MAX_UNION) {  // fitted outside the tiles, with its edges
      close();
      leftover.push_back(v);
      for (int32_t i = own_off[v]; i < own_off[v + 1]; ++i) leftover.push_back((int32_t)(nv + own_e[i]));
      continue;
    }
    take(v, 1 << group_log2(vcount[v]));
    if (d[2] == d[3]) d[2] = v;
    d[3] = v + 1;
    for (int32_t i = own_off[v]; i < own_off[v + 1]; ++i) {
      take(v, 1 << group_log2(ecount[own_e[i]]));
      if (d[4] == d[5]) d[4] = i;
      d[5] = i + 1;
    }
  }
  close();
  std::vector<int32_t>().swap(hrow);
  const MeshView mv = m->view();
  {
    // the union sizes, from the device build (the same unions)
    const int64_t ntile = (int64_t)desc.size() / TILE_DESC;
    counts.assign(ntile + 1, 0);
    if (ntile) {
      if ((rc = d_desc.ensure(sizeof(int32_t) * TILE_DESC * ntile)) ||
          (rc = m->t_tet_off.ensure(sizeof(int32_t) * (ntile + 1))))
        return rc;
      HIP_OK(hipMemcpyAsync(d_desc.p, desc.data(), sizeof(int32_t) * TILE_DESC * ntile, hipMemcpyHostToDevice, s),
             "tiles");
      if ((rc = launch_tile_build(mv, ntile, d_desc.as<int32_t>(), nullptr, 0, m->t_tet_off.as<int32_t>(), nullptr,
                                  nullptr, nullptr, d_status.as<int32_t>(), s)))
        return rc;
      HIP_OK(hipMemcpyAsync(counts.data() + 1, m->t_tet_off.as<int32_t>() + 1, sizeof(int32_t) * ntile,
                            hipMemcpyDeviceToHost, s), "tiles");
      HIP_OK(hipStreamSynchronize(s), "tiles");
    }
    int mx = 0;
    for (int64_t i = 1; i <= ntile; ++i) mx = std::max(mx, counts[i]);
    if (mx > TILE_MAX_UNION) {  // not expected: every patch fits thread per patch
      desc.clear();
      counts.assign(1, 0);
      leftover.clear();
    }
    m->max_union = mx <= TILE_MAX_UNION ? mx : 0;
  }
  const int64_t ntile = (int64_t)desc.size() / TILE_DESC;
  m->ntile = ntile;
  m->nleftover = 0;
  if (!ntile) return 0;
  for (int64_t i = 0; i < ntile; ++i) counts[i + 1] += counts[i];
  // slots: every tile's patches in descending group size, groups aligned;
  // each with the start and size of its set in the loc array (the CSR
  // offsets are the prefix sums of the valences)
  std::vector<int64_t> vfirst(nv + 1, 0), efirst(ne + 1, 0);
  for (int64_t v = 0; v < nv; ++v) vfirst[v + 1] = vfirst[v] + vcount[v];
  for (int64_t e = 0; e < ne; ++e) efirst[e + 1] = efirst[e] + ecount[e];
  std::vector<int32_t> slot_off(ntile + 1, 0);
  std::vector<int4> slot;
  slot.reserve((size_t)(4 * nv + ne));
  std::vector<std::pair<int, int32_t>> pat;
  for (int64_t k = 0; k < ntile; ++k) {
    pat.clear();
    const int32_t* dk = desc.data() + TILE_DESC * k;
    for (int32_t v = dk[2]; v < dk[3]; ++v) pat.push_back({group_log2(vcount[v]), v});
    for (int32_t i = dk[4]; i < dk[5]; ++i) pat.push_back({group_log2(ecount[own_e[i]]), (int32_t)(nv + own_e[i])});
    std::stable_sort(pat.begin(), pat.end(),
                     [](const std::pair<int, int32_t>& x, const std::pair<int, int32_t>& y) { return x.first > y.first; });
    for (const auto& pr : pat) {
      const int32_t p = pr.second;
      const int64_t first = p < nv ? vfirst[p] : 4 * nt + efirst[p - nv];
      const int32_t ns = p < nv ? vcount[p] : ecount[p - nv];
      for (int r = 0; r < (1 << pr.first); ++r) {
        int4 sl;
        sl.x = (int32_t)((uint32_t)p | ((uint32_t)pr.first << 27) | ((uint32_t)r << 29));
        sl.y = (int32_t)first;
        sl.z = ns;
        sl.w = 0;
        slot.push_back(sl);
      }
    }
    slot_off[k + 1] = (int32_t)slot.size();
  }
  if ((rc = m->t_slot_off.ensure(sizeof(int32_t) * (ntile + 1))) || (rc = m->t_slot.ensure(sizeof(int4) * slot.size())) ||
      (rc = m->t_tets.ensure(sizeof(int32_t) * ((size_t)counts[ntile] + 1))) ||
      (rc = m->t_loc.ensure(sizeof(uint16_t) * (edges ? 10 * nt : 4 * nt))) ||
      (rc = m->t_leftover.ensure(sizeof(int32_t) * (leftover.size() + 1))))
    return rc;
  HIP_OK(hipMemcpyAsync(m->t_tet_off.p, counts.data(), sizeof(int32_t) * (ntile + 1), hipMemcpyHostToDevice, s), "tiles");
  HIP_OK(hipMemcpyAsync(m->t_slot_off.p, slot_off.data(), sizeof(int32_t) * (ntile + 1), hipMemcpyHostToDevice, s), "tiles");
  HIP_OK(hipMemcpyAsync(m->t_slot.p, slot.data(), sizeof(int4) * slot.size(), hipMemcpyHostToDevice, s), "tiles");
  if (!leftover.empty())
    HIP_OK(hipMemcpyAsync(m->t_leftover.p, leftover.data(), sizeof(int32_t) * leftover.size(), hipMemcpyHostToDevice, s),
           "tiles");
  if ((rc = launch_tile_build(mv, ntile, d_desc.as<int32_t>(), edges ? d_own_e.as<int32_t>() : nullptr, 1,
                              m->t_tet_off.as<int32_t>(), m->t_tets.as<int32_t>(),
                              m->t_loc.as<uint16_t>(), m->t_loc.as<uint16_t>() + 4 * nt, d_status.as<int32_t>(),
                              s)))
    return rc;
  int32_t st = 0;
  HIP_OK(hipMemcpyAsync(&st, d_status.p, sizeof st, hipMemcpyDeviceToHost, s), "tiles");
  HIP_OK(hipStreamSynchronize(s), "tiles");
  if (st) return sprg_fail(SPRG_EINVAL, "mesh_create: tet->edge and tet->vertex connectivity disagree");
  m->nleftover = (int64_t)leftover.size();
  return 0;
}

}  // namespace

extern "C" int sprg_mesh_create(sprg_ctx* ctx, int64_t nv, int64_t nt, int64_t ne,
                                const int32_t* tet_verts, const int32_t* tet_edges,
                                const double* node_xyz, int order, sprg_mesh** out)
{
  if (!ctx || !out || !tet_verts || !node_xyz)
    return sprg_fail(SPRG_EINVAL, "mesh_create: null argument");
  *out = nullptr;
  if (order != 1 && order != 2)
    return sprg_fail(SPRG_EUNSUPPORTED, "mesh_create: coordinate order %d (1 or 2 supported)", order);
  if (nv <= 0 || nt <= 0 || ne < 0)
    return sprg_fail(SPRG_EINVAL, "mesh_create: empty mesh");
  if (order == 2 && (!tet_edges || ne <= 0))
    return sprg_fail(SPRG_EINVAL, "mesh_create: a Lagrange-2 mesh needs tet->edge");
  if (4 * nt > INT32_MAX || 6 * nt > INT32_MAX || nv + ne > INT32_MAX)
    return sprg_fail(SPRG_EUNSUPPORTED, "mesh_create: part too large for int32 indices");
  // host-side validation: every index in range
  for (int64_t i = 0; i < 4 * nt; ++i)
    if (tet_verts[i] < 0 || tet_verts[i] >= nv)
      return sprg_fail(SPRG_EINVAL, "mesh_create: tet %lld vertex index %d out of range",
                       (long long)(i / 4), tet_verts[i]);
  if (tet_edges)
    for (int64_t i = 0; i < 6 * nt; ++i)
      if (tet_edges[i] < 0 || tet_edges[i] >= ne)
        return sprg_fail(SPRG_EINVAL, "mesh_create: tet %lld edge index %d out of range",
                         (long long)(i / 6), tet_edges[i]);
  HIP_OK(hipSetDevice(ctx->device), "hipSetDevice");
  sprg_mesh* m = new (std::nothrow) sprg_mesh;
  if (!m) return sprg_fail(SPRG_ENOMEM, "mesh_create: out of host memory");
  m->ctx = ctx;
  m->ev_ok = true;
  for (int i = 0; i <= NPHASE; ++i)
    if (hipEventCreate(&m->ev[i]) != hipSuccess) m->ev_ok = false;
  m->nv = nv;
  m->nt = nt;
  m->ne = tet_edges ? ne : 0;
  m->order = order;
  m->nnodes = nv + (order == 2 ? ne : 0);
  // straight-sided test: every edge node is the midpoint (0 + a*0.5) + b*0.5
  if (order == 2) {
    for (int64_t t = 0; t < nt && m->straight; ++t)
      for (int k = 0; k < 6 && m->straight; ++k) {
        const double* a = node_xyz + 3 * (int64_t)tet_verts[4 * t + tet_edge_vert(k, 0)];
        const double* b = node_xyz + 3 * (int64_t)tet_verts[4 * t + tet_edge_vert(k, 1)];
        const double* c = node_xyz + 3 * (nv + (int64_t)tet_edges[6 * t + k]);
        for (int j = 0; j < 3; ++j) {
          double mid = 0;
          mid += a[j] * 0.5;
          mid += b[j] * 0.5;
          if (mid != c[j]) m->straight = 0;
        }
      }
  }
  hipStream_t s = ctx->stream;
  int rc = 0;
  auto up = [&](DBuf& d, const void* src, size_t bytes) {
    if (rc) return;
    if ((rc = d.ensure(bytes))) return;
    if (hipMemcpyAsync(d.p, src, bytes, hipMemcpyHostToDevice, s) != hipSuccess)
      rc = sprg_fail(SPRG_EHIP, "mesh_create: upload failed");
  };
  up(m->tv, tet_verts, sizeof(int32_t) * 4 * nt);
  if (tet_edges) up(m->te, tet_edges, sizeof(int32_t) * 6 * nt);
  up(m->xyz, node_xyz, sizeof(double) * 3 * m->nnodes);
  DBuf tmp;
  int64_t tmpn = std::max(nv, m->ne) + 4;
  if (!rc) rc = tmp.ensure(sizeof(int32_t) * tmpn);
  if (!rc) rc = m->v2t_off.ensure(sizeof(int32_t) * (nv + 1));
  if (!rc) rc = m->v2t.ensure(sizeof(int32_t) * 4 * nt);
  if (!rc) rc = launch_csr(m->view(), 0, m->v2t_off.as<int32_t>(), m->v2t.as<int32_t>(), tmp.as<int32_t>(), s);
  if (!rc && m->ne) {
    rc = m->e2t_off.ensure(sizeof(int32_t) * (m->ne + 1));
    if (!rc) rc = m->e2t.ensure(sizeof(int32_t) * 6 * nt);
    if (!rc) rc = launch_csr(m->view(), 1, m->e2t_off.as<int32_t>(), m->e2t.as<int32_t>(), tmp.as<int32_t>(), s);
  }
  if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = sprg_fail(SPRG_EHIP, "mesh_create: sync failed");
  tmp.release();
  if (!rc) rc = build_tiles(m, tet_verts, tet_edges);
  if (rc) {
    sprg_mesh_destroy(m);
    return rc;
  }
  *out = m;
  return 0;
}

extern "C" int sprg_mesh_destroy(sprg_mesh* m)
{
  if (!m) return 0;
  (void)hipSetDevice(m->ctx->device);
  (void)hipStreamSynchronize(m->ctx->stream);
  DBuf* all[] = {&m->tv, &m->te, &m->xyz, &m->v2t_off, &m->v2t, &m->e2t_off, &m->e2t,
                 &m->lists, &m->xset, &m->rec, &m->elem, &m->partial, &m->scalars,
                 &m->status, &m->setlen, &m->big_ws, &m->t_slot_off, &m->t_slot, &m->t_tet_off,
                 &m->t_tets, &m->t_loc, &m->t_leftover};
  for (DBuf* b : all) b->release();
  for (int i = 0; i <= NPHASE; ++i)
    if (m->ev[i]) (void)hipEventDestroy(m->ev[i]);
  delete m;
  return 0;
}

extern "C" int sprg_mesh_counts(sprg_mesh* m, int64_t* counts)
{
  if (!m || !counts) return sprg_fail(SPRG_EINVAL, "mesh_counts: null argument");
  counts[0] = m->nv;
  counts[1] = m->nt;
  counts[2] = m->ne;
  counts[3] = m->nnodes;
  return 0;
}

extern "C" int sprg_mesh_phase_times(sprg_mesh* m, double* ms, int n)
{
  if (!m || !ms) return sprg_fail(SPRG_EINVAL, "phase_times: null argument");
  for (int i = 0; i < n; ++i) ms[i] = i < NPHASE ? m->phase_ms[i] : 0;
  return NPHASE;
}

extern "C" int sprg_mesh_last_stats(sprg_mesh* m, int64_t* stats, int n)
{
  if (!m || (!stats && n > 0)) return sprg_fail(SPRG_EINVAL, "mesh_last_stats: null argument");
  for (int i = 0; i < n && i < 7; ++i) stats[i] = m->last_stats[i];
  return 7;
}

extern "C" int sprg_mesh_adjacency(sprg_mesh* m, int dim, const int32_t** off, const int32_t** idx)
{
  if (!m || !off || !idx) return sprg_fail(SPRG_EINVAL, "mesh_adjacency: null argument");
  if (dim == 0) {
    *off = m->v2t_off.as<int32_t>();
    *idx = m->v2t.as<int32_t>();
    return 0;
  }
  if (dim == 1 && m->ne) {
    *off = m->e2t_off.as<int32_t>();
    *idx = m->e2t.as<int32_t>();
    return 0;
  }
  return sprg_fail(SPRG_EINVAL, "mesh_adjacency: dimension %d not available", dim);
}

// ---------------------------------------------------------------------------
// recovery pipeline

namespace {

int check_ip(sprg_mesh* m, int ip_order, int nc, int* q)
{
  *q = rule_npoints(ip_order);
  if (*q < 0) return sprg_fail(SPRG_EUNSUPPORTED, "integration order %d (1..3 supported)", ip_order);
  // getSamplePoints indexes the mesh-order rule by IP index; the order-2
  // rule has 4 points (the N1 rule of an order-1 mesh ignores the index).
  if (m->order == 2 && *q > 4)
    return sprg_fail(SPRG_EUNSUPPORTED, "%d points per element exceed the order-2 mesh rule", *q);
  if (nc < 0 || nc > 27) return sprg_fail(SPRG_EINVAL, "component count %d", nc);
  return 0;
}

void collect_phases(sprg_mesh* m)
{
  for (int i = 0; i < NPHASE; ++i) {
    float ms = 0;
    m->phase_ms[i] = 0;
    if (m->ev_ok && i < m->phase_last &&
        hipEventElapsedTime(&ms, m->ev[i], m->ev[i + 1]) == hipSuccess)
      m->phase_ms[i] = ms;
  }
}

// Lays out the work lists in one device allocation.
// Waves of the exact kernel for sets above 192 rows: none unless an initial
// set of this mesh has that many rows at q points per tet.
int big_grid(const sprg_mesh* m, int q) { return (int64_t)m->max_set * q > bucket_rows(NBUCKET - 1) ? 128 : 0; }

// Bytes of the work lists of np patches: items and xitems per bucket, xq,
// xq_exact, fbq, x_patch, x_len (int32 each), x_off (int64), counters,
// cursors and the status words.
size_t list_bytes(int64_t np)
{
  const size_t n32 = (size_t)np * (2 * NBUCKET + 6) + 2 * NBUCKET + 8;
  return n32 * 4 + (size_t)np * 8 + 64;
}

// Lays the work lists of cap patches out from base (8-byte aligned); the
// expanded sets live in x_set[0, x_cap).
void carve_lists(char* base, int64_t cap, int32_t* x_set, int64_t x_cap, Lists& L)
{
  int64_t* x_off = (int64_t*)base;  // 8-byte aligned first
  int64_t* x_cursor = x_off + cap;
  int32_t* p = (int32_t*)(x_cursor + 2);
  L.x_off = x_off;
  L.x_cursor = x_cursor;
  L.count = p;
  p += 2 * NBUCKET + 2;
  L.status = p;  // [0] status bits, [1] exact-expansion count, [2] fallback entries taken before the expansion
  p += 4;
  for (int b = 0; b < NBUCKET; ++b) { L.items[b] = p; p += cap; }
  for (int b = 0; b < NBUCKET; ++b) { L.xitems[b] = p; p += cap; }
  L.xq = p; p += cap;
  L.xq_exact = p; p += cap;
  L.fbq = p; p += cap;
  L.big = p; p += cap;
  L.x_patch = p; p += cap;
  L.x_len = p; p += cap;
  L.x_set = x_set;
  L.x_cap = x_cap;
  L.cap = cap;
}

// Lays the work lists out in the mesh's own workspaces (grown as needed).
int setup_lists(sprg_mesh* m, FitArgs& a)
{
  const int64_t np = m->npatch();
  if (m->list_cap < np) {
    int rc = m->lists.ensure(list_bytes(np));
    if (rc) return rc;
    m->list_cap = np;
  }
  if (m->xset_cap == 0) {
    int64_t cap = 8 * np + (1 << 20);
    int rc = m->xset.ensure(sizeof(int32_t) * cap);
    if (rc) return rc;
    m->xset_cap = cap;
  }
  carve_lists(m->lists.as<char>(), m->list_cap, m->xset.as<int32_t>(), m->xset_cap, a.L);
  if (big_grid(m, a.q) > 0) {
    int rc = m->big_ws.ensure(big_ws_bytes(big_grid(m, a.q)));
    if (rc) return rc;
  }
  return 0;
}

int read_status(sprg_mesh* m, const FitArgs& a, int32_t* st)
{
  hipStream_t s = m->ctx->stream;
  // counts [0, 2 NBUCKET + 2), the status word and the exact-expansion
  // count follow each other
  int32_t c[2 * NBUCKET + 6];
  HIP_OK(hipMemcpyAsync(c, a.L.count, sizeof c, hipMemcpyDeviceToHost, s), "status");
  HIP_OK(hipStreamSynchronize(s), "recover sync");
  *st = c[2 * NBUCKET + 2];
  int64_t expanded = 0;
  for (int b = 0; b < NBUCKET; ++b) expanded += c[NBUCKET + b];
  m->last_stats[0] = a.npatch;
  m->last_stats[1] = c[2 * NBUCKET];
  m->last_stats[2] = expanded;
  m->last_stats[3] = c[2 * NBUCKET + 1];
  m->last_stats[4] = *st;
  m->last_stats[5] = c[2 * NBUCKET + 3];
  m->last_stats[6] = c[2 * NBUCKET + 5];
  return 0;
}

bool valid_form(int form) { return form == SPRG_FP64 || form == SPRG_FP32R || form == SPRG_FP64F; }

// Enqueues one recovery on stream s: kernels only (no allocation, no host
// synchronisation, counters cleared by a kernel rather than a memset node),
// so the sequence can be captured into a HIP graph.  FP64 form: classify ->
// exact fit of the initial sets -> expansion -> exact fit of the expanded
// sets.  Fast forms: fit of every patch -> expansion -> fit of the expanded
// sets -> FP64 fallback -> exact refits.  marks: record the phase events.
int enqueue_recover(sprg_mesh* m, const FitArgs& a, int form, hipStream_t s, bool marks)
{
  const bool fast = form != SPRG_FP64;
  const int nterms = m->order == 1 ? 4 : 10;
  int rc;
  // counters, status, the exact-expansion count and the fallback start
  // (the status words follow the counters: bits, exact count, fallback start, big count)
  if ((rc = launch_copy_i32(a.L.count, nullptr, 2 * NBUCKET + 6, s))) return rc;
  if ((rc = launch_copy_i32(reinterpret_cast<int32_t*>(a.L.x_cursor), nullptr, 4, s))) return rc;
  if (marks) {
    m->phase_last = 0;
    m->mark(0);
  }
  if (fast) {
    if (a.tiles.ntile > 0 && a.shape.tile_blocks > 0) {
      // the tiled regular pass, then the patches outside every tile
      if ((rc = launch_tile_fit(a, nterms, form, s))) return rc;
      if (a.tiles.nleftover > 0 && (rc = launch_fast_fit(a, nterms, form, 4, s))) return rc;
    } else if ((rc = launch_fast_fit(a, nterms, form, 0, s))) {
      return rc;
    }
    if (form == SPRG_FP32R) {
      // the FP64 pass over the FP32 pass's fallback queue, before the
      // expansion so that sets it cannot certify still reach the rank test;
      // the later fallback pass starts after these entries
      if ((rc = launch_fast_fit(a, nterms, form, 3, s))) return rc;
      if ((rc = launch_copy_i32(a.L.status + 2, a.L.count + 2 * NBUCKET + 1, 1, s))) return rc;
    }
  } else {
    if ((rc = launch_classify(a, nterms, s))) return rc;
    if ((rc = launch_exact_fit(a, nterms, 0, s))) return rc;
  }
  if (marks) m->mark(1);
  if ((rc = launch_expand(a, nterms, s))) return rc;
  if (marks) m->mark(2);
  if (fast) {
    if ((rc = launch_fast_fit(a, nterms, form, 1, s))) return rc;
    if ((rc = launch_fast_fit(a, nterms, form, 2, s))) return rc;
    // poorly conditioned sets: the reference's arithmetic
    if ((rc = launch_exact_fit(a, nterms, 2, s))) return rc;
  } else {
    if ((rc = launch_exact_fit(a, nterms, 1, s))) return rc;
  }
  // sets above 192 rows, from every pass above
  if ((rc = launch_exact_big(a, nterms, m->big_ws.as<double>(), big_grid(m, a.q), s))) return rc;
  if (marks) m->mark(3);
  return 0;
}

FitArgs fit_args(sprg_mesh* m, int q, int nc, const double* ip, double* out, int form)
{
  FitArgs a;
  a.shape = m->shape;
  a.m = m->view();
  a.tiles = m->tiles();
  a.fast = form != SPRG_FP64 ? 1 : 0;
  a.q = q;
  a.nc = nc;
  a.npatch = m->npatch();
  a.ip = ip;
  a.out = out;
  return a;
}

// Status word of a finished recovery -> 0 or an SPRG_E* code with message;
// *retry is set when the expanded-set storage overflowed (after growing it).
int status_code(sprg_mesh* m, int32_t st, bool* retry)
{
  *retry = false;
  if (st & ST_XSET_OVERFLOW) {
    int64_t cap = 2 * m->xset_cap;
    int rc = m->xset.ensure(sizeof(int32_t) * cap);
    if (rc) return rc;
    m->xset_cap = cap;
    *retry = true;
    return 0;
  }
  if (st & ST_PATCH_FAIL) return sprg_fail(SPRG_EPATCH, "SPR: patch construction: all hope is lost.");
  if (st & ST_TOO_LARGE)
    return sprg_fail(SPRG_EUNSUPPORTED,
                     "SPR: a patch set needs more than %d sample rows, or a set that must grow has more than %d",
                     BIG_MMAX, bucket_rows(NBUCKET - 1));
  if (st & ST_INTERNAL) return sprg_fail(SPRG_EINTERNAL, "SPR: inconsistent rank decision");
  if (st & ST_ILLCOND)
    return sprg_fail(SPRG_EUNSUPPORTED,
                     "SPR: a patch is too ill-conditioned for the fast forms; use SPRG_FP64");
  return 0;
}

int run_recover(sprg_mesh* m, int q, int nc, const double* ip, double* out, int form)
{
  if (!valid_form(form)) return sprg_fail(SPRG_EINVAL, "unknown numeric form %d", form);
  if (form != SPRG_FP64 && q > 4)
    return sprg_fail(SPRG_EUNSUPPORTED, "the fast forms take at most 4 points per element");
  hipStream_t s = m->ctx->stream;
  for (int attempt = 0; attempt < 4; ++attempt) {
    FitArgs a = fit_args(m, q, nc, ip, out, form);
    int rc = setup_lists(m, a);
    if (rc) return rc;
    if ((rc = enqueue_recover(m, a, form, s, true))) return rc;
    int32_t st = 0;
    if ((rc = read_status(m, a, &st))) return rc;
    collect_phases(m);
    bool retry;
    if ((rc = status_code(m, st, &retry))) return rc;
    if (!retry) return 0;
  }
  return sprg_fail(SPRG_ENOMEM, "SPR: expanded patch storage kept overflowing");
}

// The recovery with its work lists in the device scratch array: no
// allocation, no growth.  The lists take list_bytes(npatch); the rest of the
// scratch holds the expanded sets.
int run_recover_noalloc(sprg_mesh* m, int q, int nc, const double* ip, double* out, int form)
{
  if (!valid_form(form)) return sprg_fail(SPRG_EINVAL, "unknown numeric form %d", form);
  if (form != SPRG_FP64 && q > 4)
    return sprg_fail(SPRG_EUNSUPPORTED, "the fast forms take at most 4 points per element");
  char* base = static_cast<char*>(m->ctx->scratch);
  if (!base) return sprg_fail(SPRG_EHIP, "recover_noalloc: no scratch on this context's device");
  const int64_t np = m->npatch();
  const size_t lists = (list_bytes(np) + 255) / 256 * 256;
  // at least one expanded-set entry per patch plus slack beyond the lists
  const size_t need = lists + sizeof(int32_t) * ((size_t)np + (1 << 18));
  if (need > SPRG_NOALLOC_SCRATCH_BYTES)
    return sprg_fail(SPRG_EINVAL,
                     "recover_noalloc: %lld patches need %zu bytes of scratch, above the %zu this entry point "
                     "holds (use sprg_recover)",
                     (long long)np, need, (size_t)SPRG_NOALLOC_SCRATCH_BYTES);
  if (big_grid(m, q) > 0)
    return sprg_fail(SPRG_EUNSUPPORTED,
                     "recover_noalloc: patch sets above %d rows need the workspace of sprg_recover",
                     bucket_rows(NBUCKET - 1));
  FitArgs a = fit_args(m, q, nc, ip, out, form);
  const int64_t x_cap = (int64_t)((SPRG_NOALLOC_SCRATCH_BYTES - lists) / sizeof(int32_t));
  carve_lists(base, np, reinterpret_cast<int32_t*>(base + lists), x_cap, a.L);
  hipStream_t s = m->ctx->stream;
  int rc = enqueue_recover(m, a, form, s, true);
  if (rc) return rc;
  int32_t st = 0;
  if ((rc = read_status(m, a, &st))) return rc;
  collect_phases(m);
  if (st & ST_XSET_OVERFLOW)
    return sprg_fail(SPRG_ENOMEM, "recover_noalloc: the expanded patch sets overflowed the scratch (use sprg_recover)");
  bool retry;
  return status_code(m, st, &retry);
}

}  // namespace

extern "C" int sprg_grad_ip(sprg_mesh* m, int field_order, int nc_in, const double* nodal,
                            int ip_order, double* ip)
{
  if (!m || !nodal || !ip) return sprg_fail(SPRG_EINVAL, "grad_ip: null argument");
  if (nc_in != 1 && nc_in != 3 && nc_in != 9)
    return sprg_fail(SPRG_EINVAL, "getGradIPField: value type must be scalar, vector or matrix");
  if (field_order != 1 && field_order != 2)
    return sprg_fail(SPRG_EUNSUPPORTED, "getGradIPField: field order %d (1 or 2 supported)", field_order);
  if (field_order == 2 && !m->ne)
    return sprg_fail(SPRG_EINVAL, "getGradIPField: an order-2 field needs mesh edges");
  if (rule_npoints(ip_order) < 0)
    return sprg_fail(SPRG_EUNSUPPORTED, "getGradIPField: integration order %d (1..3 supported)", ip_order);
  HIP_OK(hipSetDevice(m->ctx->device), "hipSetDevice");
  m->phase_last = 0;
  m->mark(0);
  int rc = launch_grad(m->view(), field_order, nc_in, nodal, ip_order, ip, m->ctx->stream);
  if (rc) return rc;
  m->mark(1);
  HIP_OK(hipStreamSynchronize(m->ctx->stream), "grad sync");
  collect_phases(m);  // phase 0: the gradient kernel
  return 0;
}

extern "C" int sprg_recover(sprg_mesh* m, int ip_order, int nc, const double* ip, double* nodal,
                            int form)
{
  if (!m || !ip || !nodal) return sprg_fail(SPRG_EINVAL, "recover: null argument");
  int q, rc = check_ip(m, ip_order, nc, &q);
  if (rc) return rc;
  HIP_OK(hipSetDevice(m->ctx->device), "hipSetDevice");
  return run_recover(m, q, nc, ip, nodal, form);
}

extern "C" int sprg_recover_noalloc(sprg_mesh* m, int ip_order, int nc, const double* ip,
                                    double* nodal, int form)
{
  if (!m || !ip || !nodal) return sprg_fail(SPRG_EINVAL, "recover_noalloc: null argument");
  int q, rc = check_ip(m, ip_order, nc, &q);
  if (rc) return rc;
  HIP_OK(hipSetDevice(m->ctx->device), "hipSetDevice");
  return run_recover_noalloc(m, q, nc, ip, nodal, form);
}

namespace {

// Buffers of one size-field call (recovered field, element sizes, block
// partials, scalars): allocated before any capture.
int estimate_buffers(sprg_mesh* m, int nc)
{
  int rc;
  const int64_t nblocks = (m->nt + 255) / 256;
  if ((rc = m->rec.ensure(sizeof(double) * m->nnodes * nc))) return rc;
  if ((rc = m->elem.ensure(sizeof(double) * 2 * m->nt))) return rc;
  if ((rc = m->partial.ensure(sizeof(double) * 2 * nblocks))) return rc;
  if ((rc = m->scalars.ensure(sizeof(double) * 4))) return rc;
  return 0;
}

EstArgs est_args(sprg_mesh* m, int ip_order, int nc, const double* ip, int target, double ar,
                 double tn, double alpha, double beta, double* size)
{
  EstArgs e;
  e.m = m->view();
  e.ip_order = ip_order;
  e.nc = nc;
  e.ip = ip;
  e.rec = m->rec.as<double>();
  e.target = target;
  e.adapt_ratio = ar;
  e.target_n = tn;
  e.alpha = alpha;
  e.beta = beta;
  e.elem = m->elem.as<double>();
  e.partial = m->partial.as<double>();
  e.scalars = m->scalars.as<double>();
  e.size = size;
  e.nblocks = (int)((m->nt + 255) / 256);
  return e;
}

int run_estimate(sprg_mesh* m, int ip_order, int nc, const double* ip, int target,
                 double ar, double tn, double alpha, double beta, double* size, int form,
                 double* diag)
{
  auto t0 = std::chrono::steady_clock::now();
  int q, rc = check_ip(m, ip_order, nc, &q);
  if (rc) return rc;
  if (nc != 1 && nc != 3 && nc != 9 && nc != 27)
    return sprg_fail(SPRG_EUNSUPPORTED, "estimator: %d components", nc);
  HIP_OK(hipSetDevice(m->ctx->device), "hipSetDevice");
  if ((rc = estimate_buffers(m, nc))) return rc;
  if ((rc = run_recover(m, q, nc, ip, m->rec.as<double>(), form))) return rc;
  EstArgs e = est_args(m, ip_order, nc, ip, target, ar, tn, alpha, beta, size);
  if ((rc = launch_estimate(e, m->ctx->stream, [m](int i) { m->mark(i); }))) return rc;
  double sc[4] = {0, 0, 0, 0};
  HIP_OK(hipMemcpyAsync(sc, e.scalars, sizeof(double) * 3, hipMemcpyDeviceToHost, m->ctx->stream),
         "scalars");
  HIP_OK(hipStreamSynchronize(m->ctx->stream), "estimate sync");
  collect_phases(m);
  auto t1 = std::chrono::steady_clock::now();
  if (diag) {
    if (!target) {
      diag[0] = sc[0];
      diag[1] = sc[1];
      diag[2] = sc[2];
    } else {
      diag[0] = sc[1];
      diag[1] = sc[2];
      diag[2] = 0;
    }
    diag[3] = std::chrono::duration<double, std::milli>(t1 - t0).count();
  }
  return 0;
}

}  // namespace

extern "C" int sprg_spr_size(sprg_mesh* m, int ip_order, int nc, const double* ip,
                             double adapt_ratio, double* size, int form, double* diag)
{
  if (!m || !ip || !size) return sprg_fail(SPRG_EINVAL, "spr_size: null argument");
  return run_estimate(m, ip_order, nc, ip, 0, adapt_ratio, 0, 0, 0, size, form, diag);
}

// Sets the launch shapes of the patch kernels (for instance from a cache of
// earlier sprg_mesh_autotune results).
extern "C" int sprg_mesh_set_shape(sprg_mesh* m, const int64_t* shape, int n)
{
  if (!m || !shape || n < 4) return sprg_fail(SPRG_EINVAL, "set_shape: needs 4 or 5 values");
  if (shape[0] != 64 && shape[0] != 128 && shape[0] != 256)
    return sprg_fail(SPRG_EINVAL, "set_shape: fast-fit block size must be 64, 128 or 256");
  for (int i = 1; i < 4; ++i)
    if (shape[i] < 1 || shape[i] > (1 << 20))
      return sprg_fail(SPRG_EINVAL, "set_shape: grid sizes must lie in [1, 2^20]");
  if (n >= 5 && (shape[4] < 0 || shape[4] > (1 << 20)))
    return sprg_fail(SPRG_EINVAL, "set_shape: the tiled grid must lie in [0, 2^20]");
  m->shape.fast_threads = (int)shape[0];
  m->shape.fast_blocks = (int)shape[1];
  m->shape.exact_blocks = (int)shape[2];
  m->shape.expand_blocks = (int)shape[3];
  if (n >= 5) m->shape.tile_blocks = (int)shape[4];
  return SPRG_OK;
}

// Times recoveries of the field with candidate launch shapes, one parameter
// at a time (coordinate descent, each candidate the best of two runs), and
// keeps the fastest shape for the mesh's later calls.
extern "C" int sprg_mesh_autotune(sprg_mesh* m, int ip_order, int nc, const double* ip, int form,
                                  int64_t* shape, int n)
{
  if (!m || !ip) return sprg_fail(SPRG_EINVAL, "autotune: null argument");
  int q, rc = check_ip(m, ip_order, nc, &q);
  if (rc) return rc;
  HIP_OK(hipSetDevice(m->ctx->device), "hipSetDevice");
  if ((rc = m->rec.ensure(sizeof(double) * m->nnodes * nc))) return rc;
  auto time_once = [&](double* ms) -> int {
    auto t0 = std::chrono::steady_clock::now();
    int r = run_recover(m, q, nc, ip, m->rec.as<double>(), form);
    *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return r;
  };
  auto measure = [&](double* best) -> int {
    double a = 0, b = 0;
    int r = time_once(&a);
    if (!r) r = time_once(&b);
    *best = a < b ? a : b;
    return r;
  };
  double best_ms;
  if ((rc = measure(&best_ms))) return rc;
  const bool fast = form != SPRG_FP64;
  struct Param {
    int LaunchShape::*field;
    std::vector<int> values;
  };
  std::vector<Param> params;
  if (fast) {
    // the tiled pass is kept on or off as it is: the two regular passes sum
    // the rows of a patch in different orders
    if (m->shape.tile_blocks > 0 && m->ntile > 0)
      params.push_back({&LaunchShape::tile_blocks, {1024, 2048, 4096, 8192, 16384}});
    params.push_back({&LaunchShape::fast_threads, {64, 128, 256}});
    params.push_back({&LaunchShape::fast_blocks, {2048, 4096, 8192, 16384, 32768}});
  } else {
    params.push_back({&LaunchShape::exact_blocks, {2048, 4096, 8192, 16384}});
  }
  params.push_back({&LaunchShape::expand_blocks, {1024, 2048, 4096, 8192}});
  for (const Param& prm : params) {
    const int keep = m->shape.*prm.field;
    int chosen = keep;
    for (int v : prm.values) {
      if (v == keep) continue;
      m->shape.*prm.field = v;
      double ms;
      if ((rc = measure(&ms))) {
        m->shape.*prm.field = keep;
        return rc;
      }
      if (ms < best_ms) {
        best_ms = ms;
        chosen = v;
      }
    }
    m->shape.*prm.field = chosen;
  }
  const int64_t out[5] = {m->shape.fast_threads, m->shape.fast_blocks, m->shape.exact_blocks,
                          m->shape.expand_blocks, m->shape.tile_blocks};
  for (int i = 0; shape && i < n && i < 5; ++i) shape[i] = out[i];
  return SPRG_OK;
}

extern "C" int sprg_average_to_vertex(sprg_mesh* m, const double* elem, double* vert)
{
  if (!m || !elem || !vert) return sprg_fail(SPRG_EINVAL, "average_to_vertex: null argument");
  HIP_OK(hipSetDevice(m->ctx->device), "hipSetDevice");
  int rc = launch_vertex_mean(m->view(), elem, vert, m->ctx->stream);
  if (rc) return rc;
  HIP_OK(hipStreamSynchronize(m->ctx->stream), "average sync");
  return 0;
}

extern "C" int sprg_target_spr_size(sprg_mesh* m, int ip_order, int nc, const double* ip,
                                    uint64_t target, double alpha, double beta, double* size,
                                    int form, double* diag)
{
  if (!m || !ip || !size) return sprg_fail(SPRG_EINVAL, "target_spr_size: null argument");
  // PCU_ALWAYS_ASSERT(target > 0), PCU_ALWAYS_ASSERT(alpha < beta)
  // (spr/sprEstimateTargetError.cc:249-250)
  if (target == 0) return sprg_fail(SPRG_EINVAL, "getTargetSPRSizeField: target must be > 0");
  if (!(alpha < beta)) return sprg_fail(SPRG_EINVAL, "getTargetSPRSizeField: alpha must be < beta");
  // double N = e->target_number (spr/sprEstimateTargetError.cc:135): the
  // same size_t -> double conversion
  return run_estimate(m, ip_order, nc, ip, 1, 0, (double)target, alpha, beta, size, form, diag);
}

// ---------------------------------------------------------------------------
// patch sets (for parity checks of the patch construction)

namespace {

__global__ void k_set_len(FitArgs a, int32_t* len, int32_t* xmap)
{
  int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= a.npatch) return;
  int32_t k = xmap[p];
  if (k >= 0) {
    len[p] = a.L.x_len[k];
  } else if (p < a.m.nv) {
    len[p] = a.m.v2t_off[p + 1] - a.m.v2t_off[p];
  } else {
    len[p] = a.m.e2t_off[p - a.m.nv + 1] - a.m.e2t_off[p - a.m.nv];
  }
}

__global__ void k_xmap(FitArgs a, int32_t* xmap, int64_t nitems)
{
  int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= nitems || a.L.x_len[k] <= 0) return;  // no set found for this queued patch
  xmap[a.L.x_patch[k]] = (int32_t)k;
}

__global__ void k_set_fill(FitArgs a, const int32_t* incl, const int32_t* xmap,
                           int64_t* set_off, int32_t* set_idx, int64_t cap)
{
  int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= a.npatch) return;
  int64_t beg = p ? incl[p - 1] : 0, end = incl[p];
  set_off[p] = beg;
  if (p == a.npatch - 1) set_off[p + 1] = end;
  if (!set_idx || end > cap) return;
  const int32_t* src;
  int32_t k = xmap[p];
  if (k >= 0)
    src = a.L.x_set + a.L.x_off[k];
  else if (p < a.m.nv)
    src = a.m.v2t + a.m.v2t_off[p];
  else
    src = a.m.e2t + a.m.e2t_off[p - a.m.nv];
  for (int64_t i = 0; i < end - beg; ++i) set_idx[beg + i] = src[i];
}

}  // namespace

extern "C" int sprg_patch_sets(sprg_mesh* m, int ip_order, int64_t* set_off, int32_t* set_idx,
                               int64_t cap, int64_t* total)
{
  if (!m || !set_off || !total) return sprg_fail(SPRG_EINVAL, "patch_sets: null argument");
  int q, rc = check_ip(m, ip_order, 0, &q);
  if (rc) return rc;
  HIP_OK(hipSetDevice(m->ctx->device), "hipSetDevice");
  // the fit with zero components runs the rank tests and the expansion only
  double dummy_out = 0;
  (void)dummy_out;
  if ((rc = m->rec.ensure(sizeof(double)))) return rc;
  if ((rc = run_recover(m, q, 0, m->rec.as<double>(), m->rec.as<double>(), SPRG_FP64))) return rc;
  FitArgs a;
  a.m = m->view();
  a.fast = 0;
  a.q = q;
  a.nc = 0;
  a.npatch = m->npatch();
  if ((rc = setup_lists(m, a))) return rc;
  hipStream_t s = m->ctx->stream;
  int64_t nitems = 0;
  HIP_OK(hipMemcpyAsync(&nitems, a.L.x_cursor + 1, sizeof(int64_t), hipMemcpyDeviceToHost, s), "items");
  HIP_OK(hipStreamSynchronize(s), "sync");
  const int64_t np = a.npatch;
  if ((rc = m->setlen.ensure(sizeof(int32_t) * (3 * np + 1024)))) return rc;
  int32_t* len = m->setlen.as<int32_t>();
  int32_t* xmap = len + np;
  int32_t* work = xmap + np;
  HIP_OK(hipMemsetAsync(xmap, 0xff, sizeof(int32_t) * np, s), "memset");
  unsigned g = (unsigned)((np + 255) / 256);
  if (nitems) k_xmap<<<(unsigned)((nitems + 255) / 256), 256, 0, s>>>(a, xmap, nitems);
  k_set_len<<<g, 256, 0, s>>>(a, len, xmap);
  if ((rc = launch_scan_i32(len, np, work, s))) return rc;
  int32_t tot = 0;
  HIP_OK(hipMemcpyAsync(&tot, len + np - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s), "total");
  HIP_OK(hipStreamSynchronize(s), "sync");
  *total = tot;
  k_set_fill<<<g, 256, 0, s>>>(a, len, xmap, set_off, set_idx, cap);
  HIP_OK(hipStreamSynchronize(s), "sync");
  return 0;
}

// ---------------------------------------------------------------------------
// Batches: the size-field pipelines of several meshes recorded into one HIP
// graph and replayed.

struct sprg_batch {
  sprg_ctx* ctx = nullptr;
  std::vector<sprg_mesh*> meshes;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<uint64_t> key;  // what the recorded graph was built for
  DBuf bstat;                 // [nmesh] status word of each mesh's recovery
  void drop()
  {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    key.clear();
  }
};

extern "C" int sprg_batch_create(sprg_ctx* ctx, int nmesh, sprg_mesh* const* meshes,
                                 sprg_batch** out)
{
  if (!ctx || !out || nmesh <= 0 || !meshes) return sprg_fail(SPRG_EINVAL, "batch_create: bad argument");
  *out = nullptr;
  for (int i = 0; i < nmesh; ++i) {
    if (!meshes[i]) return sprg_fail(SPRG_EINVAL, "batch_create: mesh %d is null", i);
    if (meshes[i]->ctx != ctx)
      return sprg_fail(SPRG_EINVAL, "batch_create: mesh %d belongs to another context", i);
    for (int j = 0; j < i; ++j)
      if (meshes[j] == meshes[i])
        return sprg_fail(SPRG_EINVAL, "batch_create: mesh %d appears twice (meshes own their workspaces)", i);
  }
  HIP_OK(hipSetDevice(ctx->device), "hipSetDevice");
  sprg_batch* b = new (std::nothrow) sprg_batch;
  if (!b) return sprg_fail(SPRG_ENOMEM, "batch_create: out of host memory");
  b->ctx = ctx;
  b->meshes.assign(meshes, meshes + nmesh);
  int rc = b->bstat.ensure(sizeof(int32_t) * nmesh);
  if (rc) {
    delete b;
    return rc;
  }
  *out = b;
  return 0;
}

extern "C" int sprg_batch_destroy(sprg_batch* b)
{
  if (!b) return 0;
  (void)hipSetDevice(b->ctx->device);
  (void)hipStreamSynchronize(b->ctx->stream);
  b->drop();
  b->bstat.release();
  delete b;
  return 0;
}

namespace {

uint64_t bits(double x)
{
  uint64_t u;
  memcpy(&u, &x, sizeof u);
  return u;
}

// Records the pipelines of every mesh of the batch on the context stream.
// target: NULL (getSPRSizeField for every mesh) or one element count per
// mesh (getTargetSPRSizeField).
int record_batch(sprg_batch* b, int ip_order, int nc, const double* const* ip,
                 const uint64_t* target, double ar, double alpha, double beta,
                 double* const* size, int form)
{
  hipStream_t s = b->ctx->stream;
  const int n = (int)b->meshes.size();
  int q = rule_npoints(ip_order);
  // every allocation happens before the capture
  std::vector<FitArgs> fa(n);
  for (int i = 0; i < n; ++i) {
    sprg_mesh* m = b->meshes[i];
    int rc = estimate_buffers(m, nc);
    if (!rc) {
      fa[i] = fit_args(m, q, nc, ip[i], m->rec.as<double>(), form);
      rc = setup_lists(m, fa[i]);
    }
    if (rc) return rc;
  }
  HIP_OK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
  int rc = 0;
  for (int i = 0; i < n && !rc; ++i) {
    sprg_mesh* m = b->meshes[i];
    rc = enqueue_recover(m, fa[i], form, s, false);
    if (!rc) {
      const bool tgt = target != nullptr;
      EstArgs e = est_args(m, ip_order, nc, ip[i], tgt ? 1 : 0, ar, tgt ? (double)target[i] : 0,
                           alpha, beta, size[i]);
      rc = launch_estimate(e, s, [](int) {});
    }
    if (!rc) rc = launch_copy_i32(b->bstat.as<int32_t>() + i, fa[i].L.status, 1, s);
  }
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(s, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess || !g) return sprg_fail(SPRG_EHIP, "batch: capture failed: %s", hipGetErrorString(e));
  b->graph = g;
  if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
    b->drop();
    return sprg_fail(SPRG_EHIP, "batch: graph instantiation failed");
  }
  return 0;
}

// What a recorded graph depends on beyond the call's arguments: the launch
// shapes and every device address its kernels were given.  Workspaces only
// grow, so they are sized for this call first (no-ops when they fit); an
// eager call that grew one since the recording shows up as a new address.
int workspace_key(sprg_batch* b, int q, int nc, int form, std::vector<uint64_t>& key)
{
  for (sprg_mesh* m : b->meshes) {
    int rc = estimate_buffers(m, nc);
    FitArgs a = fit_args(m, q, nc, nullptr, nullptr, form);
    if (!rc) rc = setup_lists(m, a);
    if (rc) return rc;
    const DBuf* bufs[] = {&m->rec, &m->lists, &m->xset, &m->elem, &m->partial, &m->scalars, &m->big_ws};
    for (const DBuf* d : bufs) {
      key.push_back((uint64_t)(uintptr_t)d->p);
      key.push_back((uint64_t)d->bytes);
    }
    key.push_back((uint64_t)m->list_cap);
    key.push_back((uint64_t)m->xset_cap);
    key.push_back((uint64_t)m->shape.fast_threads);
    key.push_back((uint64_t)m->shape.fast_blocks);
    key.push_back((uint64_t)m->shape.exact_blocks);
    key.push_back((uint64_t)m->shape.expand_blocks);
    key.push_back((uint64_t)m->shape.tile_blocks);
  }
  return 0;
}

}  // namespace

extern "C" int sprg_batch_size(sprg_batch* b, int ip_order, int nc, const double* const* ip,
                               const uint64_t* target, double adapt_ratio, double alpha, double beta,
                               double* const* size, int form, double* ms)
{
  if (!b || !ip || !size) return sprg_fail(SPRG_EINVAL, "batch_size: null argument");
  if (!valid_form(form)) return sprg_fail(SPRG_EINVAL, "unknown numeric form %d", form);
  if (nc != 1 && nc != 3 && nc != 9 && nc != 27)
    return sprg_fail(SPRG_EUNSUPPORTED, "estimator: %d components", nc);
  const int n = (int)b->meshes.size();
  int q = 0;
  for (int i = 0; i < n; ++i) {
    int rc = check_ip(b->meshes[i], ip_order, nc, &q);
    if (rc) return rc;
    if (form != SPRG_FP64 && q > 4)
      return sprg_fail(SPRG_EUNSUPPORTED, "the fast forms take at most 4 points per element");
    if (!ip[i] || !size[i]) return sprg_fail(SPRG_EINVAL, "batch_size: null field of mesh %d", i);
    // PCU_ALWAYS_ASSERT(target > 0) (spr/sprEstimateTargetError.cc:249)
    if (target && target[i] == 0)
      return sprg_fail(SPRG_EINVAL, "getTargetSPRSizeField: target of mesh %d must be > 0", i);
  }
  if (target && !(alpha < beta))
    return sprg_fail(SPRG_EINVAL, "getTargetSPRSizeField: alpha must be < beta");
  HIP_OK(hipSetDevice(b->ctx->device), "hipSetDevice");
  hipStream_t s = b->ctx->stream;
  for (int attempt = 0; attempt < 4; ++attempt) {
    std::vector<uint64_t> key = {(uint64_t)ip_order, (uint64_t)nc, (uint64_t)form, bits(adapt_ratio),
                                 bits(alpha), bits(beta), target ? 1u : 0u};
    for (int i = 0; i < n; ++i) {
      key.push_back((uint64_t)(uintptr_t)ip[i]);
      key.push_back((uint64_t)(uintptr_t)size[i]);
      key.push_back(target ? target[i] : 0);
    }
    int rc = workspace_key(b, q, nc, form, key);
    if (rc) return rc;
    if (!b->exec || key != b->key) {
      b->drop();
      rc = record_batch(b, ip_order, nc, ip, target, adapt_ratio, alpha, beta, size, form);
      if (rc) return rc;
      b->key = key;
    }
    auto t0 = std::chrono::steady_clock::now();
    HIP_OK(hipGraphLaunch(b->exec, s), "graph launch");
    HIP_OK(hipStreamSynchronize(s), "batch sync");
    auto t1 = std::chrono::steady_clock::now();
    if (ms) *ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    std::vector<int32_t> st(n);
    HIP_OK(hipMemcpy(st.data(), b->bstat.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost), "status");
    bool again = false;
    for (int i = 0; i < n; ++i) {
      bool retry;
      rc = status_code(b->meshes[i], st[i], &retry);
      if (rc) {
        const std::string msg = sprg_last_error();
        return sprg_fail(rc, "batch mesh %d: %s", i, msg.c_str());
      }
      again = again || retry;
    }
    if (!again) return 0;
    b->drop();  // a mesh's expanded-set storage grew: record again
  }
  return sprg_fail(SPRG_ENOMEM, "SPR: expanded patch storage kept overflowing");
}
