// smb.cpp -- reader of MDS ".smb" mesh files (host code): one part of
// tetrahedra into the arrays of sprg_mesh_create.
//
// Format (mds/mds_smb.c:562-606, versions 0..6): big-endian 32-bit unsigned
// integers and 64-bit doubles (pcu/pcu_io.c:242-336).
//   header     magic, version, dim, number of parts (:121-134)
//   counts     entities of the 8 SMB types: vertex, edge, triangle, quad,
//              hex, prism, pyramid, tet (:27-37)
//   down       for each type but the vertex, in that order, the indices of
//              each entity's entities one dimension down (:159-186): an
//              edge's 2 vertices, a triangle's 3 edges, a tet's 4 triangles
//   coords     3 doubles per vertex; version >= 2: 2 parametric doubles
//              per vertex (:586-594)
//   then remote copies, model classification, tags, matches and (version
//   >= 5) field metadata, whose first entry is the name of the coordinate
//   field's shape (apf/apfFile.cc:102-107).
//
// A tet stores its triangles only; MDS derives its edges and vertices from
// them (mds/mds.c:652-686, tables T21 and T10 at :76-77): edge i is the
// edge shared by triangles T21[2i] and T21[2i+1], vertex i the vertex shared
// by edges T10[2i] and T10[2i+1].  The arrays here follow the same rules, so
// a tet's vertex and edge order is the order apf::Mesh::getDownward returns
// for it, entity indices are the file's (MDS iteration order), and a
// Lagrange-2 mirror puts its edge nodes at the edge midpoints (the linear
// coordinate field projected by apf::changeShape).
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "sprgpu_internal.h"

namespace {

enum { SMB_VERT, SMB_EDGE, SMB_TRI, SMB_QUAD, SMB_HEX, SMB_PRIS, SMB_PYR, SMB_TET, SMB_TYPES };
const char* const kTypeName[SMB_TYPES] = {"vertex", "edge", "triangle", "quad", "hex", "prism", "pyramid", "tet"};
// entities one dimension down per SMB type
const int kDown[SMB_TYPES] = {0, 2, 3, 4, 6, 5, 5, 4};
constexpr unsigned kMaxVersion = 6;
constexpr unsigned kMaxEntities = 100u * 1000u * 1000u;  // as mds_smb.c:50
constexpr unsigned kMaxTags = 100;

// Tet edge i lies on triangles kT21[2i], kT21[2i+1]; tet vertex i on edges
// kT10[2i], kT10[2i+1] (mds/mds.c:76-77, 90-95).
const int kT21[12] = {0, 1, 0, 2, 0, 3, 1, 3, 1, 2, 2, 3};
const int kT10[8] = {2, 0, 0, 1, 1, 2, 3, 4};

struct Reader {
  FILE* f = nullptr;
  std::string path;
  ~Reader()
  {
    if (f) fclose(f);
  }
  int open(const char* p)
  {
    path = p;
    f = fopen(p, "rb");
    if (!f) return sprg_fail(SPRG_EINVAL, "smb: cannot open \"%s\"", p);
    return 0;
  }
  int bytes(void* dst, size_t n)
  {
    if (n && fread(dst, 1, n, f) != n) return sprg_fail(SPRG_EINVAL, "smb: \"%s\" ends early", path.c_str());
    return 0;
  }
  int u32(unsigned* dst, size_t n)
  {
    int rc = bytes(dst, 4 * n);
    if (rc) return rc;
    for (size_t i = 0; i < n; ++i) dst[i] = __builtin_bswap32(dst[i]);
    return 0;
  }
  int f64(double* dst, size_t n)
  {
    int rc = bytes(dst, 8 * n);
    if (rc) return rc;
    for (size_t i = 0; i < n; ++i) {
      uint64_t u;
      memcpy(&u, dst + i, 8);
      u = __builtin_bswap64(u);
      memcpy(dst + i, &u, 8);
    }
    return 0;
  }
  int skip(size_t n)
  {
    if (n && fseek(f, (long)n, SEEK_CUR)) return sprg_fail(SPRG_EINVAL, "smb: \"%s\" ends early", path.c_str());
    return 0;
  }
  int str(std::string* s)
  {
    s->clear();
    for (;;) {
      const int c = fgetc(f);
      if (c == EOF) return sprg_fail(SPRG_EINVAL, "smb: \"%s\" ends inside a string", path.c_str());
      if (c == 0) return 0;
      if (s->size() > 4096) return sprg_fail(SPRG_EINVAL, "smb: \"%s\": string too long", path.c_str());
      s->push_back((char)c);
    }
  }
  // mds_links (mds/mds_smb.c:89-106): peer count, peers, per-peer counts,
  // per-peer entity lists
  int links()
  {
    unsigned np = 0;
    int rc = u32(&np, 1);
    if (rc || !np) return rc;
    if (np >= 10000) return sprg_fail(SPRG_EINVAL, "smb: \"%s\": %u peers", path.c_str(), np);
    std::vector<unsigned> n(np);
    if ((rc = skip(4 * (size_t)np)) || (rc = u32(n.data(), np))) return rc;
    for (unsigned i = 0; i < np; ++i)
      if ((rc = skip(4 * (size_t)n[i]))) return rc;
    return 0;
  }
};

struct Smb {
  unsigned version = 0, dim = 0, parts = 0;
  unsigned n[SMB_TYPES] = {};
  std::vector<unsigned> down[SMB_TYPES];
  std::vector<double> xyz;
  std::string shape = "Linear";
};

// Reads the whole file: header, connectivity, coordinates, and the sections
// after them up to the coordinate field's shape name.
int read_smb(const char* path, Smb& m, bool connectivity)
{
  Reader r;
  int rc = r.open(path);
  if (rc) return rc;
  unsigned h[4];
  if ((rc = r.u32(h, 4))) return rc;
  m.version = h[1];
  m.dim = h[2];
  m.parts = h[3];
  if (m.version > kMaxVersion)
    return sprg_fail(SPRG_EUNSUPPORTED, "smb: \"%s\" is version %u (at most %u read)", path, m.version, kMaxVersion);
  if ((rc = r.u32(m.n, SMB_TYPES))) return rc;
  for (int t = 0; t < SMB_TYPES; ++t)
    if (m.n[t] >= kMaxEntities)
      return sprg_fail(SPRG_EINVAL, "smb: \"%s\": %u entities of type %s", path, m.n[t], kTypeName[t]);
  if (!connectivity) return 0;
  for (int t = 1; t < SMB_TYPES; ++t) {
    m.down[t].resize((size_t)kDown[t] * m.n[t]);
    if ((rc = r.u32(m.down[t].data(), m.down[t].size()))) return rc;
  }
  m.xyz.resize(3 * (size_t)m.n[SMB_VERT]);
  if ((rc = r.f64(m.xyz.data(), m.xyz.size()))) return rc;
  if (m.version >= 2 && (rc = r.skip(16 * (size_t)m.n[SMB_VERT]))) return rc;  // parametric coordinates
  if ((rc = r.links())) return rc;                                               // remote copies
  for (int t = 0; t < SMB_TYPES; ++t)                                            // classification
    if ((rc = r.skip(8 * (size_t)m.n[t]))) return rc;
  // tags (mds/mds_smb.c:280-299, :449-475)
  unsigned ntag = 0;
  if ((rc = r.u32(&ntag, 1))) return rc;
  if (ntag >= kMaxTags) return sprg_fail(SPRG_EINVAL, "smb: \"%s\": %u tags", path, ntag);
  std::vector<unsigned> ttype(ntag), tcount(ntag), sizes(ntag);
  for (unsigned i = 0; i < ntag; ++i) {
    std::string name;
    if ((rc = r.u32(&ttype[i], 1)) || (rc = r.u32(&tcount[i], 1)) || (rc = r.str(&name))) return rc;
    if (ttype[i] > 1) return sprg_fail(SPRG_EINVAL, "smb: \"%s\": tag \"%s\" of type %u", path, name.c_str(), ttype[i]);
  }
  for (int t = 0; t < SMB_TYPES; ++t) {
    if (ntag && (rc = r.u32(sizes.data(), ntag))) return rc;
    for (unsigned j = 0; j < ntag; ++j) {
      const size_t elem = ttype[j] == 0 ? 4 : 8;
      if ((rc = r.skip(4 * (size_t)sizes[j])) || (rc = r.skip(elem * tcount[j] * (size_t)sizes[j]))) return rc;
    }
  }
  // matches (:538-552)
  if (m.version >= 3) {
    const int blocks = m.version >= 4 ? SMB_TYPES : 7;
    for (int i = 0; i < blocks; ++i)
      if ((rc = r.links())) return rc;
  }
  if (m.version >= 5 && (rc = r.str(&m.shape))) return rc;
  return 0;
}

// the one entity index both lists hold (mds/mds.c:499-511 common_down)
int64_t common(const unsigned* a, int na, const unsigned* b, int nb)
{
  for (int i = 0; i < na; ++i)
    for (int j = 0; j < nb; ++j)
      if (a[i] == b[j]) return a[i];
  return -1;
}

int check_tets_only(const char* path, const Smb& m)
{
  std::string other;
  for (int t = SMB_QUAD; t < SMB_TET; ++t)
    if (m.n[t]) other += std::string(other.empty() ? "" : ", ") + std::to_string(m.n[t]) + " " + kTypeName[t];
  if (!other.empty())
    return sprg_fail(SPRG_EUNSUPPORTED, "smb: \"%s\" holds %s elements besides %u tets (tet parts only)", path,
                     other.c_str(), m.n[SMB_TET]);
  if (m.dim != 3) return sprg_fail(SPRG_EUNSUPPORTED, "smb: \"%s\" is a %u-dimensional mesh", path, m.dim);
  if (!m.n[SMB_TET]) return sprg_fail(SPRG_EINVAL, "smb: \"%s\" holds no tets", path);
  if (m.shape != "Linear")
    return sprg_fail(SPRG_EUNSUPPORTED, "smb: \"%s\": coordinate field shape \"%s\" (Linear meshes read)", path,
                     m.shape.c_str());
  return 0;
}

}  // namespace

extern "C" int sprg_smb_counts(const char* path, int64_t* counts, int* version)
{
  if (!path || !counts) return sprg_fail(SPRG_EINVAL, "smb_counts: null argument");
  Smb m;
  int rc = read_smb(path, m, false);
  if (rc) return rc;
  for (int t = 0; t < SMB_TYPES; ++t) counts[t] = m.n[t];
  if (version) *version = (int)m.version;
  return 0;
}

extern "C" int sprg_smb_read(const char* path, int32_t* tet_verts, int32_t* tet_edges, int32_t* edge_verts,
                             double* xyz)
{
  if (!path || !tet_verts || !xyz) return sprg_fail(SPRG_EINVAL, "smb_read: null argument");
  Smb m;
  int rc = read_smb(path, m, true);
  if (rc) return rc;
  if ((rc = check_tets_only(path, m))) return rc;
  const unsigned nv = m.n[SMB_VERT], ne = m.n[SMB_EDGE], nf = m.n[SMB_TRI], nt = m.n[SMB_TET];
  const unsigned* ev = m.down[SMB_EDGE].data();
  const unsigned* fe = m.down[SMB_TRI].data();
  const unsigned* tf = m.down[SMB_TET].data();
  for (size_t i = 0; i < 2 * (size_t)ne; ++i)
    if (ev[i] >= nv) return sprg_fail(SPRG_EINVAL, "smb: \"%s\": edge vertex %u out of range", path, ev[i]);
  for (size_t i = 0; i < 3 * (size_t)nf; ++i)
    if (fe[i] >= ne) return sprg_fail(SPRG_EINVAL, "smb: \"%s\": triangle edge %u out of range", path, fe[i]);
  for (size_t i = 0; i < 4 * (size_t)nt; ++i)
    if (tf[i] >= nf) return sprg_fail(SPRG_EINVAL, "smb: \"%s\": tet triangle %u out of range", path, tf[i]);
  for (unsigned t = 0; t < nt; ++t) {
    unsigned e[6];
    for (int i = 0; i < 6; ++i) {
      const int64_t c = common(fe + 3 * (size_t)tf[4 * (size_t)t + kT21[2 * i]], 3,
                               fe + 3 * (size_t)tf[4 * (size_t)t + kT21[2 * i + 1]], 3);
      if (c < 0) return sprg_fail(SPRG_EINVAL, "smb: \"%s\": tet %u: triangles share no edge", path, t);
      e[i] = (unsigned)c;
    }
    for (int i = 0; i < 4; ++i) {
      const int64_t c = common(ev + 2 * (size_t)e[kT10[2 * i]], 2, ev + 2 * (size_t)e[kT10[2 * i + 1]], 2);
      if (c < 0) return sprg_fail(SPRG_EINVAL, "smb: \"%s\": tet %u: edges share no vertex", path, t);
      tet_verts[4 * (size_t)t + i] = (int32_t)c;
    }
    if (tet_edges)
      for (int i = 0; i < 6; ++i) tet_edges[6 * (size_t)t + i] = (int32_t)e[i];
  }
  if (edge_verts)
    for (size_t i = 0; i < 2 * (size_t)ne; ++i) edge_verts[i] = (int32_t)ev[i];
  memcpy(xyz, m.xyz.data(), sizeof(double) * m.xyz.size());
  return 0;
}

extern "C" int sprg_mesh_load_smb(sprg_ctx* ctx, const char* path, int order, sprg_mesh** out)
{
  if (!ctx || !path || !out) return sprg_fail(SPRG_EINVAL, "mesh_load_smb: null argument");
  *out = nullptr;
  if (order != 1 && order != 2)
    return sprg_fail(SPRG_EUNSUPPORTED, "mesh_load_smb: coordinate order %d (1 or 2)", order);
  int64_t n[SMB_TYPES];
  int rc = sprg_smb_counts(path, n, nullptr);
  if (rc) return rc;
  const int64_t nv = n[SMB_VERT], ne = n[SMB_EDGE], nt = n[SMB_TET];
  std::vector<int32_t> tv(4 * (size_t)nt), te(6 * (size_t)nt), ev(2 * (size_t)ne);
  std::vector<double> xyz(3 * (size_t)(nv + (order == 2 ? ne : 0)));
  if ((rc = sprg_smb_read(path, tv.data(), te.data(), ev.data(), xyz.data()))) return rc;
  if (order == 2)  // apf::changeShape(m, getLagrange(2), true): edge nodes at the midpoints
    for (int64_t e = 0; e < ne; ++e)
      for (int j = 0; j < 3; ++j) {
        double mid = 0;
        mid += xyz[3 * (size_t)ev[2 * e] + j] * 0.5;
        mid += xyz[3 * (size_t)ev[2 * e + 1] + j] * 0.5;
        xyz[3 * (size_t)(nv + e) + j] = mid;
      }
  return sprg_mesh_create(ctx, nv, nt, ne, tv.data(), te.data(), xyz.data(), order, out);
}
