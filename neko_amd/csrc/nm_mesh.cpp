// General conforming hexahedral meshes on the host: Neko's binary .nmsh
// reader, facet connectivity, the topological identification of shared GLL
// points, Neko's gs maps (same registration order and sort as the box
// builder, so the same bit-exact summation order) and the Dirichlet mask of
// every facet without a neighbour.
//
// Vertex order everywhere is Neko's symmetric (lexicographic) hex order:
// vertex v = cx + 2 cy + 4 cz for the corner (cx, cy, cz) in {0,1}^3.
// Facets 0..5 are Neko's facets 1..6: x = -1, x = +1, y = -1, y = +1,
// z = -1, z = +1 (hex.f90 face numbering).
#include "neko_amd.h"

#include <algorithm>
#include <array>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "nm_common.h"

namespace nm {

namespace {

// corners of each facet as (a, b) in-plane corner -> vertex, a fastest
inline int facet_vertex(int f, int ca, int cb) {
  const int side = f & 1;
  switch (f >> 1) {
    case 0: return side + 2 * ca + 4 * cb;    // x-facet: (y, z)
    case 1: return 2 * side + ca + 4 * cb;    // y-facet: (x, z)
    default: return 4 * side + ca + 2 * cb;   // z-facet: (x, y)
  }
}

// the 12 edges as (first vertex, last vertex), direction of increasing
// local coordinate; Neko's edge order (r-edges, s-edges, t-edges)
constexpr int kEdge[12][2] = {{0, 1}, {4, 5}, {2, 3}, {6, 7}, {0, 2}, {4, 6},
                              {1, 3}, {5, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

using Key4 = std::array<int64_t, 4>;

Key4 facet_key(const int64_t *p, int f) {
  Key4 k{p[facet_vertex(f, 0, 0)], p[facet_vertex(f, 1, 0)], p[facet_vertex(f, 0, 1)],
         p[facet_vertex(f, 1, 1)]};
  std::sort(k.begin(), k.end());
  return k;
}

// unique ids of the 6 facets of every element (slot e*6+f), plus the facet
// neighbour (1-based element, 0 on the boundary).  Fails on a facet held by
// more than two elements.
int number_facets(int64_t nel, const int64_t *pid, std::vector<int64_t> &fid,
                  int32_t *fneigh) {
  std::vector<std::pair<Key4, int64_t>> all((size_t)nel * 6);
  for (int64_t e = 0; e < nel; e++)
    for (int f = 0; f < 6; f++) all[e * 6 + f] = {facet_key(pid + e * 8, f), e * 6 + f};
  std::sort(all.begin(), all.end());
  fid.assign((size_t)nel * 6, -1);
  int64_t uid = -1;
  for (size_t a = 0; a < all.size();) {
    size_t b = a + 1;
    while (b < all.size() && all[b].first == all[a].first) b++;
    if (b - a > 2) {
      set_error("mesh: facet (%lld %lld %lld %lld) is held by %zu elements",
                (long long)all[a].first[0], (long long)all[a].first[1],
                (long long)all[a].first[2], (long long)all[a].first[3], b - a);
      return NM_ERR_ARG;
    }
    uid++;
    for (size_t c = a; c < b; c++) fid[all[c].second] = uid;
    if (fneigh != nullptr) {
      if (b - a == 2) {
        fneigh[all[a].second] = (int32_t)(all[a + 1].second / 6 + 1);
        fneigh[all[a + 1].second] = (int32_t)(all[a].second / 6 + 1);
      } else {
        fneigh[all[a].second] = 0;
      }
    }
    a = b;
  }
  return NM_OK;
}

// unique ids of the 12 edges of every element (slot e*12+q)
void number_edges(int64_t nel, const int64_t *pid, std::vector<int64_t> &eid) {
  std::vector<std::pair<std::pair<int64_t, int64_t>, int64_t>> all((size_t)nel * 12);
  for (int64_t e = 0; e < nel; e++)
    for (int q = 0; q < 12; q++) {
      const int64_t a = pid[e * 8 + kEdge[q][0]], b = pid[e * 8 + kEdge[q][1]];
      all[e * 12 + q] = {{std::min(a, b), std::max(a, b)}, e * 12 + q};
    }
  std::sort(all.begin(), all.end());
  eid.assign((size_t)nel * 12, -1);
  int64_t uid = -1;
  for (size_t a = 0; a < all.size(); a++) {
    if (a == 0 || all[a].first != all[a - 1].first) uid++;
    eid[all[a].second] = uid;
  }
}

// Topological key of local point (i,j,k) of element e: equal keys <=> the
// same global GLL point of a conforming mesh.  Vertices: the point id.
// Edge interiors: the edge and the distance from its lower-id end.  Facet
// interiors: the facet and the position in the frame whose origin is the
// facet corner with the lowest id and whose first axis points to the lower
// -id neighbour of that corner (the convention of Neko's dofmap numbering,
// dofmap.f90:316-330,742-797; any frame both elements derive from the ids
// alone would do).  Interior points never share and are not asked for.
struct KeyMaker {
  int lx, L;
  const int64_t *pid;
  const std::vector<int64_t> &eid, &fid;
  int64_t ebase, fbase;

  int64_t operator()(int64_t e, int i, int j, int k) const {
    const int l1 = lx - 1;
    const int bi = (i == 0 || i == l1), bj = (j == 0 || j == l1), bk = (k == 0 || k == l1);
    const int64_t *p = pid + e * 8;
    const int nb = bi + bj + bk;
    if (nb == 3) return p[(i == l1) + 2 * (j == l1) + 4 * (k == l1)];
    if (nb == 2) {
      int q, t;
      if (!bi) { q = (j == l1) * 2 + (k == l1); t = i; }
      else if (!bj) { q = 4 + (i == l1) * 2 + (k == l1); t = j; }
      else { q = 8 + (i == l1) + 2 * (j == l1); t = k; }
      const int pos = p[kEdge[q][0]] < p[kEdge[q][1]] ? t - 1 : L - t;
      return ebase + eid[e * 12 + q] * L + pos;
    }
    int f, a, b;
    if (bi) { f = (i == l1); a = j; b = k; }
    else if (bj) { f = 2 + (j == l1); a = i; b = k; }
    else { f = 4 + (k == l1); a = i; b = j; }
    // corner ids c[ca][cb]
    int64_t c[2][2];
    for (int ca = 0; ca < 2; ca++)
      for (int cb = 0; cb < 2; cb++) c[ca][cb] = p[facet_vertex(f, ca, cb)];
    int oa = 0, ob = 0;
    for (int ca = 0; ca < 2; ca++)
      for (int cb = 0; cb < 2; cb++)
        if (c[ca][cb] < c[oa][ob]) { oa = ca; ob = cb; }
    const int da = oa ? L - a : a - 1;  // 0-based distance from the origin
    const int db = ob ? L - b : b - 1;
    const bool a_first = c[1 - oa][ob] < c[oa][1 - ob];
    const int u = a_first ? da : db, v = a_first ? db : da;
    return fbase + fid[e * 6 + f] * (int64_t)L * L + u + (int64_t)L * v;
  }
};

}  // namespace

// Hoare sort of (dg, gd) by dg with Neko's partition scheme (nm_setup.cpp)
void hoare_sort_dofs(int32_t *dg, int32_t *gd, int64_t lo, int64_t hi);

}  // namespace nm

using namespace nm;

extern "C" {

// ---------------------------------------------------------------------------
// .nmsh (nmsh_file.f90:70-207, format nmsh.f90): int32 nelv, int32 gdim;
// nelv hex records of 228 bytes (int32 el_idx, 8 x {int32 v_idx, 3 x f64});
// int32 nzones, nzones records of 9 int32 (e, f, p_e, p_f, glb_pt_ids(4),
// type); int32 ncurves, curve records.  Little-endian, packed.
// ---------------------------------------------------------------------------
namespace {
struct NmshHead {
  int32_t nel = 0, gdim = 0, nzones = 0, ncurves = 0;
  long zone_off = 0;
};

int nmsh_head(FILE *fp, const char *path, NmshHead &h) {
  if (std::fread(&h.nel, 4, 1, fp) != 1 || std::fread(&h.gdim, 4, 1, fp) != 1) {
    set_error("nmsh %s: truncated header", path);
    return NM_ERR_ARG;
  }
  if (h.gdim != 3) {
    set_error("nmsh %s: gdim %d (only 3-D hex meshes are supported)", path, h.gdim);
    return NM_ERR_UNSUPPORTED;
  }
  if (h.nel <= 0) {
    set_error("nmsh %s: nelv %d", path, h.nel);
    return NM_ERR_ARG;
  }
  h.zone_off = 8L + 228L * h.nel;
  if (std::fseek(fp, h.zone_off, SEEK_SET) != 0 || std::fread(&h.nzones, 4, 1, fp) != 1 ||
      h.nzones < 0) {
    set_error("nmsh %s: truncated (no zone count after %d elements)", path, h.nel);
    return NM_ERR_ARG;
  }
  if (std::fseek(fp, h.zone_off + 4 + 36L * h.nzones, SEEK_SET) != 0 ||
      std::fread(&h.ncurves, 4, 1, fp) != 1) {
    set_error("nmsh %s: truncated (no curve count after %d zones)", path, h.nzones);
    return NM_ERR_ARG;
  }
  return NM_OK;
}
}  // namespace

int nm_nmsh_info(const char *path, int64_t *nel, int64_t *nzones, int64_t *nperiodic,
                 int64_t *ncurves) {
  FILE *fp = std::fopen(path, "rb");
  if (!fp) {
    set_error("nmsh %s: cannot open", path);
    return NM_ERR_ARG;
  }
  NmshHead h;
  int st = nmsh_head(fp, path, h);
  int64_t nper = 0;
  if (st == NM_OK) {
    std::vector<int32_t> z((size_t)h.nzones * 9);
    if (std::fseek(fp, h.zone_off + 4, SEEK_SET) != 0 ||
        std::fread(z.data(), 36, (size_t)h.nzones, fp) != (size_t)h.nzones) {
      set_error("nmsh %s: truncated zones", path);
      st = NM_ERR_ARG;
    } else {
      for (int32_t q = 0; q < h.nzones; q++) nper += z[q * 9 + 8] == 5;
    }
  }
  std::fclose(fp);
  if (st) return st;
  if (nel) *nel = h.nel;
  if (nzones) *nzones = h.nzones;
  if (nperiodic) *nperiodic = nper;
  if (ncurves) *ncurves = h.ncurves;
  return NM_OK;
}

// Elements in file order.  vxyz: 24 doubles per element, Neko's symmetric
// vertex order (the file's p1,p2,p3,p4,p5,p6,p7,p8 become p1,p2,p4,p3,p5,
// p6,p8,p7, nmsh_file.f90:159); pid: the 8 point ids after the periodic
// zones (type 5) have replaced the ids of their facet's points by the
// zone's glb_pt_ids (mesh_apply_periodic_facet, mesh.f90:1837-1866: a point
// is one object, so the new id holds in every element that shares it; a
// later zone overrides an earlier one).  el_idx (may be NULL): the file's
// global element numbers.  The vertices only: a file with curve records
// (ncurves > 0) needs them too (nm_nmsh_curves), else its elements are
// taken straight.
// The zone records of a .nmsh file (nmsh_file.f90:326-380): per zone
// {element (0-based, in file order), facet 1..6, type (5 periodic, 7
// labelled), label (the zone's p_f: the boundary label of a type-7 zone)},
// 4 int32 each; nzones from nm_nmsh_info.
int nm_nmsh_zones(const char *path, int64_t nzones, int32_t *out) {
  int64_t nel = 0, nz = 0, np = 0, nc = 0;
  int st = nm_nmsh_info(path, &nel, &nz, &np, &nc);
  if (st) return st;
  if (nz != nzones) {
    set_error("nmsh %s: holds %lld zones, caller expects %lld", path, (long long)nz,
              (long long)nzones);
    return NM_ERR_ARG;
  }
  std::vector<double> v((size_t)std::max<int64_t>(nel, 1) * 24);
  std::vector<int64_t> pid((size_t)std::max<int64_t>(nel, 1) * 8);
  std::vector<int32_t> el((size_t)std::max<int64_t>(nel, 1));
  if ((st = nm_nmsh_read(path, nel, v.data(), pid.data(), el.data()))) return st;
  std::unordered_map<int32_t, int64_t> glb2loc;
  for (int64_t e = 0; e < nel; e++) glb2loc[el[e]] = e;
  FILE *fp = std::fopen(path, "rb");
  if (!fp) {
    set_error("nmsh %s: cannot open", path);
    return NM_ERR_ARG;
  }
  NmshHead h;
  if ((st = nmsh_head(fp, path, h))) {
    std::fclose(fp);
    return st;
  }
  std::vector<int32_t> z((size_t)std::max<int64_t>(nz, 1) * 9);
  const bool ok = nz == 0 || (std::fseek(fp, h.zone_off + 4, SEEK_SET) == 0 &&
                              std::fread(z.data(), 36, (size_t)nz, fp) == (size_t)nz);
  std::fclose(fp);
  if (!ok) {
    set_error("nmsh %s: truncated zones", path);
    return NM_ERR_ARG;
  }
  for (int64_t q = 0; q < nz; q++) {
    const int32_t *zq = z.data() + q * 9;
    auto it = glb2loc.find(zq[0]);
    if (it == glb2loc.end()) {
      set_error("nmsh %s: zone %lld names element %d", path, (long long)q, zq[0]);
      return NM_ERR_ARG;
    }
    out[4 * q] = (int32_t)it->second;
    out[4 * q + 1] = zq[1];
    out[4 * q + 2] = zq[8];
    out[4 * q + 3] = zq[3];
  }
  return NM_OK;
}

// The curve records (nmsh_file_read_curves, nmsh_file.f90:575-639; the
// packed layout of nmsh_curve_el_t, format/nmsh.f90:39-43, as the MPI
// struct type of mpi_types.f90:225-249 writes it: int32 e, 60 f64
// curve_data(5,12), 12 int32 types = 532 bytes): per record the element
// (0-based, file order), its 60 curve values and 12 edge types.
int nm_nmsh_curves(const char *path, int64_t ncurves, int32_t *el, double *data,
                   int32_t *type) {
  int64_t nel = 0, nz = 0, np = 0, nc = 0;
  int st = nm_nmsh_info(path, &nel, &nz, &np, &nc);
  if (st) return st;
  if (nc != ncurves) {
    set_error("nmsh %s: holds %lld curve records, caller expects %lld", path, (long long)nc,
              (long long)ncurves);
    return NM_ERR_ARG;
  }
  if (nc == 0) return NM_OK;
  FILE *fp = std::fopen(path, "rb");
  if (!fp) {
    set_error("nmsh %s: cannot open", path);
    return NM_ERR_ARG;
  }
  NmshHead h;
  if ((st = nmsh_head(fp, path, h))) {
    std::fclose(fp);
    return st;
  }
  std::vector<unsigned char> hex((size_t)nel * 228), rec((size_t)nc * 532);
  const bool ok = std::fseek(fp, 8, SEEK_SET) == 0 &&
                  std::fread(hex.data(), 228, (size_t)nel, fp) == (size_t)nel &&
                  std::fseek(fp, h.zone_off + 4 + 36L * h.nzones + 4, SEEK_SET) == 0 &&
                  std::fread(rec.data(), 532, (size_t)nc, fp) == (size_t)nc;
  std::fclose(fp);
  if (!ok) {
    set_error("nmsh %s: truncated curve records", path);
    return NM_ERR_ARG;
  }
  std::unordered_map<int32_t, int64_t> glb2loc;
  glb2loc.reserve((size_t)nel * 2);
  for (int64_t e = 0; e < nel; e++) {
    int32_t idx;
    std::memcpy(&idx, hex.data() + e * 228, 4);
    glb2loc[idx] = e;
  }
  for (int64_t c = 0; c < nc; c++) {
    const unsigned char *r = rec.data() + c * 532;
    int32_t e;
    std::memcpy(&e, r, 4);
    auto it = glb2loc.find(e);
    if (it == glb2loc.end()) {
      set_error("nmsh %s: curve record %lld names element %d", path, (long long)c, e);
      return NM_ERR_ARG;
    }
    el[c] = (int32_t)it->second;
    std::memcpy(data + c * 60, r + 4, 480);
    std::memcpy(type + c * 12, r + 484, 48);
  }
  return NM_OK;
}

int nm_nmsh_read(const char *path, int64_t nel, double *vxyz, int64_t *pid,
                 int32_t *el_idx) {
  FILE *fp = std::fopen(path, "rb");
  if (!fp) {
    set_error("nmsh %s: cannot open", path);
    return NM_ERR_ARG;
  }
  NmshHead h;
  int st = nmsh_head(fp, path, h);
  if (st) {
    std::fclose(fp);
    return st;
  }
  if (h.nel != nel) {
    std::fclose(fp);
    set_error("nmsh %s: holds %d elements, caller expects %lld", path, h.nel, (long long)nel);
    return NM_ERR_ARG;
  }
  std::vector<unsigned char> rec((size_t)nel * 228);
  std::vector<int32_t> z((size_t)h.nzones * 9);
  bool ok = std::fseek(fp, 8, SEEK_SET) == 0 &&
            std::fread(rec.data(), 228, (size_t)nel, fp) == (size_t)nel &&
            std::fseek(fp, h.zone_off + 4, SEEK_SET) == 0 &&
            std::fread(z.data(), 36, (size_t)h.nzones, fp) == (size_t)h.nzones;
  std::fclose(fp);
  if (!ok) {
    set_error("nmsh %s: truncated element or zone records", path);
    return NM_ERR_ARG;
  }
  static const int kFileToNeko[8] = {0, 1, 3, 2, 4, 5, 7, 6};
  std::unordered_map<int32_t, int64_t> glb2loc;
  glb2loc.reserve((size_t)nel * 2);
  for (int64_t e = 0; e < nel; e++) {
    const unsigned char *r = rec.data() + e * 228;
    int32_t idx;
    std::memcpy(&idx, r, 4);
    if (el_idx) el_idx[e] = idx;
    glb2loc[idx] = e;
    for (int v = 0; v < 8; v++) {
      const unsigned char *pv = r + 4 + 28 * kFileToNeko[v];
      int32_t id;
      std::memcpy(&id, pv, 4);
      if (id <= 0) {
        set_error("nmsh %s: element %lld has point id %d", path, (long long)e, id);
        return NM_ERR_ARG;
      }
      pid[e * 8 + v] = id;
      std::memcpy(vxyz + e * 24 + v * 3, pv + 4, 24);
    }
  }
  // periodic zones: point object (original id) -> replacement id
  std::unordered_map<int64_t, int64_t> remap;
  for (int32_t q = 0; q < h.nzones; q++) {
    const int32_t *zq = z.data() + (size_t)q * 9;
    if (zq[8] != 5) continue;
    auto it = glb2loc.find(zq[0]);
    if (it == glb2loc.end() || zq[1] < 1 || zq[1] > 6) {
      set_error("nmsh %s: periodic zone %d names element %d facet %d", path, q, zq[0], zq[1]);
      return NM_ERR_ARG;
    }
    const int64_t e = it->second;
    const int f = zq[1] - 1;
    // facet points in Neko's face_nodes order (mesh.f90:1848-1855)
    static const int kFaceNodes[6][4] = {{0, 4, 6, 2}, {1, 5, 7, 3}, {0, 1, 5, 4},
                                         {2, 3, 7, 6}, {0, 1, 3, 2}, {4, 5, 7, 6}};
    for (int c = 0; c < 4; c++) remap[pid[e * 8 + kFaceNodes[f][c]]] = zq[4 + c];
  }
  if (!remap.empty())
    for (int64_t s = 0; s < nel * 8; s++) {
      auto it = remap.find(pid[s]);
      if (it != remap.end()) pid[s] = it->second;
    }
  return NM_OK;
}

// Facet neighbours (1-based element, 0 = boundary), slot e*6 + f.
int nm_mesh_facet_neigh(int64_t nel, const int64_t *pid, int32_t *fneigh) {
  std::vector<int64_t> fid;
  return number_facets(nel, pid, fid, fneigh);
}

int nm_mesh_gs_sizes(int64_t nel, int lx, const int32_t *fneigh, int64_t *nlocal,
                     int64_t *facet_offset) {
  if (nel <= 0 || lx < 2) {
    set_error("nm_mesh_gs_sizes: nel %lld lx %d", (long long)nel, lx);
    return NM_ERR_ARG;
  }
  const int64_t L = lx - 2;
  int64_t nf = 0;
  for (int64_t s = 0; s < nel * 6; s++) nf += fneigh[s] != 0;
  const int64_t base = nel * (8 + 12 * L);
  *nlocal = base + nf * L * L;
  *facet_offset = base + 1;
  if (*nlocal > INT32_MAX) {
    set_error("nm_mesh_gs_sizes: %lld gs entries exceed 32-bit indices", (long long)*nlocal);
    return NM_ERR_UNSUPPORTED;
  }
  return NM_OK;
}

// Neko's local gs maps of a general conforming mesh (gather_scatter.f90:
// 513-1375, single rank): the registration order of the box builder (all
// vertices element by element, then all edge interiors, then the interior
// points of every facet with a neighbour), gs ids in order of first
// registration, Neko's sort, blocks below the facet offset.
int nm_mesh_gs_map(int64_t nel, int lx, const int64_t *pid, const int32_t *fneigh,
                   int32_t *dg, int32_t *gd, int32_t *blk_len, int32_t *blk_off,
                   int64_t *nlocal_out, int64_t *facet_offset_out, int64_t *nblks_out) {
  int64_t nlocal, foff;
  int st = nm_mesh_gs_sizes(nel, lx, fneigh, &nlocal, &foff);
  if (st) return st;
  const int L = lx - 2, l1 = lx - 1;
  std::vector<int64_t> eid, fid;
  if ((st = number_facets(nel, pid, fid, nullptr))) return st;
  number_edges(nel, pid, eid);
  int64_t maxp = 0;
  for (int64_t s = 0; s < nel * 8; s++) maxp = std::max(maxp, pid[s]);
  const int64_t ne = eid.empty() ? 0 : *std::max_element(eid.begin(), eid.end()) + 1;
  const int64_t nfu = fid.empty() ? 0 : *std::max_element(fid.begin(), fid.end()) + 1;
  KeyMaker key{lx, L, pid, eid, fid, maxp + 1, maxp + 1 + ne * L};
  std::vector<int32_t> gsid((size_t)(maxp + 1 + ne * L + nfu * (int64_t)L * L), 0);
  int32_t max_id = 0;
  int64_t n = 0;
  auto push = [&](int64_t e, int i, int j, int k) {
    int32_t &s = gsid[(size_t)key(e, i, j, k)];
    if (s == 0) s = ++max_id;
    dg[n] = s;
    gd[n] = (int32_t)(i + lx * (j + (int64_t)lx * (k + (int64_t)lx * e)) + 1);
    n++;
  };
  for (int64_t e = 0; e < nel; e++) {
    push(e, 0, 0, 0); push(e, l1, 0, 0); push(e, 0, l1, 0); push(e, l1, l1, 0);
    push(e, 0, 0, l1); push(e, l1, 0, l1); push(e, 0, l1, l1); push(e, l1, l1, l1);
  }
  for (int64_t e = 0; e < nel; e++) {
    for (int a = 1; a < l1; a++) push(e, a, 0, 0);
    for (int a = 1; a < l1; a++) push(e, a, 0, l1);
    for (int a = 1; a < l1; a++) push(e, a, l1, 0);
    for (int a = 1; a < l1; a++) push(e, a, l1, l1);
    for (int a = 1; a < l1; a++) push(e, 0, a, 0);
    for (int a = 1; a < l1; a++) push(e, 0, a, l1);
    for (int a = 1; a < l1; a++) push(e, l1, a, 0);
    for (int a = 1; a < l1; a++) push(e, l1, a, l1);
    for (int a = 1; a < l1; a++) push(e, 0, 0, a);
    for (int a = 1; a < l1; a++) push(e, l1, 0, a);
    for (int a = 1; a < l1; a++) push(e, 0, l1, a);
    for (int a = 1; a < l1; a++) push(e, l1, l1, a);
  }
  for (int64_t e = 0; e < nel; e++) {
    const int32_t *nb = fneigh + e * 6;
    if (nb[0]) for (int b = 1; b < l1; b++) for (int a = 1; a < l1; a++) push(e, 0, a, b);
    if (nb[1]) for (int b = 1; b < l1; b++) for (int a = 1; a < l1; a++) push(e, l1, a, b);
    if (nb[2]) for (int b = 1; b < l1; b++) for (int a = 1; a < l1; a++) push(e, a, 0, b);
    if (nb[3]) for (int b = 1; b < l1; b++) for (int a = 1; a < l1; a++) push(e, a, l1, b);
    if (nb[4]) for (int b = 1; b < l1; b++) for (int a = 1; a < l1; a++) push(e, a, b, 0);
    if (nb[5]) for (int b = 1; b < l1; b++) for (int a = 1; a < l1; a++) push(e, a, b, l1);
  }
  if (n != nlocal) {
    set_error("nm_mesh_gs_map: internal count %lld != %lld", (long long)n, (long long)nlocal);
    return NM_ERR_STATE;
  }
  gsid.clear();
  gsid.shrink_to_fit();
  if (nlocal > 0) hoare_sort_dofs(dg, gd, 1, nlocal);
  int64_t nb = 0, i = 1;
  while (i < foff) {
    const int32_t id = dg[i - 1];
    int64_t j = i, cnt = 1;
    while (j + 1 <= nlocal && dg[j] == id) {
      j++;
      cnt++;
    }
    blk_len[nb++] = (int32_t)cnt;
    i = j + 1;
  }
  if (nb > 0) {
    blk_off[0] = 0;
    for (int64_t q = 1; q < nb; q++) blk_off[q] = blk_off[q - 1] + blk_len[q - 1];
  }
  *nlocal_out = nlocal;
  *facet_offset_out = foff;
  *nblks_out = nb;
  return NM_OK;
}

// Dirichlet mask of every facet without a neighbour (the Poisson example's
// set_bc: mark_facet where facet_neigh == 0), 1-based local indices, all
// lx*lx points of each marked facet in facet order.
int64_t nm_mesh_mask_size(int64_t nel, int lx, const int32_t *fneigh) {
  int64_t nf = 0;
  for (int64_t s = 0; s < nel * 6; s++) nf += fneigh[s] == 0;
  return nf * lx * lx;
}

int nm_mesh_mask(int64_t nel, int lx, const int32_t *fneigh, int32_t *msk, int64_t *nmsk) {
  const int l1 = lx - 1;
  int64_t m = 0;
  for (int64_t e = 0; e < nel; e++)
    for (int f = 0; f < 6; f++) {
      if (fneigh[e * 6 + f] != 0) continue;
      for (int b = 0; b < lx; b++)
        for (int a = 0; a < lx; a++) {
          int i, j, k;
          if (f < 2) { i = (f == 1) ? l1 : 0; j = a; k = b; }
          else if (f < 4) { j = (f == 3) ? l1 : 0; i = a; k = b; }
          else { k = (f == 5) ? l1 : 0; i = a; j = b; }
          msk[m++] = (int32_t)(i + lx * (j + (int64_t)lx * (k + (int64_t)lx * e)) + 1);
        }
    }
  *nmsk = m;
  return NM_OK;
}

}  // extern "C"
