// Curved hexahedra on the host: GLL coordinates of an element whose edges
// carry Neko's curve records (dofmap_generate_xyz, src/sem/dofmap.f90:
// 797-864, with dofmap_xyzquad :941-1014, gh_face_extend_3d :1023-1127 and
// arc_surface :1197-1285).
//
// Edge numbering of the records is Nek5000's cyclic one: edges 1-4 go round
// the face t = -1 (v1->v2, v2->v4, v4->v3, v3->v1 in Neko's symmetric
// vertex numbering), edges 5-8 the same round t = +1, edges 9-12 are the
// t-edges at the corners v1, v2, v4, v3.
//
// Type 4 (midpoint): the reference builds a 3x3x3 net from the trilinear
// map, puts each such edge's given midpoint into it, re-extends vertices
// and edges into faces and interior by Gordon-Hall blending and
// interpolates the net quadratically to the GLL points.  Gordon-Hall is
// linear in the edge deviations and the linear blends are reproduced
// exactly by the quadratic interpolation, so the result is, in closed form,
//   x(r,s,t) = trilinear(r,s,t)
//            + sum over type-4 edges  (1 - xi^2) h(eta) h(zeta) (M - mid),
// xi the coordinate along the edge, eta / zeta the two others with h the
// linear blend toward the edge's side, M the given point, mid the straight
// edge's midpoint.  That is what is evaluated here (only for lx > 2, as in
// the reference).
//
// Type 3 (circular arc, edges 1-8, in the x-y plane): the arc of radius
// |R| through the edge's end points (centre on the side given by the sign
// of R) replaces the straight edge; its deviation from the chord at the
// GLL points is blended linearly into the element across the edge's face
// and along t.  Other types leave the element unchanged, as in the
// reference.
#include "neko_amd.h"

#include <cmath>
#include <cstdint>
#include <vector>

#include "nm_common.h"

namespace nm {

namespace {

// cyclic edge (0-based) -> (first, last) symmetric vertex in the cyclic
// traversal direction, and the local axis the edge runs along
struct CycEdge {
  int v0, v1, axis;
};
constexpr CycEdge kCyc[12] = {
    {0, 1, 0}, {1, 3, 1}, {3, 2, 0}, {2, 0, 1},  // round t = -1
    {4, 5, 0}, {5, 7, 1}, {7, 6, 0}, {6, 4, 1},  // round t = +1
    {0, 4, 2}, {1, 5, 2}, {3, 7, 2}, {2, 6, 2},  // t-edges
};

inline double lin(int side, double z) { return side ? 0.5 * (1.0 + z) : 0.5 * (1.0 - z); }

// corner (cx, cy, cz) of vertex v
inline int bit(int v, int a) { return (v >> a) & 1; }

}  // namespace

// One element: x, y, z (lx^3 each) from its 8 vertices v (symmetric order)
// and its curve record (cd: 5 values per edge, 12 edges; type: 12).
// Returns NM_ERR_ARG when an arc's radius is too small for its chord.
int curve_element_xyz(int lx, const double *zg, const double *v, const double *cd,
                      const int32_t *type, double *x, double *y, double *z) {
  double *X[3] = {x, y, z};
  // trilinear map
  for (int k = 0; k < lx; k++)
    for (int j = 0; j < lx; j++)
      for (int i = 0; i < lx; i++) {
        const int p = i + lx * (j + lx * k);
        for (int c = 0; c < 3; c++) {
          double acc = 0.0;
          for (int q = 0; q < 8; q++)
            acc += lin(bit(q, 0), zg[i]) * lin(bit(q, 1), zg[j]) * lin(bit(q, 2), zg[k]) *
                   v[q * 3 + c];
          X[c][p] = acc;
        }
      }
  bool quad = false;
  for (int e = 0; e < 12; e++) quad |= type[e] == 4;
  if (quad && lx > 2) {
    for (int e = 0; e < 12; e++) {
      if (type[e] != 4) continue;
      const CycEdge &E = kCyc[e];
      // the two fixed axes and the sides of the edge on them
      const int a1 = (E.axis + 1) % 3, a2 = (E.axis + 2) % 3;
      const int s1 = bit(E.v0, a1), s2 = bit(E.v0, a2);
      double dev[3];
      for (int c = 0; c < 3; c++)
        dev[c] = cd[e * 5 + c] - 0.5 * (v[E.v0 * 3 + c] + v[E.v1 * 3 + c]);
      for (int k = 0; k < lx; k++)
        for (int j = 0; j < lx; j++)
          for (int i = 0; i < lx; i++) {
            const double zc[3] = {zg[i], zg[j], zg[k]};
            const double w = (1.0 - zc[E.axis] * zc[E.axis]) * lin(s1, zc[a1]) * lin(s2, zc[a2]);
            const int p = i + lx * (j + lx * k);
            for (int c = 0; c < 3; c++) X[c][p] += w * dev[c];
          }
    }
  }
  for (int e = 0; e < 8; e++) {
    if (type[e] != 3) continue;
    const CycEdge &E = kCyc[e];
    const double R = cd[e * 5];
    const double p1x = v[E.v0 * 3], p1y = v[E.v0 * 3 + 1];
    const double p2x = v[E.v1 * 3], p2y = v[E.v1 * 3 + 1];
    // chord normal (p2 - p1 turned by -90 degrees) and half angle
    const double nx = p2y - p1y, ny = p1x - p2x;
    const double len = std::sqrt(nx * nx + ny * ny);
    if (std::fabs(2.0 * R) <= len * 1.00001) {
      set_error("curved element: arc radius %g too small for an edge of length %g", R, len);
      return NM_ERR_ARG;
    }
    double half = std::fabs(std::asin(0.5 * len / R));
    const double mx = 0.5 * (p1x + p2x), my = 0.5 * (p1y + p2y);
    const double cx = mx - nx / len * R * std::cos(half);
    const double cy = my - ny / len * R * std::cos(half);
    const double th0 = std::atan2(my - cy, mx - cx);
    if (R < 0.0) half = -half;
    // deviation of the arc from the chord at the GLL points, in the
    // cyclic direction (from p1 to p2), indexed along increasing local
    // coordinate of the edge's axis
    const bool reversed = bit(E.v0, E.axis) == 1;
    std::vector<double> dx(lx), dy(lx);
    for (int q = 0; q < lx; q++) {
      const double r = zg[q], th = th0 + r * half;
      const int at = reversed ? lx - 1 - q : q;
      dx[at] = cx + std::fabs(R) * std::cos(th) - (lin(0, r) * p1x + lin(1, r) * p2x);
      dy[at] = cy + std::fabs(R) * std::sin(th) - (lin(0, r) * p1y + lin(1, r) * p2y);
    }
    const int other = E.axis == 0 ? 1 : 0;  // the in-plane axis across the edge
    const int so = bit(E.v0, other), st = bit(E.v0, 2);
    for (int k = 0; k < lx; k++)
      for (int j = 0; j < lx; j++)
        for (int i = 0; i < lx; i++) {
          const int idx[3] = {i, j, k};
          const double w = lin(so, zg[idx[other]]) * lin(st, zg[k]);
          const int p = i + lx * (j + lx * k);
          x[p] += w * dx[idx[E.axis]];
          y[p] += w * dy[idx[E.axis]];
        }
  }
  return NM_OK;
}

}  // namespace nm

extern "C" int nm_curve_xyz(int lx, const double *vxyz, const double *data,
                            const int32_t *type, double *x, double *y, double *z) {
  if (lx < 2 || lx > 16) {
    nm::set_error("nm_curve_xyz: lx=%d outside 2..16", lx);
    return nm::NM_ERR_ARG;
  }
  std::vector<double> zg(lx), w(lx), d(lx * lx), dt(lx * lx);
  int st = nm_gll(lx, zg.data(), w.data(), d.data(), dt.data());
  if (st) return st;
  return nm::curve_element_xyz(lx, zg.data(), vxyz, data, type, x, y, z);
}
