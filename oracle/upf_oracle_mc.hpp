// upf_oracle_mc.hpp -- CPU oracle for marching cubes (TEST INFRASTRUCTURE ONLY).
//
// cage's MarchingCubes (used at sources/meshGeneration.cpp:34-40) is not in the
// reference checkout, so its triangulation, vertex placement and output order are
// unpinned.  This file states the choices libupf implements, independently of the
// library's generated table:
//  * case code bit c <=> f(corner c) < iso, corners in Lorensen-Cline order;
//  * triangulation from face segments: on each face, walked counter-clockwise seen from
//    outside the cube, a segment runs from the edge where the walk enters the inside
//    set to the next edge where it leaves it; segments chain into loops, loops are
//    taken by smallest edge and fanned from the first vertex (in loop order) whose fan
//    diagonals join no two vertices on a common cube face;
//  * vertex on the edge owned by its min node, at the linear zero crossing;
//    vertices numbered by (owner node linear index, axis x<y<z);
//  * triangles in cell linear order, then table order;
//  * normal = normalised lerp of central-difference node gradients.
#pragma once

#include <cstdint>
#include <vector>

#include "upf_oracle.hpp"

namespace upfo {

static const int kCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                  {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
static const int kEdge[12][2] = {{0, 1}, {1, 2}, {3, 2}, {0, 3}, {4, 5}, {5, 6},
                                 {7, 6}, {4, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

struct McTable {
  int ntris[256];
  int tris[256][15];
};

inline int edge_between(int a, int b) {
  for (int e = 0; e < 12; e++)
    if ((kEdge[e][0] == a && kEdge[e][1] == b) || (kEdge[e][0] == b && kEdge[e][1] == a)) return e;
  return -1;
}

// Builds the face cycles from geometry: for each axis/side, order the 4 corners of the
// face by angle around the face centre, counter-clockwise about the outward normal.
inline void face_cycles(int out[6][4]) {
  int f = 0;
  for (int axis = 2; axis >= 0; axis--)
    for (int side = 0; side < 2; side++) {
      int cs[4], n = 0;
      for (int c = 0; c < 8; c++)
        if (kCorner[c][axis] == side) cs[n++] = c;
      // the two in-face axes (u, v) with u x v = +axis
      const int u = (axis + 1) % 3, v = (axis + 2) % 3;
      double ang[4];
      for (int i = 0; i < 4; i++)
        ang[i] = std::atan2(kCorner[cs[i]][v] - 0.5, kCorner[cs[i]][u] - 0.5);
      // sort by angle ascending = CCW about +axis; reverse for the -axis side
      for (int i = 0; i < 4; i++)
        for (int j = i + 1; j < 4; j++)
          if (ang[j] < ang[i]) { std::swap(ang[i], ang[j]); std::swap(cs[i], cs[j]); }
      if (side == 0) std::swap(cs[1], cs[3]);
      for (int i = 0; i < 4; i++) out[f][i] = cs[i];
      f++;
    }
}

// Bit set of the faces (index into face_cycles order) an edge lies on.
inline int edge_face_mask(int e, const int faces[6][4]) {
  int m = 0;
  for (int f = 0; f < 6; f++) {
    bool a = false, b = false;
    for (int i = 0; i < 4; i++) {
      a |= faces[f][i] == kEdge[e][0];
      b |= faces[f][i] == kEdge[e][1];
    }
    if (a && b) m |= 1 << f;
  }
  return m;
}

inline McTable build_table() {
  McTable t;
  int faces[6][4];
  face_cycles(faces);
  for (int code = 0; code < 256; code++) {
    int next[12];
    for (int e = 0; e < 12; e++) next[e] = -1;
    for (int f = 0; f < 6; f++) {
      int ent[4], entpos[4], lev[4], levpos[4], ne = 0, nl = 0;
      for (int i = 0; i < 4; i++) {
        const int a = faces[f][i], b = faces[f][(i + 1) % 4];
        const bool ia = (code >> a) & 1, ib = (code >> b) & 1;
        if (ia == ib) continue;
        if (ib) { ent[ne] = edge_between(a, b); entpos[ne++] = i; }
        else { lev[nl] = edge_between(a, b); levpos[nl++] = i; }
      }
      for (int i = 0; i < ne; i++) {
        int best = -1, bestd = 99;
        for (int j = 0; j < nl; j++) {
          const int d = (levpos[j] - entpos[i] + 4) % 4;
          if (d < bestd) { bestd = d; best = j; }
        }
        next[ent[i]] = lev[best];
      }
    }
    int n = 0;
    bool seen[12] = {};
    for (int s = 0; s < 12; s++) {
      if (next[s] < 0 || seen[s]) continue;
      int loop[12], len = 0;
      for (int e = s; !seen[e]; e = next[e]) { seen[e] = true; loop[len++] = e; }
      int apex = 0;
      for (; apex < len; apex++) {
        bool ok = true;
        for (int j = 0; j < len; j++) {
          if (j == apex || j == (apex + 1) % len || j == (apex + len - 1) % len) continue;
          if (edge_face_mask(loop[apex], faces) & edge_face_mask(loop[j], faces)) ok = false;
        }
        if (ok) break;
      }
      for (int i = 1; i + 1 < len; i++) {
        t.tris[code][3 * n + 0] = loop[apex];
        t.tris[code][3 * n + 1] = loop[(apex + i) % len];
        t.tris[code][3 * n + 2] = loop[(apex + i + 1) % len];
        n++;
      }
    }
    t.ntris[code] = n;
    for (int i = 3 * n; i < 15; i++) t.tris[code][i] = -1;
  }
  return t;
}

inline const McTable& table() {
  static const McTable t = build_table();
  return t;
}

struct McMesh {
  std::vector<float> positions, normals;
  std::vector<uint32_t> indices;
};

struct GridView {
  const float* f;
  int nx, ny, nz;
  float origin[3], spacing[3];
  float at(int i, int j, int k) const { return f[((int64_t)k * ny + j) * nx + i]; }
};

inline int case_code(const GridView& g, int i, int j, int k, float iso) {
  int code = 0;
  for (int c = 0; c < 8; c++)
    if (g.at(i + kCorner[c][0], j + kCorner[c][1], k + kCorner[c][2]) < iso) code |= 1 << c;
  return code;
}

// Central-difference gradient at a node (one-sided at the border).
inline void node_gradient(const GridView& g, int i, int j, int k, float out[3]) {
  const int idx[3] = {i, j, k};
  const int n[3] = {g.nx, g.ny, g.nz};
  for (int a = 0; a < 3; a++) {
    int lo[3] = {i, j, k}, hi[3] = {i, j, k};
    float den;
    if (idx[a] == 0) { hi[a] += 1; den = g.spacing[a]; }
    else if (idx[a] == n[a] - 1) { lo[a] -= 1; den = g.spacing[a]; }
    else { lo[a] -= 1; hi[a] += 1; den = 2.0f * g.spacing[a]; }
    const float d = g.at(hi[0], hi[1], hi[2]) - g.at(lo[0], lo[1], lo[2]);
    out[a] = d / den;
  }
}

inline McMesh mc_extract(const GridView& g, float iso, bool with_normals = true) {
  McMesh m;
  const int64_t nn = (int64_t)g.nx * g.ny * g.nz;
  std::vector<int64_t> vid(3 * nn, -1);
  int64_t nv = 0;
  const int dimv[3] = {g.nx, g.ny, g.nz};
  for (int k = 0; k < g.nz; k++)
    for (int j = 0; j < g.ny; j++)
      for (int i = 0; i < g.nx; i++) {
        const int64_t n0 = ((int64_t)k * g.ny + j) * g.nx + i;
        const float f0 = g.f[n0];
        const int ijk[3] = {i, j, k};
        for (int a = 0; a < 3; a++) {
          if (ijk[a] + 1 >= dimv[a]) continue;
          int q[3] = {i, j, k};
          q[a] += 1;
          const float f1 = g.at(q[0], q[1], q[2]);
          if ((f0 < iso) == (f1 < iso)) continue;
          vid[3 * n0 + a] = nv++;
          const float d = f1 - f0;
          const float t = (iso - f0) / d;
          float p0[3], p1a;
          for (int c = 0; c < 3; c++) p0[c] = node_coord(g.origin[c], g.spacing[c], ijk[c]);
          p1a = node_coord(g.origin[a], g.spacing[a], ijk[a] + 1);
          float pos[3] = {p0[0], p0[1], p0[2]};
          pos[a] = p0[a] + t * (p1a - p0[a]);
          m.positions.insert(m.positions.end(), pos, pos + 3);
          if (with_normals) {
            float g0[3], g1[3], gg[3];
            node_gradient(g, i, j, k, g0);
            node_gradient(g, q[0], q[1], q[2], g1);
            for (int c = 0; c < 3; c++) gg[c] = g0[c] + t * (g1[c] - g0[c]);
            const float len = std::sqrt(gg[0] * gg[0] + gg[1] * gg[1] + gg[2] * gg[2]);
            const float inv = len > 0.0f ? 1.0f / len : 0.0f;
            for (int c = 0; c < 3; c++) m.normals.push_back(gg[c] * inv);
          }
        }
      }
  const McTable& T = table();
  for (int k = 0; k + 1 < g.nz; k++)
    for (int j = 0; j + 1 < g.ny; j++)
      for (int i = 0; i + 1 < g.nx; i++) {
        const int code = case_code(g, i, j, k, iso);
        for (int t = 0; t < T.ntris[code]; t++)
          for (int v = 0; v < 3; v++) {
            const int e = T.tris[code][3 * t + v];
            const int c = kEdge[e][0], c1 = kEdge[e][1];
            int a = 0;
            while (kCorner[c][a] == kCorner[c1][a]) a++;
            const int lo = kCorner[c][a] < kCorner[c1][a] ? c : c1;
            const int64_t owner = ((int64_t)(k + kCorner[lo][2]) * g.ny + (j + kCorner[lo][1])) * g.nx +
                                  (i + kCorner[lo][0]);
            m.indices.push_back((uint32_t)vid[3 * owner + a]);
          }
      }
  return m;
}

}  // namespace upfo
