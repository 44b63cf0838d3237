// mc_cases.h -- the marching-cubes case table, generated (host code).
//
// Cube corners and edges are numbered as polygonize_MarchingCube_Kernel and
// polygoniseCell_MarchingCube number them (MarchingCube_Kernel.cu:45-58,
// 99-122): corner k of node (i, j, k) offset
//   0 (0,0,0)  1 (1,0,0)  2 (1,0,1)  3 (0,0,1)
//   4 (0,1,0)  5 (1,1,0)  6 (1,1,1)  7 (0,1,1)
// edges 0..11 = (0,1) (1,2) (2,3) (3,0) (4,5) (5,6) (6,7) (7,4) (0,4) (1,5)
// (2,6) (3,7); case bit k is set when corner k's value is <= the iso level.
//
// The triangles of each case come from the contour itself rather than a
// transcribed table.  On every face, each maximal run of "below" corners
// (walking the face's corners counter-clockwise seen from outside the cube)
// is cut off by one segment, from the crossed edge where the walk enters the
// run to the crossed edge where it leaves it.  A face with two diagonal
// below corners therefore separates them, and the two cubes sharing a face
// cut it the same way, so the surface has no cracks.  The segments chain
// into closed loops (every crossed edge starts one segment and ends one);
// a loop starts at its smallest edge, loops are taken in the order of those
// edges, and each is fanned from its start.  The winding makes a triangle's
// normal (b - a) x (c - a) point to the below side.
#pragma once

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <vector>

namespace psb {

constexpr int kMcMaxTri = 5;    // the most any case needs (the generator checks)

struct McCases {
  uint8_t ntri[256];
  int8_t tri[256][3 * kMcMaxTri];   // edge indices, -1 padded
};

inline const int (&mc_corner())[8][3] {
  static const int c[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {0, 0, 1},
                              {0, 1, 0}, {1, 1, 0}, {1, 1, 1}, {0, 1, 1}};
  return c;
}
inline const int (&mc_edge())[12][2] {
  static const int e[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                               {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
  return e;
}

inline int mc_edge_of(int a, int b) {
  for (int e = 0; e < 12; ++e)
    if ((mc_edge()[e][0] == a && mc_edge()[e][1] == b) || (mc_edge()[e][0] == b && mc_edge()[e][1] == a))
      return e;
  throw std::logic_error("mc: corners share no edge");
}

// corners of face (axis, side) counter-clockwise seen from outside
inline void mc_face(int axis, int side, int out[4]) {
  const int u = (axis + 1) % 3, w = (axis + 2) % 3;
  // (u, w, axis) is right-handed: counter-clockwise about +axis in the
  // (u, w) plane is (0,0) (1,0) (1,1) (0,1)
  const int quad[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
  for (int q = 0; q < 4; ++q) {
    const int qq = side ? q : 3 - q;   // seen from -axis: reversed
    for (int c = 0; c < 8; ++c) {
      const int* p = mc_corner()[c];
      if (p[axis] == side && p[u] == quad[qq][0] && p[w] == quad[qq][1]) out[q] = c;
    }
  }
}

inline McCases mc_generate() {
  McCases t{};
  // winding: with one below corner (case 1) the triangle must face it;
  // decided on that case and applied to all
  bool flip = false;
  for (int pass = 0; pass < 2; ++pass) {
    for (int cs = 0; cs < 256; ++cs) {
      auto below = [&](int c) { return ((cs >> c) & 1) != 0; };
      int next[12];
      for (int e = 0; e < 12; ++e) next[e] = -1;
      for (int axis = 0; axis < 3; ++axis)
        for (int side = 0; side < 2; ++side) {
          int f[4];
          mc_face(axis, side, f);
          for (int i = 0; i < 4; ++i) {
            const int prev = f[(i + 3) % 4];
            if (!below(f[i]) || below(prev)) continue;   // not the start of a below run
            int j = i;
            while (below(f[(j + 1) % 4])) j = (j + 1) % 4;
            const int ein = mc_edge_of(prev, f[i]);
            const int eout = mc_edge_of(f[j], f[(j + 1) % 4]);
            if (next[ein] != -1) throw std::logic_error("mc: edge starts two segments");
            next[ein] = eout;
          }
        }
      bool used[12] = {false};
      int nt = 0;
      for (int e0 = 0; e0 < 12; ++e0) {
        if (next[e0] == -1 || used[e0]) continue;
        std::vector<int> loop;
        for (int e = e0; !used[e]; e = next[e]) {
          if (next[e] == -1) throw std::logic_error("mc: open contour");
          used[e] = true;
          loop.push_back(e);
        }
        if (loop.size() < 3) throw std::logic_error("mc: degenerate loop");
        for (size_t k = 1; k + 1 < loop.size(); ++k) {
          if (nt >= kMcMaxTri) throw std::logic_error("mc: too many triangles");
          int a = loop[0], b = loop[k], c = loop[k + 1];
          if (flip) std::swap(b, c);
          t.tri[cs][3 * nt] = (int8_t)a;
          t.tri[cs][3 * nt + 1] = (int8_t)b;
          t.tri[cs][3 * nt + 2] = (int8_t)c;
          ++nt;
        }
      }
      for (int k = 3 * nt; k < 3 * kMcMaxTri; ++k) t.tri[cs][k] = -1;
      t.ntri[cs] = (uint8_t)nt;
    }
    if (pass == 0) {
      // case 1: corner 0 below; normal of the triangle at edge midpoints
      auto mid = [&](int e, double m[3]) {
        for (int a = 0; a < 3; ++a)
          m[a] = 0.5 * (mc_corner()[mc_edge()[e][0]][a] + mc_corner()[mc_edge()[e][1]][a]);
      };
      double A[3], B[3], Cc[3];
      mid(t.tri[1][0], A); mid(t.tri[1][1], B); mid(t.tri[1][2], Cc);
      const double u[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
      const double v[3] = {Cc[0] - A[0], Cc[1] - A[1], Cc[2] - A[2]};
      const double n[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2],
                           u[0] * v[1] - u[1] * v[0]};
      // corner 0 lies towards -(1,1,1) from the triangle
      const bool faces_corner = (n[0] + n[1] + n[2]) < 0;
      if (faces_corner) break;
      flip = true;
    }
  }
  return t;
}

}  // namespace psb
