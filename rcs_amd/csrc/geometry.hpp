// Mesh ingest and the RWG basis table, host side (src/EFIEMatrix.cpp:31-209).
//
// The reference keeps the RWG edges in a std::map keyed by the (min, max) node pair and
// finds basis i with std::next(map.begin(), i) (E:700-701), an O(N) walk per entry.
// Here the table is built once, sorted the same way, and shipped to HBM as
// structure-of-arrays so that every entry kernel reads basis i in O(1).
#pragma once
#include <algorithm>
#include <array>

#include "rcs_common.hpp"

namespace rcs {

struct HostGeometry {
  int kind = KIND_EFIE;
  double freq = 1.0e7;
  // nodes, 1-based: index 0 is the dummy origin of the reference's node list (E:32-33)
  std::vector<double> nx, ny, nz;
  // triangles
  std::vector<int> v1, v2, v3, cond;
  std::vector<double> cx, cy, cz, area, le;
  // RWG bases in lexicographic (n1, n2) order
  std::vector<int> n1, n2, tp, tm, fp, fm;
  std::vector<double> len;

  int n_unknowns() const { return kind == KIND_EFIE ? (int)n1.size() : (int)v1.size(); }

  void build(const double* nodes, int n_nodes, const int* tris, int n_tris, const int* conductor) {
    nx.assign(1, 0.0); ny.assign(1, 0.0); nz.assign(1, 0.0);
    for (int i = 0; i < n_nodes; ++i) {
      nx.push_back(nodes[3 * i]);
      ny.push_back(nodes[3 * i + 1]);
      nz.push_back(nodes[3 * i + 2]);
    }
    v1.resize(n_tris); v2.resize(n_tris); v3.resize(n_tris); cond.resize(n_tris);
    cx.resize(n_tris); cy.resize(n_tris); cz.resize(n_tris); area.resize(n_tris); le.resize(n_tris);
    for (int t = 0; t < n_tris; ++t) {
      int a = tris[3 * t], b = tris[3 * t + 1], c = tris[3 * t + 2];
      if (a < 1 || b < 1 || c < 1 || a > n_nodes || b > n_nodes || c > n_nodes)
        throw std::runtime_error("triangle references a node outside the node list");
      v1[t] = a; v2[t] = b; v3[t] = c;
      cond[t] = conductor ? conductor[t] : 0;  // E:78-79: a patch line without the column reads 0
      double ax = nx[a], ay = ny[a], az = nz[a], bx = nx[b], by = ny[b], bz = nz[b], qx = nx[c], qy = ny[c], qz = nz[c];
      cx[t] = (ax + bx + qx) / 3.0; cy[t] = (ay + by + qy) / 3.0; cz[t] = (az + bz + qz) / 3.0;
      double e1x = bx - ax, e1y = by - ay, e1z = bz - az, e2x = qx - ax, e2y = qy - ay, e2z = qz - az;
      double crx = e1y * e2z - e1z * e2y, cry = e1z * e2x - e1x * e2z, crz = e1x * e2y - e1y * e2x;
      area[t] = 0.5 * std::sqrt(crx * crx + cry * cry + crz * crz);
      double l1 = std::sqrt((ax - bx) * (ax - bx) + (ay - by) * (ay - by) + (az - bz) * (az - bz));
      double l2 = std::sqrt((bx - qx) * (bx - qx) + (by - qy) * (by - qy) + (bz - qz) * (bz - qz));
      le[t] = std::max(l1, l2);  // E:247-254 looks at two of the three edges
    }
    // edge -> patches in file order; keep edges shared by two or more patches
    std::vector<std::array<int, 3>> e;  // (lo, hi, patch)
    e.reserve(3 * (size_t)n_tris);
    for (int t = 0; t < n_tris; ++t) {
      int vv[3] = {v1[t], v2[t], v3[t]};
      for (int k = 0; k < 3; ++k) {
        int a = vv[k], b = vv[(k + 1) % 3];
        e.push_back({std::min(a, b), std::max(a, b), t});
      }
    }
    std::stable_sort(e.begin(), e.end(), [](const std::array<int, 3>& x, const std::array<int, 3>& y) {
      return x[0] != y[0] ? x[0] < y[0] : x[1] < y[1];
    });
    n1.clear(); n2.clear(); tp.clear(); tm.clear(); fp.clear(); fm.clear(); len.clear();
    for (size_t i = 0; i < e.size();) {
      size_t j = i;
      while (j < e.size() && e[j][0] == e[i][0] && e[j][1] == e[i][1]) ++j;
      if (j - i >= 2) {
        int a = e[i][0], b = e[i][1];
        n1.push_back(a); n2.push_back(b);
        tp.push_back(e[i][2]); tm.push_back(e[i + 1][2]);
        fp.push_back(third(e[i][2], a, b));
        fm.push_back(third(e[i + 1][2], a, b));
        len.push_back(std::sqrt((nx[a] - nx[b]) * (nx[a] - nx[b]) + (ny[a] - ny[b]) * (ny[a] - ny[b]) +
                                (nz[a] - nz[b]) * (nz[a] - nz[b])));
      }
      i = j;
    }
  }
  // E:353-373: sorted set difference of the triangle's nodes minus the edge's nodes
  int third(int t, int a, int b) const {
    int s[3] = {v1[t], v2[t], v3[t]};
    std::sort(s, s + 3);
    int lo = std::min(a, b), hi = std::max(a, b);
    for (int k = 0; k < 3; ++k)
      if (s[k] != lo && s[k] != hi) return s[k];
    return s[0];
  }
  // Unknown locations: RWG edge midpoints (S:143-152) or patch centroids (Laplace).
  void locations(std::vector<double>& xyz) const {
    int N = n_unknowns();
    xyz.resize(3 * (size_t)N);
    for (int i = 0; i < N; ++i) {
      if (kind == KIND_EFIE) {
        xyz[3 * i] = (nx[n1[i]] + nx[n2[i]]) / 2.0;
        xyz[3 * i + 1] = (ny[n1[i]] + ny[n2[i]]) / 2.0;
        xyz[3 * i + 2] = (nz[n1[i]] + nz[n2[i]]) / 2.0;
      } else {
        xyz[3 * i] = cx[i]; xyz[3 * i + 1] = cy[i]; xyz[3 * i + 2] = cz[i];
      }
    }
  }
};

// Device copy of the geometry (structure of arrays, read-only).
struct DevGeometry {
  const double *nx, *ny, *nz;
  const int *v1, *v2, *v3;
  const double *cx, *cy, *cz, *area, *le;
  const int *tp, *tm, *fp, *fm;
  const double* len;
  double kappa, omega, eps, mu, C0;
};

}  // namespace rcs
