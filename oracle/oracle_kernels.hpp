// ORACLE — test infrastructure only (see oracle_core.hpp).
//
// Mesh ingest, RWG numbering and MoM entry generation restated from
// src/EFIEMatrix.cpp, src/ElectrostaticsMatrix.cpp, src/includes.cpp and src/kernel.cpp.
#pragma once
#include "oracle_core.hpp"

namespace oracle {

// include/Patch.hpp:6-13
struct Patch {
  int v1, v2, v3, conductor;
  V3 centroid;
  double area, largestEdge;
};

// include/EFIEMatrix.hpp:22-25 (edgeProperties)
struct EdgeProps {
  double length;
  std::vector<int> patches;
};

// One RWG basis: the i-th key of edge2PatchMap in lexicographic order (E:129-209).
struct RWG {
  int n1, n2;       // edge nodes, n1 < n2
  double length;
  int tp, tm;       // T+ = first associated patch in file order, T- = second
  int fp, fm;       // free (non-RWG) vertex of T+ and T- (getIndexOfThirdNode, E:353-373)
};

struct Mesh {
  std::vector<V3> nodes;  // 1-based; nodes[0] is the dummy origin (E:32-33)
  std::vector<Patch> patches;
  std::vector<RWG> rwg;
  std::set<int> conductorIndices;

  // E:247-254: only the first two edge lengths enter std::max_element(l, l+2).
  static double largestEdge(const V3& a, const V3& b, const V3& c) {
    double l1 = (a - b).norm(), l2 = (b - c).norm();
    return std::max(l1, l2);
  }
  static double area(const V3& a, const V3& b, const V3& c) {  // E:262-268
    return 0.5 * (b - a).cross(c - a).norm();
  }
  // E:353-373: the node of the triangle not on the edge (sorted set difference).
  int thirdNode(int tri, int e1, int e2) const {
    int t[3] = {patches[tri].v1, patches[tri].v2, patches[tri].v3};
    int e[2] = {e1, e2};
    std::sort(t, t + 3);
    std::sort(e, e + 2);
    std::vector<int> diff(3);
    auto end = std::set_difference(t, t + 3, e, e + 2, diff.begin());
    if (end == diff.begin()) return t[0];
    return diff[0];
  }

  // nodes0: nNodes x 3 (the file lines, without the dummy); tris: nTri x 3 (1-based).
  void build(const double* nodes0, int nNodes, const int* tris, int nTri, const int* conductor) {
    nodes.clear();
    nodes.push_back(V3(0, 0, 0));
    for (int i = 0; i < nNodes; ++i) nodes.push_back(V3(nodes0[3 * i], nodes0[3 * i + 1], nodes0[3 * i + 2]));
    patches.clear();
    for (int t = 0; t < nTri; ++t) {
      Patch p;
      p.v1 = tris[3 * t];
      p.v2 = tris[3 * t + 1];
      p.v3 = tris[3 * t + 2];
      p.conductor = conductor ? conductor[t] : 0;  // E:78-79: a patch line without the column reads 0
      conductorIndices.insert(p.conductor);
      const V3 &a = nodes[p.v1], &b = nodes[p.v2], &c = nodes[p.v3];
      p.area = area(a, b, c);
      p.centroid = (a + b + c) / 3.0;
      p.largestEdge = largestEdge(a, b, c);
      patches.push_back(p);
    }
    // E:129-197 then deleteNonRWGEdges E:199-209.
    std::map<std::pair<int, int>, EdgeProps> e2p;
    for (int t = 0; t < nTri; ++t) {
      const Patch& p = patches[t];
      int idx[3][2] = {{p.v1, p.v2}, {p.v2, p.v3}, {p.v3, p.v1}};
      for (int e = 0; e < 3; ++e) {
        int a = idx[e][0], b = idx[e][1];
        std::pair<int, int> key(std::min(a, b), std::max(a, b));
        auto it = e2p.find(key);
        if (it == e2p.end()) {
          EdgeProps ep;
          ep.length = (nodes[a] - nodes[b]).norm();
          ep.patches.push_back(t);
          e2p[key] = ep;
        } else {
          it->second.patches.push_back(t);
        }
      }
    }
    rwg.clear();
    for (auto& kv : e2p) {
      if (kv.second.patches.size() == 1) continue;
      RWG r;
      r.n1 = kv.first.first;
      r.n2 = kv.first.second;
      r.length = kv.second.length;
      r.tp = kv.second.patches[0];
      r.tm = kv.second.patches[1];
      r.fp = thirdNode(r.tp, r.n1, r.n2);
      r.fm = thirdNode(r.tm, r.n1, r.n2);
      rwg.push_back(r);
    }
  }
};

// kernel base (include/kernel.hpp:13-47, src/kernel.cpp:11-61).
struct Kernel {
  Phys ph;
  Counters* ctr = nullptr;
  virtual ~Kernel() {}
  virtual int size() const = 0;
  virtual cd entry(int i, int j) const = 0;
  Mat getMatrix(const std::vector<int>& rows, const std::vector<int>& cols) const {
    Mat m((int)rows.size(), (int)cols.size());
    const int nr = (int)rows.size(), nc = (int)cols.size();
#pragma omp parallel for schedule(dynamic, 4)
    for (int j = 0; j < nc; ++j)
      for (int i = 0; i < nr; ++i) m(i, j) = entry(rows[i], cols[j]);
    if (ctr) ctr->entries += (long long)nr * nc;
    return m;
  }
  std::vector<cd> getRow(int i, const std::vector<int>& cols) const {
    std::vector<cd> v(cols.size());
    const int nc = (int)cols.size();
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nc; ++j) v[j] = entry(i, cols[j]);
    if (ctr) ctr->entries += nc;
    return v;
  }
  std::vector<cd> getCol(int j, const std::vector<int>& rows) const {
    std::vector<cd> v(rows.size());
    const int nr = (int)rows.size();
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nr; ++i) v[i] = entry(rows[i], j);
    if (ctr) ctr->entries += nr;
    return v;
  }
};

// Triangle quadrature constants, E:284-286 (truncated as in the reference).
static const double Q7A[7] = {0.3333333, 0.0597158, 0.470142, 0.470142, 0.7974269, 0.1012865, 0.1012865};
static const double Q7B[7] = {0.3333333, 0.470142, 0.0597158, 0.470142, 0.1012865, 0.7974269, 0.1012865};
static const double Q7W[7] = {0.225, 0.13239415278851, 0.13239415278851, 0.13239415278851,
                              0.12593918054483, 0.12593918054483, 0.12593918054483};

// Wilton/Graglia closed form of the integral of 1/R over a flat triangle, as coded at
// E:478-577 (vertex 1 = a, the edges a->b, b->c, c->a).
inline double analyticLaplaceTri(const V3& tc, const V3& a, const V3& b, const V3& c) {
  V3 e2 = b - a, e3 = c - a, e1 = tc - a;
  V3 n = e2.cross(e3);
  n = n * (-1.0 / n.norm());
  double d = e1.dot(n);
  V3 cp = tc - n * d;
  d = std::abs(d);
  double s = 0.0;
  const V3* P1[3] = {&a, &b, &c};
  const V3* P2[3] = {&b, &c, &a};
  for (int i = 0; i < 3; ++i) {
    const V3 &p1 = *P1[i], &p2 = *P2[i];
    double len = (p1 - p2).norm();
    double Rp = (p1 - tc).norm();
    double Rm = (p2 - tc).norm();
    V3 Pp = p1 - cp, Pm = p2 - cp;
    V3 ul = (p1 - p2) / len;
    double Lp = Pp.dot(ul), Lm = Pm.dot(ul);
    V3 uu = ul.cross(n);
    double P0 = Pp.dot(uu);
    if (P0 < 0) P0 = -P0;
    if (P0 < 1.0e-10 * len) continue;
    double R0 = std::sqrt(P0 * P0 + d * d);
    V3 up = Pp - ul * Lp;
    up = up / up.norm();
    double T1 = P0 * std::log((Rp + Lp) / (Rm + Lm));
    if (P0 < 1e-6 * len && R0 < 1e-4 * len) T1 = 0;
    double T2 = std::atan(P0 * Lp / (R0 * R0 + d * Rp));
    double T3 = std::atan(P0 * Lm / (R0 * R0 + d * Rm));
    double T4 = up.dot(uu);
    s += T4 * (T1 - d * (T2 - T3));
  }
  return s;
}

// Integral of (r' - p)/R over the triangle, p = projection of tc on its plane (E:579-668).
inline V3 analyticRhoLaplaceTri(const V3& tc, const V3& a, const V3& b, const V3& c) {
  V3 e2 = b - a, e3 = c - a, e1 = tc - a;
  V3 n = e2.cross(e3);
  n = n * (-1.0 / n.norm());
  double d = e1.dot(n);
  V3 cp = tc - n * d;
  d = std::abs(d);
  V3 s(0, 0, 0);
  const V3* P1[3] = {&a, &b, &c};
  const V3* P2[3] = {&b, &c, &a};
  for (int i = 0; i < 3; ++i) {
    const V3 &p1 = *P1[i], &p2 = *P2[i];
    double Rp = (p1 - tc).norm();
    double Rm = (p2 - tc).norm();
    V3 Pp = p1 - cp, Pm = p2 - cp;
    double len = (p1 - p2).norm();
    V3 ul = (p1 - p2) / len;
    double Lp = Pp.dot(ul), Lm = Pm.dot(ul);
    V3 uu = ul.cross(n);
    double P0 = Pp.dot(uu);
    if (P0 < 0) P0 = -P0;
    double sqR0 = P0 * P0 + d * d;
    double T1 = ((Rm + Lm) >= 1.0e-10) ? sqR0 * std::log((Rp + Lp) / (Rm + Lm)) : 0.0;
    double T2 = Lp * Rp - Lm * Rm;
    s = s + uu * (0.5 * (T1 + T2));
  }
  return s;
}

// RWG EFIE (src/EFIEMatrix.cpp).
struct EFIE : Kernel {
  const Mesh* m;
  int etaAnalInteg = 4;  // E:10
  explicit EFIE(const Mesh* mesh, double freq) : m(mesh) { ph.freq = freq; }
  int size() const override { return (int)m->rwg.size(); }

  // E:478-494: vertex 1 is the free node, vertices 2/3 follow patch order.
  void freeOrder(int tri, int fv, V3& a, V3& b, V3& c) const {
    const Patch& p = m->patches[tri];
    int i2, i3;
    if (p.v1 == fv) { i2 = p.v2; i3 = p.v3; }
    else if (p.v2 == fv) { i2 = p.v1; i3 = p.v3; }
    else { i2 = p.v1; i3 = p.v2; }
    a = m->nodes[fv];
    b = m->nodes[i2];
    c = m->nodes[i3];
  }
  double analyticLaplace(const V3& tc, int tri, int fv) const {
    V3 a, b, c;
    freeOrder(tri, fv, a, b, c);
    return analyticLaplaceTri(tc, a, b, c);
  }
  V3 analyticRhoLaplace(const V3& tc, int tri, int fv) const {
    V3 a, b, c;
    freeOrder(tri, fv, a, b, c);
    return analyticRhoLaplaceTri(tc, a, b, c);
  }
  // E:457-476: projection onto the plane through the ORIGIN with the source normal.
  // physicalProjection = true projects onto the source triangle's own plane instead
  // (a test-only switch that isolates this quirk in the Mie-series check).
  bool physicalProjection = false;
  V3 projection(const V3& tc, int tri) const {
    const Patch& p = m->patches[tri];
    const V3 &a = m->nodes[p.v1], &b = m->nodes[p.v2], &c = m->nodes[p.v3];
    V3 n = (b - a).cross(c - a);
    n = n * (-1.0 / n.norm());
    double d = physicalProjection ? (tc - a).dot(n) : tc.dot(n);
    return tc - n * d;
  }
  // includes.cpp:13-21
  cd G(const V3& ri, const V3& rj) const {
    double R = (ri - rj).norm();
    return (-1.0 / 4.0 / PI) * std::exp(-I1 * ph.kappa() * R) / R;
  }
  cd Gns(const V3& ri, const V3& rj) const {
    double R = (ri - rj).norm();
    return (-1.0 / 4.0 / PI) * (std::exp(-I1 * ph.kappa() * R) - 1.0) / R;
  }
  // E:670-685
  cd scalarIntegration(const V3& tc, int tri, int fv) const {
    const Patch& p = m->patches[tri];
    double dist = (tc - p.centroid).norm();
    const V3 &v1 = m->nodes[p.v1], &v2 = m->nodes[p.v2], &v3 = m->nodes[p.v3];
    if (dist < etaAnalInteg * p.largestEdge) {
      cd sum = 0.0;  // E:282-306 with the R->0 limit of the non-singular kernel
      for (int q = 0; q < 7; ++q) {
        V3 pt = v1 + (v2 - v1) * Q7A[q] + (v3 - v1) * Q7B[q];
        if ((pt - tc).norm() < 1e-12) sum += Q7W[q] * (I1 * ph.kappa()) / 4.0 / PI;
        else sum += Q7W[q] * Gns(tc, pt);
      }
      return (-1.0 / 4.0 / PI) * analyticLaplace(tc, tri, fv) + p.area * sum;
    } else if (dist < 10.0 * p.largestEdge) {
      cd sum = 0.0;
      for (int q = 0; q < 7; ++q) {
        V3 pt = v1 + (v2 - v1) * Q7A[q] + (v3 - v1) * Q7B[q];
        sum += Q7W[q] * G(tc, pt);
      }
      return p.area * sum;
    } else {
      return p.area * G(tc, p.centroid);  // E:270-274
    }
  }
  // E:386-446: non-singular part always 7-point (vertex order v2, v3, v1), singular
  // part analytic / 7-point / 1-point by distance.
  C3 vectorIntegration(const V3& tc, int tri, int fv) const {
    const Patch& p = m->patches[tri];
    const V3 &w1 = m->nodes[p.v2], &w2 = m->nodes[p.v3], &w3 = m->nodes[p.v1];
    const V3& r3 = m->nodes[fv];
    C3 ns;
    for (int q = 0; q < 7; ++q) {
      V3 pt = w1 + (w2 - w1) * Q7A[q] + (w3 - w1) * Q7B[q];
      double R = (tc - pt).norm();
      V3 rho = pt - r3;
      cd f;
      if (R < 1e-12) f = (-1.0 / 4.0 / PI) * (-I1 * ph.kappa());
      else f = (-1.0 / 4.0 / PI) * (std::exp(-I1 * ph.kappa() * R) - 1.0) / R;
      ns = ns + toC(rho) * (Q7W[q] * f);
    }
    ns = ns * cd(p.area);
    double dist = (tc - p.centroid).norm();
    V3 sg;
    if (dist < etaAnalInteg * p.largestEdge) {
      V3 proj = projection(tc, tri);
      V3 a = analyticRhoLaplace(tc, tri, fv);
      double l = analyticLaplace(tc, tri, fv);
      sg = (a + (proj - r3) * l) * (-1.0 / 4.0 / PI);
    } else if (dist < 10.0 * p.largestEdge) {
      V3 s(0, 0, 0);
      for (int q = 0; q < 7; ++q) {
        V3 pt = w1 + (w2 - w1) * Q7A[q] + (w3 - w1) * Q7B[q];
        double R = (tc - pt).norm();
        s = s + (pt - r3) * (Q7W[q] * (-1.0 / 4.0 / PI) / R);
      }
      sg = s * p.area;
    } else {
      double R = (tc - p.centroid).norm();
      sg = (p.centroid - r3) * ((-1.0 / 4.0 / PI) / R) * p.area;
    }
    return ns + toC(sg);
  }
  // E:448-455
  cd phiInt(int tt, int st, int sfv) const {
    const Patch& t = m->patches[tt];
    return t.area * scalarIntegration(t.centroid, st, sfv);
  }
  // E:375-384 (plain transpose product, no conjugation)
  cd aInt(int tt, int st, int tfv, int sfv) const {
    const Patch& t = m->patches[tt];
    V3 rho = t.centroid - m->nodes[tfv];
    C3 v = vectorIntegration(t.centroid, st, sfv);
    return t.area * (rho.x * v.x + rho.y * v.y + rho.z * v.z);
  }
  // E:698-783
  cd entry(int i, int j) const override {
    const RWG& te = m->rwg[i];
    const RWG& se = m->rwg[j];
    const double Lt = te.length, Ls = se.length;
    const double Atp = m->patches[te.tp].area, Atm = m->patches[te.tm].area;
    const double Asp = m->patches[se.tp].area, Asm = m->patches[se.tm].area;
    const double w = 2.0 * PI * ph.freq;
    cd phiPP = -Lt * Ls / (I1 * w * ph.epsilon * Atp * Asp);
    cd phiPM = Lt * Ls / (I1 * w * ph.epsilon * Atp * Asm);
    cd phiMP = Lt * Ls / (I1 * w * ph.epsilon * Atm * Asp);
    cd phiMM = -Lt * Ls / (I1 * w * ph.epsilon * Atm * Asm);
    cd e = phiPP * phiInt(te.tp, se.tp, se.fp) + phiPM * phiInt(te.tp, se.tm, se.fm) +
           phiMP * phiInt(te.tm, se.tp, se.fp) + phiMM * phiInt(te.tm, se.tm, se.fm);
    cd APP = -(I1 * w * ph.mu * Lt * Ls) / (4.0 * Atp * Asp);
    cd APM = (I1 * w * ph.mu * Lt * Ls) / (4.0 * Atp * Asm);
    cd AMP = (I1 * w * ph.mu * Lt * Ls) / (4.0 * Atm * Asp);
    cd AMM = -(I1 * w * ph.mu * Lt * Ls) / (4.0 * Atm * Asm);
    e += APP * aInt(te.tp, se.tp, te.fp, se.fp) + APM * aInt(te.tp, se.tm, te.fp, se.fm) +
         AMP * aInt(te.tm, se.tp, te.fm, se.fp) + AMM * aInt(te.tm, se.tm, te.fm, se.fm);
    return e;
  }
  // Plane-wave excitation, E:792-829. k_v = (kappa sin(t) cos(p), kappa sin(t) sin(p),
  // cos(t)) as coded (phaseQuirk=1, the reference hard-codes t = p = 0), or the physical
  // kappa*(...) phase (phaseQuirk=0).
  cd rhs(const V3& Einc, const V3& kdir, int i) const {
    const RWG& te = m->rwg[i];
    const Patch &P = m->patches[te.tp], &M = m->patches[te.tm];
    V3 rp = P.centroid - m->nodes[te.fp];
    V3 rm = M.centroid - m->nodes[te.fm];
    double kp = kdir.dot(P.centroid), km = kdir.dot(M.centroid);
    cd ep = std::exp(I1 * kp), em = std::exp(I1 * km);
    return te.length / 2.0 * (Einc.dot(rp) * ep - Einc.dot(rm) * em);
  }
  // E:1100-1130
  void jAtCentroids(const std::vector<cd>& coef, std::vector<C3>& J) const {
    J.assign(m->patches.size(), C3());
    for (size_t i = 0; i < coef.size(); ++i) {
      const RWG& te = m->rwg[i];
      double Ap = m->patches[te.tp].area, An = m->patches[te.tm].area;
      V3 rp = m->patches[te.tp].centroid - m->nodes[te.fp];
      V3 rn = m->patches[te.tm].centroid - m->nodes[te.fm];
      J[te.tp] = J[te.tp] + toC(rp) * (coef[i] * te.length / 2.0 / Ap);
      J[te.tm] = J[te.tm] + toC(rn) * (-coef[i] * te.length / 2.0 / An);
    }
  }
  // E:916-930 (neta = 377 as an int)
  double rcs(double theta, double phi, const std::vector<C3>& J) const {
    const int neta = 377;
    cd st = 0, sp = 0;
    double k = ph.kappa();
    V3 kv(k * std::sin(theta) * std::cos(phi), k * std::sin(theta) * std::sin(phi), k * std::cos(theta));
    for (size_t j = 0; j < m->patches.size(); ++j) {
      cd ph_ = std::exp(I1 * kv.dot(m->patches[j].centroid));
      double A = m->patches[j].area;
      st += (J[j].x * std::cos(theta) * std::cos(phi) + J[j].y * std::cos(theta) * std::sin(phi) -
             J[j].z * std::sin(theta)) * ph_ * A;
      sp += (-J[j].x * std::sin(phi) + J[j].y * std::cos(phi)) * ph_ * A;
    }
    double a = k * k / (4.0 * PI) * neta * std::abs(st) * neta * std::abs(st);
    double b = k * k / (4.0 * PI) * neta * std::abs(sp) * neta * std::abs(sp);
    return a + b;
  }
};

// Piecewise-constant Laplace single layer, src/ElectrostaticsMatrix.cpp:293-340.
struct Laplace : Kernel {
  const Mesh* m;
  explicit Laplace(const Mesh* mesh) : m(mesh) {}
  int size() const override { return (int)m->patches.size(); }
  cd entry(int i, int j) const override {
    const Patch& pi = m->patches[i];
    const Patch& pj = m->patches[j];
    const V3 &a = m->nodes[pj.v1], &b = m->nodes[pj.v2], &c = m->nodes[pj.v3];
    double d = (pi.centroid - pj.centroid).norm();
    double v;
    if (d < 4.0 * pj.largestEdge) {
      v = ph.C0() * analyticLaplaceTri(pi.centroid, a, b, c);
    } else if (d < 10.0 * pj.largestEdge) {
      double s = 0;
      for (int q = 0; q < 7; ++q) {
        V3 pt = a + (b - a) * Q7A[q] + (c - a) * Q7B[q];
        s += pj.area * Q7W[q] * (1.0 / (pi.centroid - pt).norm());
      }
      v = ph.C0() * s;
    } else {
      v = ph.C0() * (pj.area * (1.0 / (pi.centroid - pj.centroid).norm()));
    }
    return cd(v, 0.0);
  }
};

}  // namespace oracle
