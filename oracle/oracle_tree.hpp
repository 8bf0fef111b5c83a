// ORACLE — test infrastructure only (see oracle_core.hpp).
//
// FMM2DTree<K> of include/AIFMMTree2D_3D_memory_optimized.hpp (the variant the
// reference compiles, include/AIFMM.hpp:9) and the AIFMM<K> facade (src/AIFMM.cpp),
// restated: tree setup, nested ACA skeletons, L2P/P2M, M2L and leaf P2P assembly,
// the serial level-by-level elimination with RRQR fill-in recompression, the rhs
// elimination and the back substitution. Vectors are n x 1 Mats.
//
// Two structural changes leave the arithmetic unchanged: the LU factors of P2P_kk
// and L2M_f computed in the factorization are kept and reused by the solve (the
// reference recomputes them from the same, unchanged matrices at T:3622, T:4120,
// T:4825, T:4843), and solve(P2P[k][pp]) is formed once per pp instead of once per
// (nn, pp) (T:3934-4019 recompute the identical product).
#pragma once
#include <cstdio>
#include <cstdlib>
#include "oracle_lowrank.hpp"

namespace oracle {

struct SolverInputs {  // include/includes.hpp:73-77
  double xmin, xmax, ymin, ymax, zmin, zmax;
  int leafSize, tol_pow, fillin_rank, fillin_tol;
  std::vector<V3> locations;
  int variant = 0;  // 0: memory_optimized tree (T); 1: AIFMMTree2D_3D parameters (T3)
};

struct Box {  // include/AIFMMBox.hpp:13-71
  bool ILActive = true;
  double maxPivot_L2P = 0.0, maxPivot_P2M = 0.0;
  V3 center;
  std::vector<int> interactionList, neighborList;
  bool Eliminated = false;
  int NumMultipoles = 0, NumLocals = 0;
  Mat particles, multipoles, locals, particle_rhs, multipole_rhs, local_rhs;
  std::map<int, Mat> M2L, P2P, M2P, P2L, L2P_f, P2M_f, M2M_f;
  Mat L2M_f, L2P, P2M;
  std::vector<int> chargeLocations;
  std::vector<int> incoming_chargePoints, incoming_checkPoints, outgoing_chargePoints, outgoing_checkPoints;
  Mat Ac, Ar;
  LU luP2P, luL2M;
  bool hasLuP2P = false, hasLuL2M = false;
};

inline Mat colvec(int n) { return Mat(n, 1); }

struct Tree {
  const Kernel* K;
  Counters ctr;
  int nLevels, N, startLevel = 0;
  double L, eta;
  std::vector<int> nBoxesPerLevel;
  std::vector<double> boxSideLength;
  std::vector<std::vector<Box>> tree;
  int TOL_POW, fill_in_TOL_POW;
  double fillInTolerance;
  int fillin_rankVec[32];
  int variant;
  std::vector<std::pair<int, int>> P2Plist, P2L_M2P;
  std::vector<V3> particles;
  double dom[6];
  int nChild = 8;

  // T:52-165
  Tree(const SolverInputs& in, const Kernel* k) : K(k) {
    particles = in.locations;
    variant = in.variant;
    TOL_POW = in.tol_pow;
    fill_in_TOL_POW = in.fillin_tol;
    fillInTolerance = std::pow(10.0, -1.0 * TOL_POW);
    dom[0] = in.xmin; dom[1] = in.xmax; dom[2] = in.ymin; dom[3] = in.ymax; dom[4] = in.zmin; dom[5] = in.zmax;
    L = std::max({in.xmax - in.xmin, in.ymax - in.ymin, in.zmax - in.zmin});
    N = (int)in.locations.size();
    int q = N / in.leafSize;  // integer division, as at T:132
    int lv = q > 0 ? (int)std::floor(std::log((double)q) / std::log(8.0)) : 0;
    if (variant == 0) {
      nLevels = std::max(2, lv);
      eta = std::sqrt(3.0) * 2;
    } else {  // T3:131-132
      nLevels = std::max(3, lv);
      eta = std::sqrt(3.0);
    }
    nBoxesPerLevel.push_back(1);
    boxSideLength.push_back(L);
    for (int j = 1; j <= nLevels; ++j) {
      nBoxesPerLevel.push_back(nChild * nBoxesPerLevel[j - 1]);
      boxSideLength.push_back(0.5 * boxSideLength[j - 1]);
    }
    for (int j = 0; j < 32; ++j) fillin_rankVec[j] = 0;
    int temp = 5, i = 0;  // T:148-156
    for (int j = nLevels; j >= 2; --j) {
      fillin_rankVec[j] = temp;
      ++i;
      if (i % 3 == 0) temp += 5;
    }
    if (variant == 1)
      for (int j = 0; j < 32; ++j) fillin_rankVec[j] = 1 << 28;  // T3: no fill-in rank cap
  }

  Box& B(int j, int k) { return tree[j][k]; }

  // getAvgLeafSize T:490-501 (leaf boxes holding points), getAvgMultipoles T:531-551 (boxes
  // with P2M rows and an interaction list, levels nLevels..startLevel), integer averages.
  int getAvgLeafSize() const {
    long long avg = 0, n = 0;
    for (const Box& b : tree[nLevels])
      if (!b.chargeLocations.empty()) { ++n; avg += (long long)b.chargeLocations.size(); }
    return n ? (int)(avg / n) : 0;
  }
  int getAvgMultipoles() const {
    long long avg = 0, n = 0;
    for (int j = nLevels; j >= startLevel; --j)
      for (const Box& b : tree[j])
        if (b.P2M.rows() > 0 && !b.interactionList.empty()) { ++n; avg += b.P2M.rows(); }
    return n ? (int)(avg / n) : 0;
  }
  static double entriesOf(const std::map<int, Mat>& m, int key) {
    auto it = m.find(key);
    return it == m.end() ? 0.0 : (double)it->second.size();
  }
  // findMemory T:4511-4539: L2P and P2M of every box, M2L over interaction lists, P2P over
  // neighbour lists at the leaf level; entries x 8 bytes, in GB.
  double findMemory() const {
    double sum = 0;
    for (int j = nLevels; j >= startLevel; --j)
      for (const Box& b : tree[j]) sum += (double)b.L2P.size() + (double)b.P2M.size();
    for (int j = nLevels; j >= startLevel; --j)
      for (const Box& b : tree[j])
        for (int n : b.interactionList) sum += entriesOf(b.M2L, n);
    for (const Box& b : tree[nLevels])
      for (int n : b.neighborList) sum += entriesOf(b.P2P, n);
    return sum * 8 * 1e-9;
  }
  // findMemoryAfterFactorize T:4541-4571: L2P and P2M, and P2P over neighbour lists at every
  // level (the interaction-list part is commented out there).
  double findMemoryAfterFactorize() const {
    double sum = 0;
    for (int j = nLevels; j >= startLevel; --j)
      for (const Box& b : tree[j]) sum += (double)b.L2P.size() + (double)b.P2M.size();
    for (int j = nLevels; j >= startLevel; --j)
      for (const Box& b : tree[j])
        for (int n : b.neighborList) sum += entriesOf(b.P2P, n);
    return sum * 8 * 1e-9;
  }

  void createTree() {  // T:167-192
    tree.assign(nLevels + 1, std::vector<Box>());
    for (int j = 0; j <= nLevels; ++j) tree[j].resize(nBoxesPerLevel[j]);
  }
  void assign_Center_Location() {  // T:425-463
    B(0, 0).center = V3((dom[0] + dom[1]) / 2.0, (dom[2] + dom[3]) / 2.0, (dom[4] + dom[5]) / 2.0);
    static const int sx[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
    static const int sy[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
    static const int sz[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
    for (int j = 0; j < nLevels; ++j) {
      double shift = 0.5 * boxSideLength[j + 1];
      for (int k = 0; k < nBoxesPerLevel[j]; ++k) {
        const V3& c = B(j, k).center;
        for (int l = 0; l < 8; ++l)
          B(j + 1, 8 * k + l).center = V3(c.x + sx[l] * shift, c.y + sy[l] * shift, c.z + sz[l] * shift);
      }
    }
  }
  bool admissibilityCondition(int j, int child, int other) {  // T:277-293
    V3 d = B(j, child).center - B(j, other).center;
    double dist = d.norm();
    double etaLocal = eta;
    if (j == nLevels) etaLocal = eta / 2.0;
    if (dist > etaLocal * boxSideLength[j] + boxSideLength[nLevels] / 2.0) {
      B(j, child).interactionList.push_back(other);
      return true;
    }
    B(j, child).neighborList.push_back(other);
    return false;
  }
  void assign_Tree_Interactions() {  // T:250-331
    B(0, 0).neighborList.push_back(0);
    for (int j = 0; j < nLevels; ++j)
      for (int k = 0; k < nBoxesPerLevel[j]; ++k)
        for (int i = 0; i < nChild; ++i) {
          int child = nChild * k + i;
          for (int nn : B(j, k).neighborList)
            for (int c = 0; c < nChild; ++c) admissibilityCondition(j + 1, child, nn * nChild + c);
        }
    int flag = -1;
    for (int j = 0; j <= nLevels; ++j)
      for (int k = 0; k < nBoxesPerLevel[j]; ++k)
        if (flag != 0 && !B(j, k).interactionList.empty()) { startLevel = j; flag = 0; }
  }
  void assignChargeLocations() {  // T:576-625
    for (int i = 0; i < N; ++i) B(0, 0).chargeLocations.push_back(i);
    for (int j = 0; j < nLevels; ++j)
      for (int k = 0; k < nBoxesPerLevel[j]; ++k) {
        const V3& c = B(j, k).center;
        int Kp = 8 * k;
        for (int index : B(j, k).chargeLocations) {
          const V3& p = particles[index];
          int ch;
          if (p.z <= c.z) {
            if (p.x <= c.x) ch = (p.y <= c.y) ? 0 : 3;
            else ch = (p.y <= c.y) ? 1 : 2;
          } else {
            if (p.x <= c.x) ch = (p.y <= c.y) ? 4 : 7;
            else ch = (p.y <= c.y) ? 5 : 6;
          }
          B(j + 1, Kp + ch).chargeLocations.push_back(index);
        }
      }
  }
  void assignNonLeafChargeLocations() {  // T:659-668
    for (int j = nLevels - 1; j >= 0; --j)
      for (int k = 0; k < nBoxesPerLevel[j]; ++k) {
        auto& cl = B(j, k).chargeLocations;
        cl.clear();
        for (int i = 0; i < nChild; ++i) {
          auto& c = B(j + 1, nChild * k + i).chargeLocations;
          cl.insert(cl.end(), c.begin(), c.end());
        }
      }
  }

  // ---------------- skeletons (T:917-1256) ----------------
  void childrenPoints(int j, int k, bool incomingCheck, std::vector<int>& out) {
    if (j == nLevels) {
      out.insert(out.end(), B(j, k).chargeLocations.begin(), B(j, k).chargeLocations.end());
      return;
    }
    for (int c = 0; c < nChild; ++c) {
      Box& ch = B(j + 1, nChild * k + c);
      auto& v = incomingCheck ? ch.incoming_checkPoints : ch.outgoing_chargePoints;
      out.insert(out.end(), v.begin(), v.end());
    }
  }
  void getNodes_incoming_box(int j, int k) {  // T:1155-1256
    Box& b = B(j, k);
    std::vector<int> boxA, IL;
    childrenPoints(j, k, true, boxA);
    for (int kIL : b.interactionList) childrenPoints(j, kIL, false, IL);
    int n_rows = (int)boxA.size(), n_cols = (int)IL.size();
    if (n_rows > 0 && n_cols > 0) {
      ACAResult r = aca_only_nodes2(*K, std::pow(10.0, -TOL_POW), boxA, IL);
      b.Ac = r.Ac;
      if (r.rank > 0) {
        for (int rb : r.row_bases) b.incoming_checkPoints.push_back(boxA[rb]);
        for (int cb : r.col_bases) b.incoming_chargePoints.push_back(IL[cb]);
      }
    } else if (n_cols == 0) {
      b.ILActive = false;
      b.incoming_checkPoints = boxA;
    }
  }
  void getNodes_outgoing_box(int j, int k) {  // T:994-1077
    Box& b = B(j, k);
    std::vector<int> boxA, IL;
    childrenPoints(j, k, false, boxA);
    for (int kIL : b.interactionList) childrenPoints(j, kIL, true, IL);
    int n_rows = (int)IL.size(), n_cols = (int)boxA.size();
    if (n_cols > 0 && n_rows > 0) {
      ACAResult r = aca_only_nodes2(*K, std::pow(10.0, -TOL_POW), IL, boxA);
      b.Ar = r.Ar;
      if (r.rank > 0) {
        for (int rb : r.row_bases) b.outgoing_checkPoints.push_back(IL[rb]);
        for (int cb : r.col_bases) b.outgoing_chargePoints.push_back(boxA[cb]);
      }
    } else if (n_rows == 0) {
      b.ILActive = false;
      b.outgoing_chargePoints = boxA;
    }
  }
  static double maxColNorm(const Mat& A) {  // |R(0,0)| of a column-pivoted QR
    double best = 0;
    for (int j = 0; j < A.c; ++j) {
      double s = 0;
      for (int i = 0; i < A.r; ++i) s += std::norm(A(i, j));
      best = std::max(best, std::sqrt(s));
    }
    return best;
  }
  void get_L2P_P2M_box(int j, int k) {  // T:681-732
    Box& b = B(j, k);
    size_t nin = b.incoming_checkPoints.size(), nout = b.outgoing_chargePoints.size();
    if (nin < nout) {
      b.outgoing_checkPoints.resize(std::min(b.outgoing_checkPoints.size(), nin));
      b.outgoing_chargePoints.resize(nin);
    } else {
      b.incoming_checkPoints.resize(nout);
      if (b.incoming_chargePoints.size() > nout) b.incoming_chargePoints.resize(nout);
    }
    const int r = (int)b.incoming_checkPoints.size();
    if (b.ILActive) {
      Mat temp1 = b.Ac.block(0, 0, b.Ac.r, std::min(r, b.Ac.c));
      b.L2P = Mat(b.Ac.r, r);
      if (r > 0) {
        Mat D1 = K->getMatrix(b.incoming_checkPoints, b.incoming_chargePoints);
        LU lu(adjoint(D1));
        b.L2P = adjoint(lu.solve(adjoint(temp1)));
        ctr.getrf_flops += 8.0 / 3.0 * r * (double)r * r;
        ctr.getrs_flops += 8.0 * r * (double)r * b.Ac.r;
      }
      const int ro = (int)b.outgoing_checkPoints.size();
      Mat temp2 = b.Ar.block(0, 0, std::min(ro, b.Ar.r), b.Ar.c);
      b.P2M = Mat(ro, b.Ar.c);
      if (ro > 0) {
        Mat D2 = K->getMatrix(b.outgoing_checkPoints, b.outgoing_chargePoints);
        LU lu(D2);
        b.P2M = lu.solve(temp2);
        ctr.getrf_flops += 8.0 / 3.0 * ro * (double)ro * ro;
        ctr.getrs_flops += 8.0 * ro * (double)ro * b.Ar.c;
      }
    } else {
      b.L2P = Mat::Identity((int)b.incoming_checkPoints.size());
      b.P2M = Mat::Identity((int)b.outgoing_chargePoints.size());
    }
    if (b.Ac.size() > 0) b.maxPivot_L2P = maxColNorm(b.Ac.block(0, 0, b.Ac.r, std::min(r, b.Ac.c)));
    if (b.Ar.size() > 0) {
      int ro = (int)b.outgoing_checkPoints.size();
      b.maxPivot_P2M = maxColNorm(adjoint(b.Ar.block(0, 0, std::min(ro, b.Ar.r), b.Ar.c)));
    }
  }
  void getNodes() {  // T:917-923 (boxes are independent within a level)
    for (int j = nLevels; j >= startLevel; --j) {
      for (int k = 0; k < nBoxesPerLevel[j]; ++k) getNodes_incoming_box(j, k);
      for (int k = 0; k < nBoxesPerLevel[j]; ++k) getNodes_outgoing_box(j, k);
      for (int k = 0; k < nBoxesPerLevel[j]; ++k) get_L2P_P2M_box(j, k);
    }
  }
  void assemble_M2L() {  // T:1258-1273
    for (int j = startLevel; j <= nLevels; ++j)
      for (int k = 0; k < nBoxesPerLevel[j]; ++k)
        for (int kIL : B(j, k).interactionList)
          B(j, k).M2L[kIL] = K->getMatrix(B(j, k).incoming_checkPoints, B(j, kIL).outgoing_chargePoints);
  }
  void free_charge_and_checkPoints() {  // T:1275-1284
    for (int j = startLevel; j <= nLevels; ++j)
      for (int k = 0; k < nBoxesPerLevel[j]; ++k) {
        Box& b = B(j, k);
        b.outgoing_chargePoints.clear();
        b.incoming_checkPoints.clear();
        b.Ar = Mat();
        b.Ac = Mat();
      }
  }
  void initialise_phase() {  // T:3343-3370
    for (int j = startLevel; j <= nLevels; ++j)
      for (int k = 0; k < nBoxesPerLevel[j]; ++k) {
        Box& b = B(j, k);
        b.NumMultipoles = b.P2M.r;
        b.NumLocals = b.L2P.c;
        b.local_rhs = colvec(b.NumMultipoles);
        b.multipole_rhs = colvec(b.NumLocals);
      }
  }
  void initialise_P2P_Leaf_Level() {  // T:3372-3386
    int j = nLevels;
    for (int k = 0; k < nBoxesPerLevel[j]; ++k) {
      for (int nn : B(j, k).neighborList)
        if (nn != k) B(j, k).P2P[nn] = K->getMatrix(B(j, k).chargeLocations, B(j, nn).chargeLocations);
      B(j, k).P2P[k] = K->getMatrix(B(j, k).chargeLocations, B(j, k).chargeLocations);
    }
  }

  // ---------------- helpers with Eigen's shape rules ----------------
  Mat gemm(const Mat& A, const Mat& Bm) {
    ctr.gemm_flops += 8.0 * A.r * (double)A.c * Bm.c;
    return mul(A, Bm);
  }
  Mat solveLU(const LU& lu, const Mat& Bm) {
    ctr.getrs_flops += 8.0 * lu.n * (double)lu.n * Bm.c;
    return lu.solve(Bm);
  }
  // X = X - A*S, where an empty X (size 0) starts from zero (T:3936-3947 pattern)
  void subAssign(Mat& X, const Mat& prod) {
    if (X.size() == 0 && (X.r == 0 || X.c == 0) && !(X.r == prod.r && X.c == prod.c)) {
      X = neg(prod);
      return;
    }
    X = sub(X, prod);
  }
  void assignOrSub(Mat& X, const Mat& prod) {  // if (X.size()==0) X = -prod else X -= prod
    if (X.size() == 0) X = neg(prod);
    else X = sub(X, prod);
  }
  bool is_well_separated(int j, int k, int l) {  // T:4044-4055
    if (k == l) return false;
    for (int nn : B(j, k).neighborList)
      if (l == nn) return false;
    return true;
  }

  // ---------------- elimination (T:3413-4509, 4599-4747) ----------------
  void initialise_P2P_NonLeafLevel(int j) {  // T:3413-3476
    for (int k = 0; k < nBoxesPerLevel[j]; ++k) {
      for (int nn : B(j, k).neighborList) {
        if (nn == k) continue;
        int n_rows = 0, n_cols = 0;
        for (int a = 0; a < nChild; ++a) {
          int t = 0;
          for (int b = 0; b < nChild; ++b) t += B(j + 1, nChild * k + a).M2L[nChild * nn + b].c;
          n_cols = std::max(n_cols, t);
        }
        for (int b = 0; b < nChild; ++b) {
          int t = 0;
          for (int a = 0; a < nChild; ++a) t += B(j + 1, nChild * k + a).M2L[nChild * nn + b].r;
          n_rows = std::max(n_rows, t);
        }
        Mat P(n_rows, n_cols);
        int row_index = 0, col_index = 0;
        for (int a = 0; a < nChild; ++a) {
          for (int b = 0; b < nChild; ++b) {
            Mat& M = B(j + 1, nChild * k + a).M2L[nChild * nn + b];
            if (M.size() > 0) {
              if (row_index + M.r > P.r || col_index + M.c > P.c)
                throw std::runtime_error("oracle: P2P non-leaf block out of range");
              P.setBlock(row_index, col_index, M);
              col_index += M.c;
            }
          }
          row_index += B(j + 1, nChild * k + a).M2L[nChild * nn + 0].r;
          col_index = 0;
        }
        B(j, k).P2P[nn] = P;
      }
      {  // self
        int n_rows = 0, n_cols = 0;
        for (int a = 0; a < nChild; ++a) n_rows += B(j + 1, nChild * k + a).NumLocals;
        for (int b = 0; b < nChild; ++b) n_cols += B(j + 1, nChild * k + b).NumMultipoles;
        Mat P(n_rows, n_cols);
        int row_index = 0, col_index = 0;
        for (int a = 0; a < nChild; ++a) {
          Box& ca = B(j + 1, nChild * k + a);
          for (int b = 0; b < nChild; ++b) {
            Box& cb = B(j + 1, nChild * k + b);
            Mat& M = ca.M2L[nChild * k + b];
            if (M.r == ca.NumLocals && M.c == cb.NumMultipoles) P.setBlock(row_index, col_index, M);
            else if (!(M.size() == 0 && (size_t)ca.NumLocals * cb.NumMultipoles == 0))
              throw std::runtime_error("oracle: P2P self block size mismatch");
            col_index += cb.NumMultipoles;
          }
          row_index += ca.NumLocals;
          col_index = 0;
        }
        B(j, k).P2P[k] = P;
      }
    }
  }
  void deleteM2L(int j) {  // T:3478-3504
    for (int k = 0; k < nBoxesPerLevel[j]; ++k)
      for (int nn : B(j, k).neighborList)
        for (int a = 0; a < nChild; ++a)
          for (int b = 0; b < nChild; ++b) B(j + 1, nChild * k + a).M2L[nChild * nn + b] = Mat();
  }

  void eliminate_x(int j, int k) {  // T:3561-3611, 3653-3791, 3813-4042
    Box& bk = B(j, k);
    Mat& Pkk = bk.P2P[k];
    if (Pkk.size() != 0) {
      bk.luP2P.compute(Pkk);
      bk.hasLuP2P = true;
      ctr.getrf_flops += 8.0 / 3.0 * Pkk.r * (double)Pkk.r * Pkk.r;
      const LU& lu = bk.luP2P;
      Mat X_L2P = solveLU(lu, bk.L2P);
      bk.L2M_f = neg(gemm(bk.P2M, X_L2P));
      std::map<int, Mat> Xn;  // solve(P2P[k][pp]) or solve(M2P[k][pp]) per neighbor pp
      for (int nn : bk.neighborList) {
        if (nn == k) continue;
        if (!B(j, nn).Eliminated) Xn[nn] = solveLU(lu, bk.P2P[nn]);
        else Xn[nn] = solveLU(lu, bk.M2P[nn]);
        if (!B(j, nn).Eliminated) bk.P2M_f[nn] = neg(gemm(bk.P2M, Xn[nn]));
        else bk.M2M_f[nn] = neg(gemm(bk.P2M, Xn[nn]));
      }
      for (int nn : bk.neighborList) {
        if (nn == k) continue;
        Box& bn = B(j, nn);
        const Mat& Ank = bn.Eliminated ? bn.P2L[k] : bn.P2P[k];
        bn.L2P_f[k] = neg(gemm(Ank, X_L2P));
        for (int pp : bk.neighborList) {
          if (pp == k) continue;
          Mat prod = gemm(Ank, Xn[pp]);
          if (!bn.Eliminated) {
            if (!B(j, pp).Eliminated) assignOrSub(bn.P2P[pp], prod);
            else assignOrSub(bn.M2P[pp], prod);
          } else {
            if (!B(j, pp).Eliminated) assignOrSub(bn.P2L[pp], prod);
            else assignOrSub(bn.M2L[pp], prod);
          }
        }
      }
    } else {  // T:3656-3791: shapes only
      bk.L2M_f = Mat(bk.P2M.r, bk.L2P.c);
      for (int nn : bk.neighborList) {
        if (nn == k) continue;
        if (!B(j, nn).Eliminated) bk.P2M_f[nn] = Mat(bk.P2M.r, bk.P2P[nn].c);
        else bk.M2M_f[nn] = Mat(bk.P2M.r, bk.M2P[nn].c);
      }
      for (int nn : bk.neighborList) {
        if (nn == k) continue;
        Box& bn = B(j, nn);
        if (!bn.Eliminated) bn.L2P_f[k] = Mat(bn.P2P[k].r, bk.L2P.c);
        else bn.L2P_f[k] = Mat(bn.P2L[k].r, bk.L2P.c);
        for (int pp : bk.neighborList) {
          if (pp == k) continue;
          if (!bn.Eliminated) {
            if (!B(j, pp).Eliminated) {
              if (bn.P2P[pp].size() == 0) bn.P2P[pp] = Mat(bn.P2P[k].r, bk.P2P[pp].c);
            } else {
              if (bn.M2P[pp].size() == 0) bn.M2P[pp] = Mat(bn.P2P[k].r, bk.M2P[pp].c);
            }
          } else {
            if (!B(j, pp).Eliminated) {
              if (bn.P2L[pp].size() == 0) bn.P2L[pp] = Mat(bn.P2L[k].r, bk.P2P[pp].c);
            } else {
              (void)bn.M2L[pp];
            }
          }
        }
      }
    }
  }

  void eliminate_z(int j, int k) {  // T:4081-4113, 4150-4509
    Box& bk = B(j, k);
    if (bk.L2M_f.size() > 0) {
      bk.luL2M.compute(bk.L2M_f);
      bk.hasLuL2M = true;
      ctr.getrf_flops += 8.0 / 3.0 * bk.L2M_f.r * (double)bk.L2M_f.r * bk.L2M_f.r;
      const LU& lu = bk.luL2M;
      Mat minusI = neg(Mat::Identity(bk.NumMultipoles));
      Mat Yself = solveLU(lu, minusI);
      bk.M2L[k] = Yself;
      std::map<int, Mat> Y;
      for (int nn : bk.neighborList) {
        if (nn == k) continue;
        if (!B(j, nn).Eliminated) {
          Y[nn] = solveLU(lu, bk.P2M_f[nn]);
          bk.P2L[nn] = Y[nn];
          if (is_well_separated(j, k, nn)) {  // T:4342 (never true for a neighbor)
            double mn = bk.M2L[nn].norm();
            bool push = variant == 0 ? (mn > 0 && bk.P2L[nn].norm() / mn != 0.0)
                                     : (mn > 0 && bk.P2L[nn].norm() / mn > fillInTolerance);
            if (push) P2L_M2P.push_back({k, nn});
          }
        } else {
          Y[nn] = solveLU(lu, bk.M2M_f[nn]);
          bk.M2L[nn] = Y[nn];
        }
      }
      for (int nn : bk.neighborList) {
        if (nn == k) continue;
        Box& bn = B(j, nn);
        Mat& Lf = bn.L2P_f[k];
        Mat py = gemm(Lf, Yself);
        if (!bn.Eliminated) assignOrSub(bn.M2P[k], py);
        else assignOrSub(bn.M2L[k], py);
        for (int pp : bk.neighborList) {
          if (pp == k) continue;
          Mat prod = gemm(Lf, Y[pp]);
          if (!bn.Eliminated) {
            if (!B(j, pp).Eliminated) {
              bn.P2P[pp] = sub(bn.P2P[pp], prod);
              if (is_well_separated(j, nn, pp) && nn < pp) {
                double mn = bn.M2L[pp].norm();
                if (mn > 0 && bn.P2P[pp].norm() / mn > fillInTolerance) P2Plist.push_back({nn, pp});
              }
            } else {
              bn.M2P[pp] = sub(bn.M2P[pp], prod);
            }
          } else {
            if (!B(j, pp).Eliminated) {
              bn.P2L[pp] = sub(bn.P2L[pp], prod);
              if (is_well_separated(j, nn, pp)) {
                double mn = bn.M2L[pp].norm();
                if (mn > 0 && bn.P2L[pp].norm() / mn > fillInTolerance) P2L_M2P.push_back({nn, pp});
              }
            } else {
              bn.M2L[pp] = sub(bn.M2L[pp], prod);
            }
          }
        }
      }
    } else {  // T:4150-4296
      bk.M2L[k] = Mat(0, bk.NumMultipoles);
      for (int nn : bk.neighborList) {
        if (nn == k) continue;
        if (!B(j, nn).Eliminated) bk.P2L[nn] = Mat(0, bk.P2M_f[nn].c);
        else bk.M2L[nn] = Mat(0, bk.M2M_f[nn].c);
      }
      for (int nn : bk.neighborList) {
        if (nn == k) continue;
        Box& bn = B(j, nn);
        if (!bn.Eliminated) {
          if (bn.M2P[k].size() == 0) bn.M2P[k] = Mat(bn.L2P_f[k].r, bk.NumMultipoles);
        } else {
          if (bn.M2L[k].size() == 0) bn.M2L[k] = Mat(bn.L2P_f[k].r, bk.NumMultipoles);
        }
      }
    }
  }

  // ---------------- fill-in recompression (T:1573-1656, 2116-2306, 2454-2596, 2878-2959) --------
  void update_other_M2Ls_at_k_old(int j, int k, int b, const Mat& newL2P) {  // T:3159-3188
    Box& bk = B(j, k);
    Mat R = gemm(adjoint(newL2P), bk.L2P);
    for (auto& kv : bk.P2L) kv.second = gemm(R, kv.second);
    for (int kIL : bk.interactionList) {
      if (kIL == b) continue;
      bk.M2L[kIL] = gemm(R, bk.M2L[kIL]);
    }
    for (int kN : bk.neighborList)
      if (kN != k && B(j, kN).Eliminated && bk.Eliminated) bk.M2L[kN] = gemm(R, bk.M2L[kN]);
    if (bk.Eliminated) bk.M2L[k] = gemm(R, bk.M2L[k]);
  }
  void update_other_M2Ls_at_b_old(int j, int k, int b, const Mat& newP2Mt) {  // T:3252-3288
    Box& bb = B(j, b);
    Mat R = gemm(bb.P2M, newP2Mt);
    for (auto& kv : bb.P2L) {
      Mat& M = B(j, kv.first).M2P[b];
      M = gemm(M, R);
    }
    for (int kIL : bb.interactionList) {
      if (kIL == k) continue;
      Mat& M = B(j, kIL).M2L[b];
      M = gemm(M, R);
    }
    for (int kN : bb.neighborList)
      if (kN != b && B(j, kN).Eliminated && bb.Eliminated) {
        Mat& M = B(j, kN).M2L[b];
        M = gemm(M, R);
      }
    if (bb.Eliminated) bb.M2L[b] = gemm(bb.M2L[b], R);
  }
  void update_parent_L2P_old(int j, int k, const Mat& newL2P) {  // T:3290-3314
    int kP = k / nChild, ck = k % nChild;
    Box& par = B(j - 1, kP);
    std::vector<Mat> U(nChild);
    int row = 0;
    for (int c = 0; c < nChild; ++c) {
      int nl = B(j, kP * nChild + c).NumLocals;
      U[c] = par.L2P.block(row, 0, nl, par.L2P.c);
      row += nl;
    }
    U[ck] = gemm(gemm(adjoint(newL2P), B(j, k).L2P), U[ck]);
    int nr = 0;
    for (auto& u : U) nr += u.r;
    Mat P(nr, U[0].c);
    row = 0;
    for (auto& u : U) { P.setBlock(row, 0, u); row += u.r; }
    par.L2P = P;
  }
  void update_parent_P2M_old(int j, int b, const Mat& newP2Mt) {  // T:3316-3340
    int bP = b / nChild, cb = b % nChild;
    Box& par = B(j - 1, bP);
    std::vector<Mat> V(nChild);
    int col = 0;
    for (int c = 0; c < nChild; ++c) {
      int nm = B(j, bP * nChild + c).NumMultipoles;
      V[c] = par.P2M.block(0, col, par.P2M.r, nm);
      col += nm;
    }
    V[cb] = gemm(gemm(V[cb], B(j, b).P2M), newP2Mt);
    int nc = 0;
    for (auto& v : V) nc += v.c;
    Mat P(V[0].r, nc);
    col = 0;
    for (auto& v : V) { P.setBlock(0, col, v); col += v.c; }
    par.P2M = P;
  }
  static Mat hcat(const Mat& A, const Mat& Bm) {
    if (A.r != Bm.r) throw std::runtime_error("oracle: hcat row mismatch");
    Mat C(A.r, A.c + Bm.c);
    C.setBlock(0, 0, A);
    C.setBlock(0, A.c, Bm);
    return C;
  }
  // T:1573-1656
  void compress_P2P_myRRQR(int j, int k, int b, const Mat& oldL2P, const Mat& oldP2M, const Mat& F,
                           Mat& newL2P, Mat& newP2Mt, Mat& M2Lkb) {
    Mat Acol = hcat(oldL2P, F);
    if (std::getenv("RCS_DEBUG_LOG"))
      std::fprintf(stderr, "P2P j%d k%d b%d A %dx%d part %d ip %.6e nF %.6e\n", j, k, b, Acol.r, Acol.c, oldL2P.c,
                   B(j, k).maxPivot_L2P, F.norm());
    RRQRResult r1 = rrqr(Acol, TOL_POW, fill_in_TOL_POW, oldL2P.c, B(j, k).maxPivot_L2P, fillin_rankVec[j], &ctr);
    if (std::getenv("RCS_DEBUG_LOG")) std::fprintf(stderr, "  r1 %d mp %.6e\n", r1.rank, r1.maxPivot);
    B(j, k).maxPivot_L2P = std::max(B(j, k).maxPivot_L2P, r1.maxPivot);
    newL2P = r1.Q;
    Mat Rk = gemm(adjoint(newL2P), oldL2P);
    Mat Rkp = gemm(adjoint(newL2P), F);
    Mat Arow = hcat(adjoint(oldP2M), adjoint(Rkp));
    RRQRResult r2 = rrqr(Arow, TOL_POW, fill_in_TOL_POW, oldP2M.r, B(j, b).maxPivot_P2M, fillin_rankVec[j], &ctr);
    if (std::getenv("RCS_DEBUG_LOG"))
      std::fprintf(stderr, "  r2 %d mp %.6e ip %.6e A %dx%d\n", r2.rank, r2.maxPivot, B(j, b).maxPivot_P2M, Arow.r, Arow.c);
    B(j, b).maxPivot_P2M = std::max(B(j, b).maxPivot_P2M, r2.maxPivot);
    newP2Mt = r2.Q;
    Mat Rn = gemm(adjoint(newP2Mt), adjoint(oldP2M));
    Mat Rnp = gemm(adjoint(newP2Mt), adjoint(Rkp));
    Mat& M = B(j, k).M2L[b];
    if (M.size() > 0) M2Lkb = add(gemm(gemm(Rk, M), adjoint(Rn)), adjoint(Rnp));
    else M2Lkb = adjoint(Rnp);
  }
  static void squareUp(const Mat& L2P, const Mat& P2Mt, Mat& L2Pm, Mat& P2Mtm) {  // T:2168-2183
    if (L2P.c < P2Mt.c) {
      P2Mtm = P2Mt.block(0, 0, P2Mt.r, L2P.c);
      L2Pm = L2P;
    } else {
      L2Pm = L2P.block(0, 0, L2P.r, P2Mt.c);
      P2Mtm = P2Mt;
    }
  }
  void compress_P2P_Two_Ways(int j, int k, int b) {  // T:2116-2306
    if (B(j, b).M2L[k].size() == 0 || B(j, k).M2L[b].size() == 0) return;
    ctr.n_compress_p2p++;
    Mat newL2Pk, newP2Mtb, M2Lkb;
    compress_P2P_myRRQR(j, k, b, B(j, k).L2P, B(j, b).P2M, B(j, k).P2P[b], newL2Pk, newP2Mtb, M2Lkb);
    Mat newL2Pb, newP2Mtk, M2Lbk;
    compress_P2P_myRRQR(j, b, k, B(j, b).L2P, B(j, k).P2M, B(j, b).P2P[k], newL2Pb, newP2Mtk, M2Lbk);
    Mat L2Pk_m, P2Mtk_m, L2Pb_m, P2Mtb_m;
    squareUp(newL2Pk, newP2Mtk, L2Pk_m, P2Mtk_m);
    squareUp(newL2Pb, newP2Mtb, L2Pb_m, P2Mtb_m);
    B(j, k).M2L[b] = M2Lkb.block(0, 0, L2Pk_m.c, P2Mtb_m.c);
    update_other_M2Ls_at_k_old(j, k, b, L2Pk_m);
    update_other_M2Ls_at_b_old(j, k, b, P2Mtb_m);
    if (j > startLevel) {
      update_parent_L2P_old(j, k, L2Pk_m);
      update_parent_P2M_old(j, b, P2Mtb_m);
    }
    B(j, k).L2P = L2Pk_m;
    B(j, b).P2M = adjoint(P2Mtb_m);
    B(j, k).NumLocals = B(j, k).L2P.c;
    B(j, b).NumMultipoles = B(j, b).P2M.r;
    B(j, b).M2L[k] = M2Lbk.block(0, 0, L2Pb_m.c, P2Mtk_m.c);
    update_other_M2Ls_at_k_old(j, b, k, L2Pb_m);
    update_other_M2Ls_at_b_old(j, b, k, P2Mtk_m);
    if (j > startLevel) {
      update_parent_L2P_old(j, b, L2Pb_m);
      update_parent_P2M_old(j, k, P2Mtk_m);
    }
    B(j, b).L2P = L2Pb_m;
    B(j, k).P2M = adjoint(P2Mtk_m);
    B(j, b).NumLocals = B(j, b).L2P.c;
    B(j, k).NumMultipoles = B(j, k).P2M.r;
  }
  void compress_M2P_P2L(int j, int k, int b) {  // T:2454-2596, 2878-2959
    if (B(j, b).M2L[k].size() == 0 || B(j, k).M2L[b].size() == 0) return;
    ctr.n_compress_p2l++;
    Box& bk = B(j, k);
    Box& bb = B(j, b);
    // compress_M2P_myRRQR(j, k, b)
    Mat Acol = hcat(bk.L2P, bk.M2P[b]);
    RRQRResult r1 = rrqr(Acol, TOL_POW, fill_in_TOL_POW, bk.L2P.c, bk.maxPivot_L2P, fillin_rankVec[j], &ctr);
    Mat newL2Pk = r1.Q;
    bk.maxPivot_L2P = std::max(bk.maxPivot_L2P, r1.maxPivot);
    Mat Rk = gemm(adjoint(newL2Pk), bk.L2P);
    Mat Rkp = gemm(adjoint(newL2Pk), bk.M2P[b]);
    Mat M2Lkb = add(gemm(Rk, bk.M2L[b]), Rkp);
    // compress_P2L_myRRQR(j, b, k)
    Mat Arow = hcat(adjoint(bk.P2M), adjoint(bb.P2L[k]));
    RRQRResult r2 = rrqr(Arow, TOL_POW, fill_in_TOL_POW, bk.P2M.r, bk.maxPivot_P2M, fillin_rankVec[j], &ctr);
    Mat newP2Mtk = r2.Q;
    bk.maxPivot_P2M = std::max(bk.maxPivot_P2M, r2.maxPivot);
    Mat Rn = gemm(adjoint(newP2Mtk), adjoint(bk.P2M));
    Mat Rnp = gemm(adjoint(newP2Mtk), adjoint(bb.P2L[k]));
    Mat M2Lbk = add(gemm(bb.M2L[k], adjoint(Rn)), adjoint(Rnp));
    Mat L2Pm, P2Mtm;
    squareUp(newL2Pk, newP2Mtk, L2Pm, P2Mtm);
    bk.M2L[b] = M2Lkb.block(0, 0, L2Pm.c, M2Lkb.c);
    update_other_M2Ls_at_k_old(j, k, b, L2Pm);
    if (j > startLevel) update_parent_L2P_old(j, k, L2Pm);
    bk.L2P = L2Pm;
    bk.NumLocals = bk.L2P.c;
    bb.M2L[k] = M2Lbk.block(0, 0, M2Lbk.r, P2Mtm.c);
    update_other_M2Ls_at_b_old(j, b, k, P2Mtm);
    if (j > startLevel) update_parent_P2M_old(j, k, P2Mtm);
    bk.P2M = adjoint(P2Mtm);
    bk.NumMultipoles = bk.P2M.r;
  }

  void eliminate_phase_efficient() {  // T:4599-4747
    for (int j = nLevels; j >= startLevel; --j) {
      if (j != nLevels) {
        initialise_P2P_NonLeafLevel(j);
        deleteM2L(j);
      }
      for (int k = 0; k < nBoxesPerLevel[j]; ++k) {
        for (size_t i = 0; i < P2Plist.size(); ++i) {
          int nn = P2Plist[i].first, pp = P2Plist[i].second;
          if (k == nn || k == pp) {
            P2Plist.erase(P2Plist.begin() + i);
            --i;
            if (B(j, nn).P2P[pp].size() != 0) {
              compress_P2P_Two_Ways(j, nn, pp);
              B(j, nn).P2P[pp] = Mat();
              B(j, pp).P2P[nn] = Mat();
            }
          }
        }
        for (size_t i = 0; i < P2L_M2P.size(); ++i) {
          int nn = P2L_M2P[i].first, pp = P2L_M2P[i].second;
          if (k == nn || k == pp) {
            P2L_M2P.erase(P2L_M2P.begin() + i);
            --i;
            if (B(j, nn).P2L[pp].size() != 0) {
              compress_M2P_P2L(j, pp, nn);
              B(j, nn).P2L[pp] = Mat();
              B(j, pp).M2P[nn] = Mat();
            }
          }
        }
        B(j, k).Eliminated = true;
        eliminate_x(j, k);
        eliminate_z(j, k);
      }
    }
  }

  // ---------------- solve (T:670-678, 3506-3535, 3613-3653, 4115-4148, 4749-4907) ----------------
  void assign_Leaf_rhs(const std::vector<cd>& b) {
    for (int k = 0; k < nBoxesPerLevel[nLevels]; ++k) {
      Box& bx = B(nLevels, k);
      bx.particle_rhs = colvec((int)bx.chargeLocations.size());
      for (size_t i = 0; i < bx.chargeLocations.size(); ++i) bx.particle_rhs(i, 0) = b[bx.chargeLocations[i]];
    }
  }
  void rhs_eliminate_phase_efficient() {
    for (int j = nLevels; j >= startLevel; --j)
      for (int k = 0; k < nBoxesPerLevel[j]; ++k) B(j, k).Eliminated = false;
    for (int j = nLevels; j >= startLevel; --j) {
      if (j != nLevels) {  // T:3506-3523
        for (int k = 0; k < nBoxesPerLevel[j]; ++k) {
          int len = 0;
          for (int c = 0; c < nChild; ++c) len += B(j + 1, nChild * k + c).multipole_rhs.r;
          Mat pr(len, 1);
          int off = 0;
          for (int c = 0; c < nChild; ++c) {
            Mat& m = B(j + 1, nChild * k + c).multipole_rhs;
            pr.setBlock(off, 0, m);
            off += m.r;
          }
          B(j, k).particle_rhs = pr;
        }
      }
      for (int k = 0; k < nBoxesPerLevel[j]; ++k) {
        Box& bk = B(j, k);
        bk.Eliminated = true;
        if (bk.P2P[k].size() != 0) {  // rhs_eliminate_x
          Mat s = solveLU(bk.luP2P, bk.particle_rhs);
          bk.local_rhs = neg(gemm(bk.P2M, s));
          for (int nn : bk.neighborList) {
            if (nn == k) continue;
            Box& bn = B(j, nn);
            if (!bn.Eliminated) bn.particle_rhs = sub(bn.particle_rhs, gemm(bn.P2P[k], s));
            else bn.multipole_rhs = sub(bn.multipole_rhs, gemm(bn.P2L[k], s));
          }
        } else {
          bk.local_rhs = colvec(bk.P2M.r);
        }
        if (bk.L2M_f.size() > 0) {  // rhs_eliminate_z
          Mat s = solveLU(bk.luL2M, bk.local_rhs);
          bk.multipole_rhs = s;
          for (int nn : bk.neighborList) {
            if (nn == k) continue;
            Box& bn = B(j, nn);
            if (!bn.Eliminated) bn.particle_rhs = sub(bn.particle_rhs, gemm(bn.L2P_f[k], s));
            else bn.multipole_rhs = sub(bn.multipole_rhs, gemm(bn.L2P_f[k], s));
          }
        } else {
          bk.multipole_rhs = Mat(0, 1);
        }
      }
    }
  }
  void solve_particles_at_base_level(int j) {  // T:4772-4810
    int NM = 0, NL = 0;
    for (int k = 0; k < nBoxesPerLevel[j]; ++k) { NM += B(j, k).NumMultipoles; NL += B(j, k).NumLocals; }
    if (NL == 0 || NM == 0) return;
    Mat A(NL, NM), b(NL, 1);
    int ri = 0;
    for (int r = 0; r < nBoxesPerLevel[j]; ++r) {
      int si = 0;
      for (int s = 0; s < nBoxesPerLevel[j]; ++s) {
        auto it = B(j, r).M2L.find(s);
        if (it != B(j, r).M2L.end() && it->second.r == B(j, r).NumLocals && it->second.c == B(j, s).NumMultipoles)
          A.setBlock(ri, si, it->second);
        si += B(j, s).NumMultipoles;
      }
      b.setBlock(ri, 0, B(j, r).multipole_rhs);
      ri += B(j, r).NumLocals;
    }
    LU lu(A);
    ctr.getrf_flops += 8.0 / 3.0 * NL * (double)NL * NL;
    Mat x = solveLU(lu, b);
    int idx = 0;
    for (int k = 0; k < nBoxesPerLevel[j]; ++k) {
      B(j, k).multipoles = x.block(idx, 0, B(j, k).NumMultipoles, 1);
      idx += B(j, k).NumMultipoles;
    }
  }
  void back_substitution_phase() {  // T:4812-4865
    solve_particles_at_base_level(startLevel);
    for (int j = startLevel; j <= nLevels; ++j) {
      for (int k = nBoxesPerLevel[j] - 1; k >= 0; --k) {
        Box& bk = B(j, k);
        if (bk.P2P[k].size() != 0) {
          Mat rhs1 = bk.particle_rhs;
          for (int nn : bk.neighborList) {
            if (nn == k) continue;
            Box& bn = B(j, nn);
            if (!bn.Eliminated) rhs1 = sub(rhs1, gemm(bk.P2P[nn], bn.particles));
            else rhs1 = sub(rhs1, gemm(bk.M2P[nn], bn.multipoles));
          }
          Mat rhs2 = sub(bk.multipoles, gemm(bk.P2M, solveLU(bk.luP2P, rhs1)));
          if (bk.L2M_f.size() != 0) bk.locals = solveLU(bk.luL2M, rhs2);  // Atemp == L2M_f
          else bk.locals = Mat(0, 1);
          bk.particles = solveLU(bk.luP2P, sub(rhs1, gemm(bk.L2P, bk.locals)));
        } else {
          bk.particles = Mat(0, 1);
        }
        bk.Eliminated = false;
        if (nLevels != j) {
          int idx = 0;
          for (int c = 0; c < nChild; ++c) {
            Box& ch = B(j + 1, nChild * k + c);
            ch.multipoles = bk.particles.block(idx, 0, ch.NumMultipoles, 1);
            idx += ch.NumMultipoles;
          }
        }
      }
    }
  }
  // T:4874-4907 (box order) and T:466-476 (back to the input order)
  std::vector<cd> getx() {
    std::vector<cd> x(N, cd(0));
    int off = 0;
    for (int k = 0; k < nBoxesPerLevel[nLevels]; ++k) {
      Box& b = B(nLevels, k);
      for (int i = 0; i < b.particles.r; ++i) x[off + i] = b.particles(i, 0);
      off += (int)b.chargeLocations.size();
    }
    return x;
  }
  void reorder(std::vector<cd>& p) {
    std::vector<cd> t = p;
    int start = 0;
    for (int k = 0; k < nBoxesPerLevel[nLevels]; ++k)
      for (int idx : B(nLevels, k).chargeLocations) p[idx] = t[start++];
  }
};

// AIFMM<K> (src/AIFMM.cpp:6-147)
struct AIFMM {
  Tree* A;
  AIFMM(const SolverInputs& in, const Kernel* K) {
    A = new Tree(in, K);
    A->createTree();
    A->assign_Center_Location();
    A->assign_Tree_Interactions();
    A->assignChargeLocations();
    A->assignNonLeafChargeLocations();
    A->getNodes();
    A->assemble_M2L();
    A->free_charge_and_checkPoints();
    A->initialise_phase();
    A->initialise_P2P_Leaf_Level();
  }
  ~AIFMM() { delete A; }
  void factorize() { A->eliminate_phase_efficient(); }
  void backSubstitute1(const std::vector<cd>& rhs) { A->assign_Leaf_rhs(rhs); }
  void backSubstitute2() { A->rhs_eliminate_phase_efficient(); }
  void backSubstitute3() { A->back_substitution_phase(); }
  std::vector<cd> solve(const std::vector<cd>& rhs) {  // no reorder (src/AIFMM.cpp:103-110)
    A->assign_Leaf_rhs(rhs);
    A->rhs_eliminate_phase_efficient();
    A->back_substitution_phase();
    return A->getx();
  }
  std::vector<cd> getPhi() {
    std::vector<cd> p = A->getx();
    A->reorder(p);
    return p;
  }
};

}  // namespace oracle
