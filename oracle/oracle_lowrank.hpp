// ORACLE — test infrastructure only (see oracle_core.hpp).
//
// Adaptive cross approximation (src/ACA.cpp:25-38, 638-859) and the two-partition
// column-pivoted Householder QR (src/rrqr.cpp:46-182), restated.
#pragma once
#include <cfloat>
#include "oracle_kernels.hpp"

namespace oracle {

// Eigen's a.dot(b) conjugates its first argument.
inline cd dotc(const std::vector<cd>& a, const std::vector<cd>& b) {
  cd s = 0;
  for (size_t i = 0; i < a.size(); ++i) s += std::conj(a[i]) * b[i];
  return s;
}
inline double vnorm(const std::vector<cd>& a) {
  double s = 0;
  for (auto& v : a) s += std::norm(v);
  return std::sqrt(s);
}

// A:25-38: first index (in ascending order) of maximal |v| over the allowed set.
inline int maxAbsIndex(const std::vector<cd>& v, const std::set<int>& allowed) {
  if (allowed.empty()) return -1;
  int index = *allowed.begin();
  double best = std::abs(v[index]);
  for (int i : allowed) {
    double a = std::abs(v[i]);
    if (a > best) { best = a; index = i; }
  }
  return index;
}

struct ACAResult {
  std::vector<int> row_bases, col_bases;
  int rank = 0;
  Mat Ac, Ar;  // Ac = K(rows, pivot cols) (N1 x r); Ar = K(pivot rows, cols) (r x N2)
};

// LowRank<K>::ACA_only_nodes2, A:639-859. The stopping test compares |u||v| with
// tol*|S|_F where |S|_F is updated as at A:750-755 (the last term of the sum includes
// the newest pair itself).
inline ACAResult aca_only_nodes2(const Kernel& K, double tol, const std::vector<int>& rows,
                                 const std::vector<int>& cols) {
  ACAResult res;
  const int N1 = (int)rows.size(), N2 = (int)cols.size();
  const int mn = std::min(N1, N2);
  std::set<int> rr, rc;
  for (int l = 0; l < N1; ++l) rr.insert(l);
  for (int l = 0; l < N2; ++l) rc.insert(l);
  std::vector<std::vector<cd>> U, V, AcV, ArV;
  int rank = 0;
  if (N1 < N2) {
    std::vector<cd> row, col, u, v;
    int ri = 0, ci = 0;
    int l;
    for (l = 0; l < N1; ++l) {
      ri = l;
      row = K.getRow(rows[ri], cols);
      ci = maxAbsIndex(row, rc);
      if (std::abs(row[ci]) > 1e-16) break;
    }
    if (l == N1) {
      res.Ac = Mat(N1, 0);
      res.Ar = Mat(0, N2);
      return res;
    }
    v.resize(N2);
    cd piv = row[ci];
    for (int s = 0; s < N2; ++s) v[s] = row[s] / piv;
    res.row_bases.push_back(ri);
    res.col_bases.push_back(ci);
    col = K.getCol(cols[ci], rows);
    u = col;
    U.push_back(u);
    V.push_back(v);
    AcV.push_back(col);
    ArV.push_back(row);
    rc.erase(ci);
    rr.erase(ri);
    rank = 1;
    double normS = 0.0;
    ri = maxAbsIndex(col, rr);
    while (vnorm(u) * vnorm(v) > tol * normS && rank < mn) {
      res.row_bases.push_back(ri);
      row = K.getRow(rows[ri], cols);
      std::vector<cd> row_temp = row;
      for (int q = 0; q < rank; ++q)
        for (int s = 0; s < N2; ++s) row[s] -= U[q][ri] * V[q][s];
      ci = maxAbsIndex(row, rc);
      res.col_bases.push_back(ci);
      piv = row[ci];
      for (int s = 0; s < N2; ++s) v[s] = row[s] / piv;
      col = K.getCol(cols[ci], rows);
      std::vector<cd> col_temp = col;
      for (int q = 0; q < rank; ++q)
        for (int s = 0; s < N1; ++s) col[s] -= V[q][ci] * U[q][s];
      u = col;
      U.push_back(u);
      V.push_back(v);
      AcV.push_back(col_temp);
      ArV.push_back(row_temp);
      ++rank;
      rc.erase(ci);
      rr.erase(ri);
      double nu = vnorm(u), nv = vnorm(v);
      normS = normS * normS + (nu * nv) * (nu * nv);
      for (int q = 0; q < rank; ++q) normS += 2 * std::abs(dotc(U[q], u) * dotc(V[q], v));
      normS = std::sqrt(normS);
      if (rr.empty()) break;  // the reference would dereference an empty set here
      ri = maxAbsIndex(col, rr);
    }
  } else {
    std::vector<cd> row, col, u, v;
    int ri = 0, ci = 0;
    int l;
    for (l = 0; l < N2; ++l) {
      ci = l;
      col = K.getCol(cols[ci], rows);
      ri = maxAbsIndex(col, rr);
      if (std::abs(col[ri]) > 1e-16) break;
    }
    if (l == N2) {
      res.Ac = Mat(N1, 0);
      res.Ar = Mat(0, N2);
      return res;
    }
    res.col_bases.push_back(ci);
    v.resize(N1);
    cd piv = col[ri];
    for (int s = 0; s < N1; ++s) v[s] = col[s] / piv;
    res.row_bases.push_back(ri);
    row = K.getRow(rows[ri], cols);
    u = row;
    U.push_back(u);
    V.push_back(v);
    AcV.push_back(col);
    ArV.push_back(row);
    rank = 1;
    rr.erase(ri);
    rc.erase(ci);
    double normS = 0.0;
    ci = maxAbsIndex(row, rc);
    while (vnorm(u) * vnorm(v) > tol * normS && rank < mn) {
      res.col_bases.push_back(ci);
      col = K.getCol(cols[ci], rows);
      std::vector<cd> col_temp = col;
      for (int q = 0; q < rank; ++q)
        for (int s = 0; s < N1; ++s) col[s] -= U[q][ci] * V[q][s];
      ri = maxAbsIndex(col, rr);
      res.row_bases.push_back(ri);
      piv = col[ri];
      for (int s = 0; s < N1; ++s) v[s] = col[s] / piv;
      row = K.getRow(rows[ri], cols);
      std::vector<cd> row_temp = row;
      for (int q = 0; q < rank; ++q)
        for (int s = 0; s < N2; ++s) row[s] -= V[q][ri] * U[q][s];
      u = row;
      U.push_back(u);
      V.push_back(v);
      AcV.push_back(col_temp);
      ArV.push_back(row_temp);
      ++rank;
      rr.erase(ri);
      rc.erase(ci);
      double nu = vnorm(u), nv = vnorm(v);
      normS = normS * normS + (nu * nv) * (nu * nv);
      for (int q = 0; q < rank; ++q) normS += 2 * std::abs(dotc(U[q], u) * dotc(V[q], v));
      normS = std::sqrt(normS);
      if (rc.empty()) break;
      ci = maxAbsIndex(row, rc);
    }
  }
  res.rank = rank;
  res.Ac = Mat(N1, rank);
  res.Ar = Mat(rank, N2);
  for (int q = 0; q < rank; ++q) {
    for (int i = 0; i < N1; ++i) res.Ac(i, q) = AcV[q][i];
    for (int i = 0; i < N2; ++i) res.Ar(q, i) = ArV[q][i];
  }
  return res;
}

// RRQR::colPivHouseholderQRWithPartition, Q:69-182. Columns [0, partition) are the
// old basis; they are pivoted among themselves and stop at pivot/max1 < tol1 with
// rank j-1 (Q:115-117). The remaining columns are then pivoted over everything left
// and stop at pivot/max(inputPivot, first pivot) < tol2 or after fillInRank columns.
// Column norms are downdated (Q:124-127). Q_trunc is formed from the reflectors.
struct RRQRResult {
  Mat Q;  // m x rank
  Mat R;  // rank x n (columns in pivoted order)
  std::vector<int> perm;  // column i of A*P is column perm[i] of A
  int rank = 0;
  int rank1 = 0, rank2 = 0;
  double maxPivot = 0.0;
};

inline void householder(const Mat& R, int col, int i0, std::vector<cd>& v) {  // Q:46-67
  const int n = R.r;
  v.assign(n, cd(0));
  double sn = 0;
  for (int i = i0; i < n; ++i) { v[i] = R(i, col); sn += std::norm(v[i]); }
  sn = std::sqrt(sn);
  if (i0 < n) v[i0] = v[i0] + std::exp(I1 * std::arg(R(i0, col))) * sn;
  double vn = 0;
  for (auto& x : v) vn += std::norm(x);
  vn = std::sqrt(vn);
  if (vn > 0) {
    double s = std::sqrt(2.0) / vn;
    for (auto& x : v) x *= s;
  }
}

inline RRQRResult rrqr(const Mat& A, int tolPow1, int tolPow2, int partition, double inputPivot,
                       int fillInRank, Counters* ctr = nullptr) {
  RRQRResult res;
  const int m = A.r, n = A.c;
  const double tol1 = std::pow(10.0, -tolPow1), tol2 = std::pow(10.0, -tolPow2);
  int steps = std::min(n, m);
  Mat R = A;
  std::vector<int> P(n);
  for (int i = 0; i < n; ++i) P[i] = i;
  std::vector<double> ns(n);
  for (int j = 0; j < n; ++j) {
    double s = 0;
    for (int i = 0; i < m; ++i) s += std::norm(R(i, j));
    ns[j] = s;
  }
  std::vector<std::vector<cd>> refl;
  double maxPivot = 0.0, max1 = 0.0;
  int rank1 = std::min(partition, m);

  auto pivot_sort = [&](int j, int end) {
    // stable order by descending norm over [j, end): the reference uses std::sort
    std::vector<int> idx(end - j);
    for (int t = 0; t < end - j; ++t) idx[t] = j + t;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return ns[a] > ns[b]; });
    Mat Rn = R;
    std::vector<double> nsn = ns;
    std::vector<int> Pn = P;
    for (int t = 0; t < end - j; ++t) {
      int src = idx[t];
      for (int i = 0; i < m; ++i) Rn(i, j + t) = R(i, src);
      nsn[j + t] = ns[src];
      Pn[j + t] = P[src];
    }
    R.a.swap(Rn.a);
    ns.swap(nsn);
    P.swap(Pn);
  };
  auto apply = [&](const std::vector<cd>& v, int j) {
    // R <- R - v (v^H R) over all columns; returns |R(j,j)| afterwards via R
    for (int c = 0; c < n; ++c) {
      cd w = 0;
      for (int i = 0; i < m; ++i) w += std::conj(v[i]) * R(i, c);
      if (w == cd(0)) continue;
      for (int i = 0; i < m; ++i) R(i, c) -= v[i] * w;
    }
  };
  for (int j = 0; j < rank1; ++j) {
    pivot_sort(j, partition);
    std::vector<cd> v;
    householder(R, j, j, v);
    Mat Rt = R;
    apply(v, j);
    if (j == 0) max1 = std::abs(R(j, j));
    maxPivot = max1;
    double pivot = std::abs(R(j, j));
    if (pivot / max1 < tol1) {
      std::swap(R, Rt);  // the reference discards Rtemp on exit
      rank1 = j - 1;
      break;
    }
    for (int i = j + 1; i < m; ++i) R(i, j) = 0.0;
    refl.push_back(v);
    for (int i = j + 1; i < n; ++i) ns[i] = ns[i] - std::abs(R(j, i)) * std::abs(R(j, i));
  }
  if (rank1 < 0) rank1 = 0;  // unreachable in exact arithmetic (pivot 0 equals max1)
  for (int i = rank1; i < partition && i < n; ++i) ns[i] = 0.0;
  steps = steps - (partition - rank1);
  const int first = rank1;
  int rank2 = std::min(steps - first, m);
  if (rank2 < 0) rank2 = 0;
  // After an early stop at step j the reflectors 0..j-1 stay in Q although rank1 is
  // j-1; the second partition restarts at position rank1 on top of them (Q:115-134).
  std::vector<std::vector<cd>> refl2;
  for (int j = first; j < steps; ++j) {
    pivot_sort(j, n);
    std::vector<cd> v;
    householder(R, j, j, v);
    Mat Rt = R;
    apply(v, j);
    if (j == first) maxPivot = std::max(inputPivot, std::abs(R(first, first)));
    double pivot = std::abs(R(j, j));
    if (pivot / maxPivot < tol2 || j - first >= fillInRank) {
      std::swap(R, Rt);
      rank2 = j - first;
      break;
    }
    for (int i = j + 1; i < m; ++i) R(i, j) = 0.0;
    refl2.push_back(v);
    for (int i = j + 1; i < n; ++i) ns[i] = ns[i] - std::abs(R(j, i)) * std::abs(R(j, i));
  }
  for (auto& v : refl2) refl.push_back(v);
  const int rankA = rank1 + rank2;
  // Q = H_1 H_2 ... H_s; Q_trunc = Q(:, 0:rankA)
  Mat Q(m, rankA);
  for (int c = 0; c < rankA; ++c) Q(c, c) = 1.0;
  for (int s = (int)refl.size() - 1; s >= 0; --s) {
    const std::vector<cd>& v = refl[s];
    for (int c = 0; c < rankA; ++c) {
      cd w = 0;
      for (int i = 0; i < m; ++i) w += std::conj(v[i]) * Q(i, c);
      if (w == cd(0)) continue;
      for (int i = 0; i < m; ++i) Q(i, c) -= v[i] * w;
    }
  }
  res.Q = Q;
  res.R = R.block(0, 0, rankA, n);
  res.perm = P;
  res.rank = rankA;
  res.rank1 = rank1;
  res.rank2 = rank2;
  res.maxPivot = maxPivot;
  if (ctr) {
    ctr->n_rrqr++;
    double r = rankA;
    ctr->qr_flops += 8.0 * (2.0 * m * n * r - r * r * (m + n) + 2.0 * r * r * r / 3.0) +
                     8.0 * (2.0 * m * r * r - 2.0 * r * r * r / 3.0);
  }
  return res;
}

}  // namespace oracle
