/*
 * C++ host interface of the MI355X AIFMM library, mirroring the reference's classes
 * (vaishna77/RCS) so that a driver written against them keeps its shape:
 *
 *   kernel            include/kernel.hpp:13-47, src/kernel.cpp:11-73
 *   EFIEMatrix        include/EFIEMatrix.hpp:29-118 (mesh files, N, getRhs, RCS)
 *   ElectrostaticsMatrix include/ElectrostaticsMatrix.hpp:8-40
 *   LowRank<K>        include/ACA.hpp:7-30 (ACA_only_nodes2)
 *   RRQR              include/rrqr.hpp:5-27
 *   AIFMM<K>          include/AIFMM.hpp:11-36 (factorize, solve, backSubstitute, backSubstitute1/2/3, getPhi,
 *                     and A for the tree statistics the driver prints)
 *
 * Header only, over the C ABI of rcs_amd.h (link with -lrcs_amd). Matrices are
 * column-major complex (rcs::Mat); the reference's Eigen Mat/Vec map onto data()/rows()/
 * cols() one to one. Differences from the reference, on purpose: failures throw
 * std::runtime_error with the library's message (the reference has no error path);
 * freq is a constructor argument (the reference compiles it in, include/includes.hpp:58);
 * single entries (getMatrixEntry) are served by the GPU one at a time and are meant
 * for checks, the batched getMatrix/getRow/getCol for real use.
 */
#ifndef RCS_AMD_HPP
#define RCS_AMD_HPP

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <fstream>
#include <map>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "rcs_amd.h"

namespace rcs {

using cdouble = std::complex<double>;
using Vec = std::vector<cdouble>;

// Column-major complex matrix (the role of Eigen::MatrixXcd in the reference).
struct Mat {
  int rows_ = 0, cols_ = 0;
  std::vector<cdouble> v;
  Mat() = default;
  Mat(int r, int c) : rows_(r), cols_(c), v((size_t)r * c) {}
  int rows() const { return rows_; }
  int cols() const { return cols_; }
  cdouble& operator()(int i, int j) { return v[(size_t)i + (size_t)j * rows_]; }
  const cdouble& operator()(int i, int j) const { return v[(size_t)i + (size_t)j * rows_]; }
  double* data() { return reinterpret_cast<double*>(v.data()); }
  const double* data() const { return reinterpret_cast<const double*>(v.data()); }
};

inline void check(int rc) {
  if (rc != 0) throw std::runtime_error(std::string("rcs_amd: ") + rcs_last_error());
}

// Node (include/includes.hpp:20, an Eigen::Vector3d there): x, y, z with indexing and the
// vector arithmetic the reference driver uses on nodes.
struct Node {
  double x = 0, y = 0, z = 0;
  Node() = default;
  Node(double a, double b, double c) : x(a), y(b), z(c) {}
  double& operator()(int i) { return i == 0 ? x : (i == 1 ? y : z); }
  double operator()(int i) const { return i == 0 ? x : (i == 1 ? y : z); }
  double& operator[](int i) { return (*this)(i); }
  double operator[](int i) const { return (*this)(i); }
  Node operator+(const Node& o) const { return {x + o.x, y + o.y, z + o.z}; }
  Node operator-(const Node& o) const { return {x - o.x, y - o.y, z - o.z}; }
  Node operator*(double a) const { return {x * a, y * a, z * a}; }
  Node operator/(double a) const { return {x / a, y / a, z / a}; }
  double dot(const Node& o) const { return x * o.x + y * o.y + z * o.z; }
  double norm() const { return std::sqrt(dot(*this)); }
};

// Patch (include/Patch.hpp): 1-based vertex indices, conductor index, centroid, area and
// largestEdge as the reference computes them.
struct Patch {
  int indexOfVertex1 = 0, indexOfVertex2 = 0, indexOfVertex3 = 0, conductorIndex = 0;
  Node centroid;
  double area = 0, largestEdge = 0;
};

// edgeProperties (include/EFIEMatrix.hpp:24-27): edge length and the two patches of an RWG edge.
struct edgeProperties {
  double length = 0;
  std::vector<int> indicesOfAssociatedPatches;
};

// include/includes.hpp:73-84
struct inputsToKernelClass {
  std::string nodeListFilename;
  std::string patchListFilename;
  int n = 0;
  double L = 0;
};
struct inputsToSolverClass {
  double xmin = 0, xmax = 0, ymin = 0, ymax = 0, zmin = 0, zmax = 0;
  int leafSize = 64, tol_pow = 6, nLevelsInput = 0, fillin_rank = 0, fillin_tol = 6;
  std::vector<double> locations;  // N x 3, row by row (the rows of the reference's locations matrix)
};

// Reference mesh files: "x y z" per node line, "i j k [conductor]" (1-based) per patch line
// (E:31-44, E:72-95). A line without the conductor column gives conductor 0, as the
// reference's stream extraction does.
inline void readMeshFiles(const std::string& nodes, const std::string& patches, std::vector<double>& xyz,
                          std::vector<int>& tris, std::vector<int>* conductor = nullptr) {
  std::ifstream fn(nodes), fp(patches);
  if (!fn || !fp) throw std::runtime_error("rcs_amd: cannot open mesh files");
  double a, b, c;
  while (fn >> a >> b >> c) { xyz.push_back(a); xyz.push_back(b); xyz.push_back(c); }
  std::string line;
  while (std::getline(fp, line)) {
    std::istringstream ls(line);
    int i, j, k, cond = 0;
    if (!(ls >> i >> j >> k)) continue;
    if (!(ls >> cond)) cond = 0;
    tris.push_back(i); tris.push_back(j); tris.push_back(k);
    if (conductor) conductor->push_back(cond);
  }
}

// kernel (include/kernel.hpp): batched entry generation on the GPU.
class kernel {
 public:
  int N = 0;
  virtual ~kernel() { if (mesh_) rcs_mesh_destroy(mesh_); }
  kernel(const kernel&) = delete;
  kernel& operator=(const kernel&) = delete;

  virtual cdouble getMatrixEntry(const unsigned i, const unsigned j) {
    int r = (int)i, c = (int)j;
    double out[2];
    check(rcs_entries(mesh_, RCS_PREC_C128, &r, 1, &c, 1, out));
    return {out[0], out[1]};
  }
  Mat getMatrix(std::vector<int> row_indices, std::vector<int> col_indices) {  // src/kernel.cpp:49-61
    Mat m((int)row_indices.size(), (int)col_indices.size());
    if (m.v.empty()) return m;
    check(rcs_entries(mesh_, RCS_PREC_C128, row_indices.data(), m.rows(), col_indices.data(), m.cols(), m.data()));
    return m;
  }
  Mat getMatrix(int row_start_index, int col_start_index, int row_end_index, int col_end_index) {
    std::vector<int> r, c;  // src/kernel.cpp:63-73 (absolute indices, valid from 0 as there)
    for (int i = row_start_index; i < row_end_index; ++i) r.push_back(i);
    for (int j = col_start_index; j < col_end_index; ++j) c.push_back(j);
    return getMatrix(r, c);
  }
  void getMatrix(Mat& mat, int n) { mat = getMatrix(0, 0, n, n); }
  Vec getRow(const int j, std::vector<int> col_indices) {  // src/kernel.cpp:11-19
    Mat m = getMatrix(std::vector<int>{j}, col_indices);
    return Vec(m.v.begin(), m.v.end());
  }
  Vec getCol(const int k, std::vector<int> row_indices) {  // src/kernel.cpp:21-29
    Mat m = getMatrix(row_indices, std::vector<int>{k});
    return Vec(m.v.begin(), m.v.end());
  }
  Vec getRow(const int n, const int k) {  // row k over all n columns
    std::vector<int> c(n);
    for (int i = 0; i < n; ++i) c[i] = i;
    return getRow(k, c);
  }
  Vec getCol(const int n, const int k) {  // column k over all n rows
    std::vector<int> r(n);
    for (int i = 0; i < n; ++i) r[i] = i;
    return getCol(k, r);
  }
  rcs_mesh_t handle() const { return mesh_; }
  double frequency() const { return freq_; }

  // Public data of EFIEMatrix / ElectrostaticsMatrix that drivers read
  // (include/EFIEMatrix.hpp:31-34, include/ElectrostaticsMatrix.hpp:10-14): nodeList is
  // 1-based with the dummy origin at index 0 (E:32-33).
  std::vector<Node> nodeList;
  std::vector<Patch> patchList;
  std::set<int> conductorIndices;
  int numConductors = 0;

 protected:
  kernel() = default;
  // hostOnly: host tables only (rcs_mesh_create_host), for post-processing without a GPU;
  // entries, ACA and AIFMM then report an error.
  void create(const std::vector<double>& xyz, const std::vector<int>& tris, int kind, double freq,
              const std::vector<int>* conductor = nullptr, bool hostOnly = false) {
    freq_ = freq;
    const int nt = (int)tris.size() / 3;
    const int* cond = (conductor && (int)conductor->size() == nt) ? conductor->data() : nullptr;
    check((hostOnly ? rcs_mesh_create_host : rcs_mesh_create)(xyz.data(), (int)xyz.size() / 3, tris.data(), nt, cond,
                                                               kind, freq, &mesh_));
    N = rcs_mesh_size(mesh_);
    nodeList.assign(1, Node());
    for (size_t i = 0; i + 2 < xyz.size(); i += 3) nodeList.push_back(Node(xyz[i], xyz[i + 1], xyz[i + 2]));
    std::vector<double> p5(5 * (size_t)nt);
    std::vector<int> pc(nt);
    check(rcs_mesh_patches(mesh_, p5.data(), pc.data()));
    patchList.resize(nt);
    conductorIndices.clear();
    for (int t = 0; t < nt; ++t) {
      Patch& p = patchList[t];
      p.indexOfVertex1 = tris[3 * t];
      p.indexOfVertex2 = tris[3 * t + 1];
      p.indexOfVertex3 = tris[3 * t + 2];
      p.conductorIndex = pc[t];
      p.centroid = Node(p5[5 * t], p5[5 * t + 1], p5[5 * t + 2]);
      p.area = p5[5 * t + 3];
      p.largestEdge = p5[5 * t + 4];
      conductorIndices.insert(pc[t]);
    }
    numConductors = (int)conductorIndices.size();
  }
  // Dense LU solve of the whole N x N system (getCoefficients E:831-837 and
  // getchargeDensityUsingLU src/ElectrostaticsMatrix.cpp:342-362): every entry on the GPU,
  // then the GPU LU; meant for small N, as in the reference.
  Vec denseSolve(const Vec& rhs) {
    Mat mat = getMatrix(0, 0, N, N);
    Vec x(rhs.size());
    const int nrhs = N > 0 ? (int)(rhs.size() / N) : 0;
    check(rcs_lu_solve(RCS_PREC_C128, mat.data(), N, reinterpret_cast<const double*>(rhs.data()), nrhs,
                       reinterpret_cast<double*>(x.data())));
    return x;
  }
  rcs_mesh_t mesh_ = nullptr;
  double freq_ = 0;
};

// EFIEMatrix (include/EFIEMatrix.hpp:29-118): RWG EFIE operator of a closed PEC mesh.
class EFIEMatrix : public kernel {
 public:
  EFIEMatrix(inputsToKernelClass in, double freq, bool hostOnly = false) {
    std::vector<double> xyz;
    std::vector<int> tris, cond;
    readMeshFiles(in.nodeListFilename, in.patchListFilename, xyz, tris, &cond);
    create(xyz, tris, RCS_KERNEL_EFIE, freq, &cond, hostOnly);
    buildEdgeMap();
  }
  EFIEMatrix(const std::vector<double>& xyz, const std::vector<int>& tris, double freq) {
    create(xyz, tris, RCS_KERNEL_EFIE, freq);
    buildEdgeMap();
  }
  // edge2PatchMap (E:127-209): RWG edges keyed by their sorted node pair; iterating it in
  // order visits the bases in the library's numbering.
  std::map<std::pair<int, int>, edgeProperties> edge2PatchMap;
  // getCoefficients (E:831-837): dense LU solution of the full system.
  void getCoefficients(Vec& rhs, Vec& coefficients) { coefficients = denseSolve(rhs); }
  // getJAtCentroids (E:1100-1130): n_patch x 3 surface current at the patch centroids.
  void getJAtCentroids(Vec& coefficients, Mat& JAtCentroids) {
    JAtCentroids = Mat((int)patchList.size(), 3);
    check(rcs_j_at_centroids(mesh_, reinterpret_cast<const double*>(coefficients.data()), JAtCentroids.data()));
  }
  // getRCS_OnePointQuadrature (E:916-930), m^2.
  double getRCS_OnePointQuadrature(double theta, double phi, Mat& JAtCentroids) {
    double out = 0;
    check(rcs_rcs_from_j(mesh_, JAtCentroids.data(), &theta, &phi, 1, &out));
    return out;
  }
  // getRhs(Node& Einc, VectorXcd& rhs) (E:785-829); kvec defaults to the reference's
  // hard-coded phase vector (0, 0, 1)
  void getRhs(Node& Einc, Vec& rhs, Node kvec = Node{0, 0, 1}) {
    rhs.assign(N, cdouble());
    const double E[3] = {Einc.x, Einc.y, Einc.z}, k[3] = {kvec.x, kvec.y, kvec.z};
    check(rcs_rhs(mesh_, E, k, reinterpret_cast<double*>(rhs.data())));
  }
  // getJAtCentroids + getRCS_OnePointQuadrature (E:916-930, E:1100-1130), m^2
  std::vector<double> getRCS(const Vec& coefficients, const std::vector<double>& theta,
                             const std::vector<double>& phi) {
    std::vector<double> out(theta.size());
    check(rcs_rcs(mesh_, reinterpret_cast<const double*>(coefficients.data()), theta.data(), phi.data(),
                  (int)theta.size(), out.data()));
    return out;
  }
  std::vector<double> locations() {
    std::vector<double> xyz(3 * (size_t)N);
    check(rcs_mesh_locations(mesh_, xyz.data()));
    return xyz;
  }

 private:
  void buildEdgeMap() {
    std::vector<int> t6(6 * (size_t)N);
    std::vector<double> len(N);
    check(rcs_mesh_rwg(mesh_, t6.data(), len.data()));
    edge2PatchMap.clear();
    for (int i = 0; i < N; ++i) {
      edgeProperties e;
      e.length = len[i];
      e.indicesOfAssociatedPatches = {t6[6 * i + 2], t6[6 * i + 3]};
      edge2PatchMap[std::make_pair(t6[6 * i], t6[6 * i + 1])] = e;
    }
  }
};

// ElectrostaticsMatrix (include/ElectrostaticsMatrix.hpp:8-40): piecewise-constant
// single layer C0 * int 1/R at patch centroids.
class ElectrostaticsMatrix : public kernel {
 public:
  explicit ElectrostaticsMatrix(inputsToKernelClass in, bool hostOnly = false) {
    std::vector<double> xyz;
    std::vector<int> tris, cond;
    readMeshFiles(in.nodeListFilename, in.patchListFilename, xyz, tris, &cond);
    create(xyz, tris, RCS_KERNEL_LAPLACE, 1.0e9, &cond, hostOnly);
  }
  // getRhs(conductorIndex, rhs) (src/ElectrostaticsMatrix.cpp:374-381): 1 on the patches of
  // that conductor, 0 elsewhere.
  void getRhs(int conductorIndex, std::vector<double>& rhs) {
    rhs.assign(N, 0.0);
    for (int i = 0; i < N; ++i)
      if (patchList[i].conductorIndex == conductorIndex) rhs[i] = 1.0;
  }
  // getchargeDensityUsingLU (src/ElectrostaticsMatrix.cpp:342-346): dense LU solution.
  void getchargeDensityUsingLU(std::vector<double>& rhs, std::vector<double>& chargeDensity) {
    Vec b(rhs.begin(), rhs.end());
    Vec x = denseSolve(b);
    chargeDensity.resize(x.size());
    for (size_t i = 0; i < x.size(); ++i) chargeDensity[i] = x[i].real();
  }
  // getCapacitance (src/ElectrostaticsMatrix.cpp:366-372): sum of charge density times area.
  double getCapacitance(std::vector<double>& chargeDensity) {
    double c = 0;
    for (int i = 0; i < N; ++i) c += chargeDensity[i] * patchList[i].area;
    return c;
  }
  // getCentroids (src/ElectrostaticsMatrix.cpp:82): N x 3, row by row.
  void getCentroids(std::vector<double>& centroids) {
    centroids.resize(3 * (size_t)N);
    for (int i = 0; i < N; ++i) {
      centroids[3 * i] = patchList[i].centroid.x;
      centroids[3 * i + 1] = patchList[i].centroid.y;
      centroids[3 * i + 2] = patchList[i].centroid.z;
    }
  }
  ElectrostaticsMatrix(const std::vector<double>& xyz, const std::vector<int>& tris) {
    create(xyz, tris, RCS_KERNEL_LAPLACE, 1.0e9);
  }
  std::vector<double> locations() {
    std::vector<double> xyz(3 * (size_t)N);
    check(rcs_mesh_locations(mesh_, xyz.data()));
    return xyz;
  }
};

// LowRank<K> (include/ACA.hpp): ACA_only_nodes2 (src/ACA.cpp:639-859). row_bases and
// col_bases are positions within row_indices / col_indices, as in the reference; Ac and
// Ar are the original kernel columns and rows at those pivots.
template <typename kerneltype>
class LowRank {
 public:
  kerneltype* K;
  double tolerance_or_rank;
  std::vector<int> row_indices, col_indices;
  LowRank(kerneltype* K_, double tol, std::vector<int>& rows, std::vector<int>& cols)
      : K(K_), tolerance_or_rank(tol), row_indices(rows), col_indices(cols) {}
  void ACA_only_nodes2(std::vector<int>& row_bases, std::vector<int>& col_bases, int& computed_rank, Mat& Ac,
                       Mat& Ar) {
    const int nr = (int)row_indices.size(), nc = (int)col_indices.size();
    const int mn = std::min(nr, nc);
    row_bases.assign(std::max(mn, 1), 0);
    col_bases.assign(std::max(mn, 1), 0);
    computed_rank = 0;
    if (mn > 0)
      check(rcs_aca(K->handle(), RCS_PREC_C128, tolerance_or_rank, row_indices.data(), nr, col_indices.data(), nc,
                    &computed_rank, row_bases.data(), col_bases.data()));
    row_bases.resize(computed_rank);
    col_bases.resize(computed_rank);
    std::vector<int> pr, pc;
    for (int r : row_bases) pr.push_back(row_indices[r]);
    for (int c : col_bases) pc.push_back(col_indices[c]);
    Ac = K->getMatrix(row_indices, pc);
    Ar = K->getMatrix(pr, col_indices);
  }
};

// RRQR (include/rrqr.hpp): the partitioned column-pivoted Householder QR of
// src/rrqr.cpp:69-182. Produces Q_trunc and maxPivot (what the tree uses) and errorInQR;
// R_trunc and P are not formed (the tree reads them only in commented-out lines, T:1593-1594).
class RRQR {
 public:
  Mat A;
  int fillInRank, partition, flag;
  double inputPivot, maxPivot = 0;
  int TOL_POW1, TOL_POW2;
  Mat Q_trunc;
  int rank = 0;
  RRQR(Mat& A_, int tol_pow1, int tol_pow2, int partition_, double inputPivot_, int flag_, int fillInRank_)
      : A(A_), fillInRank(fillInRank_), partition(partition_), flag(flag_), inputPivot(inputPivot_),
        TOL_POW1(tol_pow1), TOL_POW2(tol_pow2) {}
  void colPivHouseholderQRWithPartition() {
    const int m = A.rows(), n = A.cols();
    Mat Q(m, std::max(1, std::min(m, n)));
    if (m > 0 && n > 0)
      check(rcs_rrqr(RCS_PREC_C128, A.data(), m, n, TOL_POW1, TOL_POW2, partition, inputPivot, fillInRank, Q.data(),
                     &rank, &maxPivot));
    Q_trunc = Mat(m, rank);
    for (int j = 0; j < rank; ++j)
      for (int i = 0; i < m; ++i) Q_trunc(i, j) = Q(i, j);
  }
  // errorInQR (src/rrqr.cpp:183-186): |A P - Q_trunc R_trunc|_F / |A|_F. R_trunc is the first
  // rank rows of Q^H A P for the orthonormal Householder Q, so the difference is
  // (I - Q_trunc Q_trunc^H) A P, whose norm does not depend on the column order P; it is
  // computed here on the host from Q_trunc and A.
  double errorInQR() const {
    const int m = A.rows(), n = A.cols(), r = Q_trunc.cols();
    double num = 0, den = 0;
    std::vector<cdouble> c(r);
    for (int j = 0; j < n; ++j) {
      for (int q = 0; q < r; ++q) {
        cdouble t = 0;
        for (int i = 0; i < m; ++i) t += std::conj(Q_trunc(i, q)) * A(i, j);
        c[q] = t;
      }
      for (int i = 0; i < m; ++i) {
        cdouble v = A(i, j);
        for (int q = 0; q < r; ++q) v -= Q_trunc(i, q) * c[q];
        num += std::norm(v);
        den += std::norm(A(i, j));
      }
    }
    return den > 0 ? std::sqrt(num / den) : 0.0;
  }
};

// What the reference driver reads from the tree through AIFMM::A (test/sphere.cpp:217-245):
// getAvgLeafSize (T:490-501), getAvgMultipoles (T:531-551), findMemory and
// findMemoryAfterFactorize (T:4511-4571), each of the solver's current state.
class TreeStatistics {
 public:
  explicit TreeStatistics(rcs_solver_t s) : s_(s) {}
  int getAvgLeafSize() { return (int)summary()[0]; }
  int getAvgMultipoles() { return (int)summary()[1]; }
  double findMemory() { return summary()[2]; }
  double findMemoryAfterFactorize() { return summary()[3]; }

 private:
  std::array<double, 4> summary() {
    std::array<double, 4> out{};
    check(rcs_solver_tree_summary(s_, out.data()));
    return out;
  }
  rcs_solver_t s_;
};

// AIFMM<K> (include/AIFMM.hpp): construction assembles on the GPU, factorize() runs the
// elimination, solve() returns the solution in box order (src/AIFMM.cpp:103-110) and
// getPhi() the solution in the input order (src/AIFMM.cpp:135-147) of the last solve.
template <typename kerneltype>
class AIFMM {
 public:
  AIFMM(inputsToSolverClass in, kerneltype* mykernel, int precision = RCS_PREC_C128) : K(mykernel) {
    const double bbox[6] = {in.xmin, in.xmax, in.ymin, in.ymax, in.zmin, in.zmax};
    check(rcs_solver_create(mykernel->handle(), bbox, in.leafSize, in.tol_pow, in.fillin_tol, in.fillin_rank,
                            precision, 0, &s_));
    A = new TreeStatistics(s_);
  }
  ~AIFMM() {
    delete A;
    if (s_) rcs_solver_destroy(s_);
  }
  AIFMM(const AIFMM&) = delete;
  AIFMM& operator=(const AIFMM&) = delete;
  void factorize() { check(rcs_solver_factorize(s_)); }
  Vec solve(Vec& rhs) {
    const int n = (int)rhs.size();
    phi_.assign(n, cdouble());
    Vec xb(n);
    check(rcs_solver_solve(s_, reinterpret_cast<const double*>(rhs.data()), 1, reinterpret_cast<double*>(phi_.data()),
                           reinterpret_cast<double*>(xb.data())));
    return xb;
  }
  // backSubstitute(rhs) (src/AIFMM.cpp:112-117): a solve whose box-order result is not returned.
  void backSubstitute(Vec& rhs) { (void)solve(rhs); }
  // The reference splits a solve into assigning the right-hand side (backSubstitute1,
  // src/AIFMM.cpp:119-122), the forward elimination (backSubstitute2, :124-127) and the back
  // substitution (backSubstitute3, :129-132). Here backSubstitute1 keeps the right-hand side,
  // backSubstitute2 runs both phases on the GPU in one call and backSubstitute3 has nothing
  // left to do, so a driver that times 2 and 3 together (test/sphere.cpp:253-258) still times
  // the whole solve; getPhi then returns its solution in the input order.
  void backSubstitute1(Vec& rhs) {
    rhs_ = rhs;
    pending_ = true;
  }
  void backSubstitute2() {
    if (!pending_) throw std::runtime_error("rcs_amd: backSubstitute2 called before backSubstitute1");
    (void)solve(rhs_);
    pending_ = false;
  }
  void backSubstitute3() {}
  void getPhi(Vec& phi) { phi = phi_; }
  rcs_solver_t handle() const { return s_; }
  kerneltype* K;
  TreeStatistics* A = nullptr;  // the reference's FMM2DTree* A, for the statistics drivers print

 private:
  rcs_solver_t s_ = nullptr;
  Vec phi_, rhs_;
  bool pending_ = false;
};

}  // namespace rcs

#endif
