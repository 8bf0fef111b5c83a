// ORACLE — test infrastructure only.
//
// CPU FP64 restatement of the vaishna77/RCS AIFMM factor-and-solve path. It is the
// reference answer for the HIP implementation in rcs_amd/: only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg may load it, never the product path.
//
// How it is pinned (tests/test_oracle.py; the reference cannot be built here and ships no
// numbers): RWG numbering on the reference's 78-patch fixture; panel integrals, the EFIE
// integration primitives and the Laplace entries against brute-force quadrature in every
// distance branch; the EFIE operator against the Mie series; the Laplace operator against the
// sphere's capacitance; the ACA stopping rule by hand-worked known answers and exact-rank
// recovery; the partitioned RRQR against LAPACK's pivoted QR and hand-worked thresholds; the
// elimination and solve against dense LU at tight tolerance; the tree statistics against
// other views of the same tree. Parity unpinned: bit-level rounding and pivot choices at
// exact ties of the reference's Eigen calls (PartialPivLU, ColPivHouseholderQR's maxPivot,
// std::sort's tie order in the RRQR); here the first maximal element is taken.
//
// Dense types and the Eigen operations the reference relies on (column-major complex
// matrices, PartialPivLU with first-max partial pivoting, products, adjoints, norms).
#pragma once
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace oracle {

using cd = std::complex<double>;
static const cd I1(0.0, 1.0);
static const double PI = 4.0 * std::atan(1.0);

// Physical constants, include/includes.hpp:53-63 (freq is a runtime argument here;
// the reference hard-codes 1e7 at includes.hpp:58).
struct Phys {
  double epsilon = 8.854187817e-12;
  double c = 3.0e8;
  double freq = 1.0e7;
  double mu = 4 * PI * 1e-7;
  double lambda() const { return c / freq; }
  double kappa() const { return 2 * PI / lambda(); }
  double C0() const { return 1.0 / (4 * PI * epsilon); }
};

struct V3 {
  double x = 0, y = 0, z = 0;
  V3() {}
  V3(double a, double b, double c) : x(a), y(b), z(c) {}
  V3 operator+(const V3& o) const { return V3(x + o.x, y + o.y, z + o.z); }
  V3 operator-(const V3& o) const { return V3(x - o.x, y - o.y, z - o.z); }
  V3 operator*(double s) const { return V3(x * s, y * s, z * s); }
  V3 operator/(double s) const { return V3(x / s, y / s, z / s); }
  double dot(const V3& o) const { return x * o.x + y * o.y + z * o.z; }
  V3 cross(const V3& o) const { return V3(y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x); }
  double norm() const { return std::sqrt(x * x + y * y + z * z); }
};

struct C3 {
  cd x, y, z;
  C3() : x(0), y(0), z(0) {}
  C3(cd a, cd b, cd c) : x(a), y(b), z(c) {}
  C3 operator+(const C3& o) const { return C3(x + o.x, y + o.y, z + o.z); }
  C3 operator*(cd s) const { return C3(x * s, y * s, z * s); }
};
inline C3 toC(const V3& v) { return C3(v.x, v.y, v.z); }

// Column-major dense complex matrix (the Eigen::MatrixXcd subset used by the path).
struct Mat {
  int r = 0, c = 0;
  std::vector<cd> a;
  Mat() {}
  Mat(int rows, int cols) : r(rows), c(cols), a((size_t)rows * cols, cd(0)) {}
  static Mat Zero(int rows, int cols) { return Mat(rows, cols); }
  static Mat Identity(int n) {
    Mat m(n, n);
    for (int i = 0; i < n; ++i) m(i, i) = 1.0;
    return m;
  }
  int rows() const { return r; }
  int cols() const { return c; }
  size_t size() const { return (size_t)r * c; }
  cd& operator()(int i, int j) { return a[(size_t)i + (size_t)j * r]; }
  const cd& operator()(int i, int j) const { return a[(size_t)i + (size_t)j * r]; }
  cd* col(int j) { return a.data() + (size_t)j * r; }
  const cd* col(int j) const { return a.data() + (size_t)j * r; }
  Mat block(int i0, int j0, int nr, int nc) const {
    Mat m(nr, nc);
    for (int j = 0; j < nc; ++j)
      for (int i = 0; i < nr; ++i) m(i, j) = (*this)(i0 + i, j0 + j);
    return m;
  }
  void setBlock(int i0, int j0, const Mat& m) {
    for (int j = 0; j < m.c; ++j)
      for (int i = 0; i < m.r; ++i) (*this)(i0 + i, j0 + j) = m(i, j);
  }
  double norm() const {  // Frobenius
    double s = 0;
    for (auto& v : a) s += std::norm(v);
    return std::sqrt(s);
  }
};

inline Mat adjoint(const Mat& A) {
  Mat B(A.c, A.r);
  for (int j = 0; j < A.c; ++j)
    for (int i = 0; i < A.r; ++i) B(j, i) = std::conj(A(i, j));
  return B;
}

// Threads of the dense products, LU and solves (1 by default: the reference driver runs
// factorize and solve on one thread, test/sphere.cpp:206-207). Every parallel loop below is
// over independent output columns, so results do not depend on the count.
inline int& dense_threads() {
  static int t = 1;
  return t;
}
inline bool dense_parallel(double work) { return dense_threads() > 1 && work > 5.0e5; }

// C = A*B (inner dimensions must agree, as Eigen asserts).
inline Mat mul(const Mat& A, const Mat& B) {
  if (A.c != B.r) throw std::runtime_error("oracle: invalid matrix product");
  Mat C(A.r, B.c);
#pragma omp parallel for schedule(static) num_threads(dense_threads()) \
    if (dense_parallel((double)A.r * A.c * B.c))
  for (int j = 0; j < B.c; ++j) {
    cd* cc = C.col(j);
    for (int k = 0; k < A.c; ++k) {
      cd b = B(k, j);
      if (b == cd(0)) continue;
      const cd* ac = A.col(k);
      for (int i = 0; i < A.r; ++i) cc[i] += ac[i] * b;
    }
  }
  return C;
}

inline Mat add(const Mat& A, const Mat& B) {
  if (A.r != B.r || A.c != B.c) throw std::runtime_error("oracle: add size mismatch");
  Mat C = A;
  for (size_t i = 0; i < C.a.size(); ++i) C.a[i] += B.a[i];
  return C;
}
inline Mat sub(const Mat& A, const Mat& B) {
  if (A.r != B.r || A.c != B.c) throw std::runtime_error("oracle: sub size mismatch");
  Mat C = A;
  for (size_t i = 0; i < C.a.size(); ++i) C.a[i] -= B.a[i];
  return C;
}
inline Mat neg(const Mat& A) {
  Mat C = A;
  for (auto& v : C.a) v = -v;
  return C;
}

// Partial-pivoting LU, Eigen::PartialPivLU semantics: at column k the pivot is the
// first row of maximal |a_ik| (Eigen scores with abs and keeps the first max).
struct LU {
  Mat lu;
  std::vector<int> perm;  // row i of P*A is row perm[i] of A
  int n = 0;
  LU() {}
  explicit LU(const Mat& A) { compute(A); }
  void compute(const Mat& A) {
    if (A.r != A.c) throw std::runtime_error("oracle: LU of non-square matrix");
    n = A.r;
    lu = A;
    perm.resize(n);
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int k = 0; k < n; ++k) {
      int p = k;
      double best = std::abs(lu(k, k));
      for (int i = k + 1; i < n; ++i) {
        double v = std::abs(lu(i, k));
        if (v > best) { best = v; p = i; }
      }
      if (p != k) {
        for (int j = 0; j < n; ++j) std::swap(lu(k, j), lu(p, j));
        std::swap(perm[k], perm[p]);
      }
      cd piv = lu(k, k);
      if (piv != cd(0)) {
        for (int i = k + 1; i < n; ++i) lu(i, k) /= piv;
      }
#pragma omp parallel for schedule(static) num_threads(dense_threads()) \
    if (dense_parallel((double)(n - k) * (n - k)))
      for (int j = k + 1; j < n; ++j) {
        cd f = lu(k, j);
        if (f == cd(0)) continue;
        cd* cj = lu.col(j);
        const cd* ck = lu.col(k);
        for (int i = k + 1; i < n; ++i) cj[i] -= ck[i] * f;
      }
    }
  }
  Mat solve(const Mat& B) const {
    if (B.r != n) throw std::runtime_error("oracle: LU solve size mismatch");
    Mat X(n, B.c);
#pragma omp parallel for schedule(static) num_threads(dense_threads()) \
    if (dense_parallel((double)n * n * B.c))
    for (int j = 0; j < B.c; ++j) {
      cd* x = X.col(j);
      for (int i = 0; i < n; ++i) x[i] = B(perm[i], j);
      for (int k = 0; k < n; ++k) {  // unit lower
        cd v = x[k];
        if (v == cd(0)) continue;
        const cd* l = lu.col(k);
        for (int i = k + 1; i < n; ++i) x[i] -= l[i] * v;
      }
      for (int k = n - 1; k >= 0; --k) {  // upper
        x[k] /= lu(k, k);
        cd v = x[k];
        if (v == cd(0)) continue;
        const cd* u = lu.col(k);
        for (int i = 0; i < k; ++i) x[i] -= u[i] * v;
      }
    }
    return X;
  }
};

// Work counters: algorithmic flops per reference operation (complex GEMM 8mnk,
// GETRF (8/3)n^3, GETRS 8 n^2 nrhs), shared by the CPU baseline report.
struct Counters {
  double gemm_flops = 0, getrf_flops = 0, getrs_flops = 0, qr_flops = 0;
  long long entries = 0;
  long long n_rrqr = 0, n_compress_p2p = 0, n_compress_p2l = 0;
};

}  // namespace oracle
