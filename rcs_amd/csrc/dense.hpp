// Dense complex linear algebra on HBM-resident blocks (column-major, explicit ld).
//
// These kernels replace the Eigen operations of the reference's factor/solve path:
// products (GEMM), PartialPivLU (blocked GETRF with first-max partial pivoting), LU
// solves (GETRS), Frobenius norms, adjoints and block copies, and the partitioned
// column-pivoted Householder QR of src/rrqr.cpp.
#pragma once
#include "ctx.hpp"

namespace rcs {

template <typename T>
struct DMat {  // view of a device matrix
  T* p = nullptr;
  int r = 0, c = 0, ld = 0;
  size_t size() const { return (size_t)r * c; }
  T* col(int j) const { return p + (size_t)j * ld; }
  DMat cols(int j0, int n) const { DMat m = *this; m.p = p + (size_t)j0 * ld; m.c = n; return m; }
  DMat rows(int i0, int n) const { DMat m = *this; m.p = p + i0; m.r = n; return m; }
};

enum Op { OP_N = 0, OP_C = 1 };  // no transpose / conjugate transpose

template <typename T>
struct GemmOp {
  const T* A;
  const T* B;
  T* C;
  int lda, ldb, ldc;
  int m, n, k;
  int opA, opB;
  T alpha, beta;
  // chain = 1: accumulate into the same C as the previous op of the group (same m, n,
  // alpha); the products are summed inside one tile and C is written once, so no two
  // workgroups ever update the same output block. An op whose C differs from the
  // previous op's starts a new output block with its own beta.
  int chain;
};

// C_g = alpha*sum_s op(A_s)*op(B_s) + beta*C_g for every output block of the group, in
// one launch. Two output blocks of one group must not share C (checked on the host).
// role 1 marks the Schur-complement updates of the elimination (a separate kernel
// instantiation, so traces attribute them); 0 is every other product. prof_tag >= 0
// times the kernel alone with HIP events under that profiler tag (algorithmic flops
// 8*m*n*k per product).
template <typename T>
void gemm_grouped(std::vector<GemmOp<T>>& ops, Ctx& cx, int role = 0, int prof_tag = -1);

// In-place LU with partial pivoting (first maximal |a| in the column). ipiv is a device
// array of 2n ints: ipiv[0, n) = 0-based absolute row swapped with each row (LAPACK
// style), ipiv[n, 2n) = the composed permutation (row i of PA is row perm[i] of A),
// which lets solves gather the right-hand side in parallel.
template <typename T>
void getrf(DMat<T> A, int* ipiv, Ctx& cx);
// Solve (LU) X = B in place of B.
template <typename T>
void getrs(DMat<T> LU, const int* ipiv, DMat<T> B, Ctx& cx);

template <typename T>
void copy2d(DMat<T> dst, DMat<T> src, hipStream_t st);            // dst = src
template <typename T>
void adjoint2d(DMat<T> dst, DMat<T> src, hipStream_t st);         // dst = src^H
template <typename T>
void fill_identity(DMat<T> dst, T diag, hipStream_t st);          // dst = diag*I (zeros elsewhere)
template <typename T>
void zero2d(DMat<T> dst, hipStream_t st);
template <typename T>
void axpy2d(DMat<T> y, DMat<T> x, T a, hipStream_t st);            // y += a*x
// Squared Frobenius norms of many blocks into out[i] (device doubles).
template <typename T>
void sqnorms(const std::vector<DMat<T>>& mats, double* d_out, Ctx& cx);
// Many independent block copies dst_i = src_i (optionally adjoint) in one launch.
template <typename T>
struct CopyOp {
  const T* src;
  T* dst;
  int lds, ldd, r, c;  // shape of dst
  int adj;             // dst = src^H when set
  int tile0;
};
template <typename T>
void copy_batched(std::vector<CopyOp<T>>& ops, Ctx& cx);
// dst(i, c) = src(idx[i], c) for i < dst.r (gather rows); scatter: dst(idx[i], c) = src(i, c).
template <typename T>
void gather_rows(DMat<T> dst, DMat<T> src, const int* d_idx, hipStream_t st);
template <typename T>
void scatter_rows(DMat<T> dst, DMat<T> src, const int* d_idx, hipStream_t st);
// Cast between precisions (same shape).
void cast_c128_to_c64(DMat<cpx<float>> dst, DMat<cpx<double>> src, hipStream_t st);
void cast_c64_to_c128(DMat<cpx<double>> dst, DMat<cpx<float>> src, hipStream_t st);

// Two-partition column-pivoted Householder QR (RRQR::colPivHouseholderQRWithPartition,
// src/rrqr.cpp:69-182). A (m x n) is read only. Q (m x min(m,n)) receives Q_trunc.
struct RRQRInfo {
  int rank, rank1, rank2;
  double maxPivot;
};
template <typename T>
RRQRInfo rrqr(DMat<T> A, int partition, double tol1, double tol2, double inputPivot, int fillInRank, DMat<T> Q,
              Ctx& cx);
// Several independent RRQRs with the same tolerances through one launch per pivot step.
template <typename T>
struct RRQRTask {
  DMat<T> A;         // input, read only
  int partition = 0;
  double inputPivot = 0;
  int fillInRank = 0;
  DMat<T> Q;         // m x (>= rank) output
  RRQRInfo info{0, 0, 0, 0.0};
};
template <typename T>
void rrqr_batched(std::vector<RRQRTask<T>>& tasks, double tol1, double tol2, Ctx& cx);

}  // namespace rcs
