// Dense complex kernels for gfx950 (CDNA4, 64-wide wavefronts).
#include <algorithm>

#include "dense.hpp"

namespace rcs {

template <typename T> struct RealOf;
template <> struct RealOf<cpx<double>> { using type = double; };
template <> struct RealOf<cpx<float>> { using type = float; };

__device__ inline double dabs(cpx<double> a) { return hypot(a.re, a.im); }
__device__ inline float dabs(cpx<float> a) { return hypotf(a.re, a.im); }

template <typename T>
__device__ inline T ld_op(const T* A, int lda, int i, int k, int op) {
  // element (i, k) of op(A)
  if (op == OP_N) return A[(size_t)i + (size_t)k * lda];
  T v = A[(size_t)k + (size_t)i * lda];
  v.im = -v.im;
  return v;
}

// ---------------------------------------------------------------- grouped GEMM
// 64x64 output tile per 256-thread workgroup, K staged through LDS in chunks of 16.
// Thread (tx, ty) accumulates rows ty+16i and columns tx+16j (i, j < 4). An output
// block may sum several products (segments); the tile walks them all before writing.
constexpr int GT = 64, GK = 16;

template <typename T>
struct GemmSeg {  // one product op(A)*op(B) of an output block
  const T* A;
  const T* B;
  int lda, ldb, k, opA, opB;
};
template <typename T>
struct GemmOut {  // one output block
  T* C;
  int ldc, m, n;
  T alpha, beta;
  int tile0, seg0, nseg;
};

template <typename T>
__global__ void __launch_bounds__(256) gemm_grouped_kernel(const GemmOut<T>* outs, int nouts,
                                                           const GemmSeg<T>* segs) {
  using R = typename RealOf<T>::type;
  __shared__ T As[GK][GT + 1];
  __shared__ T Bs[GK][GT + 1];
  const int tile = blockIdx.x;
  int lo = 0, hi = nouts - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (outs[mid].tile0 <= tile) lo = mid;
    else hi = mid - 1;
  }
  const GemmOut<T> out = outs[lo];
  const int t = tile - out.tile0;
  const int tm = (out.m + GT - 1) / GT;
  const int m0 = (t % tm) * GT, n0 = (t / tm) * GT;
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  T acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = T(0, 0);
  for (int sgi = 0; sgi < out.nseg; ++sgi) {
    const GemmSeg<T> op = segs[out.seg0 + sgi];
    for (int k0 = 0; k0 < op.k; k0 += GK) {
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        int i, kk;
        if (op.opA == OP_N) { i = tid & 63; kk = (tid >> 6) + 4 * s; }
        else { kk = tid & 15; i = (tid >> 4) + 16 * s; }
        T v(0, 0);
        if (m0 + i < out.m && k0 + kk < op.k) v = ld_op(op.A, op.lda, m0 + i, k0 + kk, op.opA);
        As[kk][i] = v;
        int j2, k2;
        if (op.opB == OP_N) { k2 = tid & 15; j2 = (tid >> 4) + 16 * s; }
        else { j2 = tid & 63; k2 = (tid >> 6) + 4 * s; }
        T w(0, 0);
        if (n0 + j2 < out.n && k0 + k2 < op.k) w = ld_op(op.B, op.ldb, k0 + k2, n0 + j2, op.opB);
        Bs[k2][j2] = w;
      }
      __syncthreads();
#pragma unroll 4
      for (int kk = 0; kk < GK; ++kk) {
        T a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = As[kk][ty + 16 * i];
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx + 16 * j];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) cfma(acc[i][j], a[i], b[j]);
      }
      __syncthreads();
    }
  }
  const bool zb = (out.beta.re == R(0) && out.beta.im == R(0));
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    int r = m0 + ty + 16 * i;
    if (r >= out.m) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      int c = n0 + tx + 16 * j;
      if (c >= out.n) continue;
      T* cp = out.C + (size_t)r + (size_t)c * out.ldc;
      T v = out.alpha * acc[i][j];
      if (!zb) v = v + out.beta * (*cp);
      *cp = v;
    }
  }
}

template <typename T>
void gemm_grouped(std::vector<GemmOp<T>>& ops, Ctx& cx) {
  if (ops.empty()) return;
  std::vector<GemmOut<T>> outs;
  std::vector<GemmSeg<T>> segs;
  outs.reserve(ops.size());
  segs.reserve(ops.size());
  int tiles = 0;
  for (const auto& o : ops) {
    if (o.m <= 0 || o.n <= 0) continue;
    const bool join = o.chain && !outs.empty() && outs.back().C == o.C;
    if (join) {
      const GemmOut<T>& h = outs.back();
      if (h.m != o.m || h.n != o.n || h.ldc != o.ldc || h.alpha.re != o.alpha.re || h.alpha.im != o.alpha.im)
        throw std::runtime_error("gemm_grouped: chained op differs from its output block");
    } else {
      GemmOut<T> h{};
      h.C = o.C; h.ldc = o.ldc; h.m = o.m; h.n = o.n;
      h.alpha = o.alpha; h.beta = o.beta;
      h.tile0 = tiles; h.seg0 = (int)segs.size(); h.nseg = 0;
      tiles += ((o.m + GT - 1) / GT) * ((o.n + GT - 1) / GT);
      outs.push_back(h);
    }
    if (o.k > 0) {  // k = 0 contributes nothing: the block becomes beta*C
      segs.push_back(GemmSeg<T>{o.A, o.B, o.lda, o.ldb, o.k, o.opA, o.opB});
      outs.back().nseg++;
    }
  }
  if (outs.empty()) return;
  if (outs.size() > 1) {  // workgroups of distinct output blocks run concurrently
    std::vector<const T*> cs(outs.size());
    for (size_t i = 0; i < outs.size(); ++i) cs[i] = outs[i].C;
    std::sort(cs.begin(), cs.end());
    if (std::adjacent_find(cs.begin(), cs.end()) != cs.end())
      throw std::runtime_error("gemm_grouped: two output blocks of one launch share C (chain them)");
  }
  if (segs.empty()) segs.push_back(GemmSeg<T>{});
  const GemmOut<T>* d_outs = (const GemmOut<T>*)cx.push(outs.data(), sizeof(GemmOut<T>) * outs.size());
  const GemmSeg<T>* d_segs = (const GemmSeg<T>*)cx.push(segs.data(), sizeof(GemmSeg<T>) * segs.size());
  hipLaunchKernelGGL(gemm_grouped_kernel<T>, dim3(tiles), dim3(256), 0, cx.st, d_outs, (int)outs.size(), d_segs);
  HIPCHK(hipGetLastError());
}

// ---------------------------------------------------------------- LU factorization
constexpr int LU_NB = 32;

// Factor the panel A[j:n, j:j+nb] (one workgroup). Pivot = first row of maximal |a|.
template <typename T>
__global__ void __launch_bounds__(1024) lu_panel_kernel(T* A, int lda, int n, int j, int nb, int* ipiv) {
  using R = typename RealOf<T>::type;
  __shared__ R sval[1024];
  __shared__ int sidx[1024];
  __shared__ int piv_s;
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int jj = 0; jj < nb; ++jj) {
    const int c = j + jj;
    T* colc = A + (size_t)c * lda;
    R best = R(-1);
    int bi = n;
    for (int r = c + tid; r < n; r += nt) {
      R v = dabs(colc[r]);
      if (v > best) { best = v; bi = r; }
    }
    sval[tid] = best;
    sidx[tid] = bi;
    __syncthreads();
    for (int s = nt / 2; s > 0; s >>= 1) {
      if (tid < s) {
        R v2 = sval[tid + s];
        int i2 = sidx[tid + s];
        if (v2 > sval[tid] || (v2 == sval[tid] && i2 < sidx[tid])) { sval[tid] = v2; sidx[tid] = i2; }
      }
      __syncthreads();
    }
    if (tid == 0) {
      int p = sidx[0];
      if (p >= n) p = c;
      piv_s = p;
      ipiv[c] = p;
    }
    __syncthreads();
    const int p = piv_s;
    if (p != c) {
      for (int q = tid; q < nb; q += nt) {
        T* cq = A + (size_t)(j + q) * lda;
        T tmp = cq[c];
        cq[c] = cq[p];
        cq[p] = tmp;
      }
    }
    __syncthreads();
    const T piv = colc[c];
    const bool nz = !(piv.re == R(0) && piv.im == R(0));
    if (nz) {
      for (int r = c + 1 + tid; r < n; r += nt) colc[r] = cdiv(colc[r], piv);
    }
    __syncthreads();
    // rank-1 update of the remaining panel columns
    const int rows = n - c - 1, cols = nb - jj - 1;
    for (int e = tid; e < rows * cols; e += nt) {
      int r = c + 1 + e % rows, q = c + 1 + e / rows;
      T* cq = A + (size_t)q * lda;
      T f = cq[c];
      T l = colc[r];
      cq[r] = cq[r] - l * f;
    }
    __syncthreads();
  }
}

// Apply the row interchanges of rows [j, j+nb) to the columns outside the panel.
template <typename T>
__global__ void laswp_kernel(T* A, int lda, int ncols, int j, int nb, const int* ipiv, int skip0, int skip1) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncols || (c >= skip0 && c < skip1)) return;
  T* col = A + (size_t)c * lda;
  for (int r = j; r < j + nb; ++r) {
    int p = ipiv[r];
    if (p != r) {
      T t = col[r];
      col[r] = col[p];
      col[p] = t;
    }
  }
}

// B[j:j+nb, c] <- L11^{-1} B[j:j+nb, c] (unit lower L11 = A[j:j+nb, j:j+nb]).
template <typename T>
__global__ void trsm_lower_unit_kernel(const T* L, int ldl, int j, int nb, T* B, int ldb, int ncols) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncols) return;
  T* b = B + (size_t)c * ldb + j;
  T x[LU_NB];
  for (int i = 0; i < nb; ++i) x[i] = b[i];
  for (int k = 0; k < nb; ++k) {
    const T* lk = L + (size_t)(j + k) * ldl + j;
    T xk = x[k];
    for (int i = k + 1; i < nb; ++i) x[i] = x[i] - lk[i] * xk;
  }
  for (int i = 0; i < nb; ++i) b[i] = x[i];
}

// B[j:j+nb, c] <- U11^{-1} B[j:j+nb, c] (upper, non-unit).
template <typename T>
__global__ void trsm_upper_kernel(const T* U, int ldu, int j, int nb, T* B, int ldb, int ncols) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncols) return;
  T* b = B + (size_t)c * ldb + j;
  T x[LU_NB];
  for (int i = 0; i < nb; ++i) x[i] = b[i];
  for (int k = nb - 1; k >= 0; --k) {
    const T* uk = U + (size_t)(j + k) * ldu + j;
    x[k] = cdiv(x[k], uk[k]);
    T xk = x[k];
    for (int i = 0; i < k; ++i) x[i] = x[i] - uk[i] * xk;
  }
  for (int i = 0; i < nb; ++i) b[i] = x[i];
}

template <typename T>
void getrf(DMat<T> A, int* ipiv, Ctx& cx) {
  hipStream_t st = cx.st;
  const int n = A.r;
  if (n == 0) return;
  if (A.c != n) throw std::runtime_error("getrf: matrix is not square");
  for (int j = 0; j < n; j += LU_NB) {
    const int nb = std::min(LU_NB, n - j);
    int threads = 1024;
    if (n - j <= 256) threads = 256;
    else if (n - j <= 512) threads = 512;
    hipLaunchKernelGGL(lu_panel_kernel<T>, dim3(1), dim3(threads), 0, st, A.p, A.ld, n, j, nb, ipiv);
    hipLaunchKernelGGL(laswp_kernel<T>, dim3((n + 127) / 128), dim3(128), 0, st, A.p, A.ld, n, j, nb,
                       (const int*)ipiv, j, j + nb);
    const int rest = n - j - nb;
    if (rest > 0) {
      hipLaunchKernelGGL(trsm_lower_unit_kernel<T>, dim3((rest + 63) / 64), dim3(64), 0, st, (const T*)A.p, A.ld,
                         j, nb, A.p + (size_t)(j + nb) * A.ld, A.ld, rest);
      std::vector<GemmOp<T>> g(1);
      g[0].A = A.p + (size_t)j * A.ld + j + nb; g[0].lda = A.ld; g[0].opA = OP_N;
      g[0].B = A.p + (size_t)(j + nb) * A.ld + j; g[0].ldb = A.ld; g[0].opB = OP_N;
      g[0].C = A.p + (size_t)(j + nb) * A.ld + j + nb; g[0].ldc = A.ld;
      g[0].m = rest; g[0].n = rest; g[0].k = nb;
      g[0].alpha = T(-1, 0); g[0].beta = T(1, 0);
      gemm_grouped(g, cx);
    }
  }
  HIPCHK(hipGetLastError());
}

// One workgroup per right-hand side column: row swaps, unit-lower then upper solve.
template <typename T>
__global__ void __launch_bounds__(256) getrs_vec_kernel(const T* LU, int ld, int n, const int* ipiv, T* B, int ldb) {
  T* b = B + (size_t)blockIdx.x * ldb;
  const int tid = threadIdx.x;
  if (tid == 0)
    for (int r = 0; r < n; ++r) {
      int p = ipiv[r];
      if (p != r) { T t = b[r]; b[r] = b[p]; b[p] = t; }
    }
  __syncthreads();
  __shared__ T xs[64];
  // forward, blocks of 64
  for (int j0 = 0; j0 < n; j0 += 64) {
    const int nb = min(64, n - j0);
    if (tid < 64) {  // one wavefront solves the diagonal block
      T x = tid < nb ? b[j0 + tid] : T(0, 0);
      for (int k = 0; k < nb; ++k) {
        T xk;
        xk.re = __shfl(x.re, k);
        xk.im = __shfl(x.im, k);
        if (tid > k && tid < nb) x = x - LU[(size_t)(j0 + k) * ld + j0 + tid] * xk;
      }
      if (tid < nb) { b[j0 + tid] = x; xs[tid] = x; }
    }
    __syncthreads();
    for (int r = j0 + nb + tid; r < n; r += blockDim.x) {
      T s = b[r];
      for (int k = 0; k < nb; ++k) s = s - LU[(size_t)(j0 + k) * ld + r] * xs[k];
      b[r] = s;
    }
    __syncthreads();
  }
  // backward
  for (int j1 = n; j1 > 0; j1 -= 64) {
    const int j0 = max(0, j1 - 64), nb = j1 - j0;
    if (tid < 64) {
      T x = tid < nb ? b[j0 + tid] : T(0, 0);
      for (int k = nb - 1; k >= 0; --k) {
        if (tid == k) x = cdiv(x, LU[(size_t)(j0 + k) * ld + j0 + k]);
        T xk;
        xk.re = __shfl(x.re, k);
        xk.im = __shfl(x.im, k);
        if (tid < k) x = x - LU[(size_t)(j0 + k) * ld + j0 + tid] * xk;
      }
      if (tid < nb) { b[j0 + tid] = x; xs[tid] = x; }
    }
    __syncthreads();
    for (int r = tid; r < j0; r += blockDim.x) {
      T s = b[r];
      for (int k = 0; k < nb; ++k) s = s - LU[(size_t)(j0 + k) * ld + r] * xs[k];
      b[r] = s;
    }
    __syncthreads();
  }
}

template <typename T>
__global__ void laswp_all_kernel(T* B, int ldb, int ncols, int n, const int* ipiv) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncols) return;
  T* col = B + (size_t)c * ldb;
  for (int r = 0; r < n; ++r) {
    int p = ipiv[r];
    if (p != r) { T t = col[r]; col[r] = col[p]; col[p] = t; }
  }
}

template <typename T>
void getrs(DMat<T> LU, const int* ipiv, DMat<T> B, Ctx& cx) {
  hipStream_t st = cx.st;
  const int n = LU.r;
  if (n == 0 || B.c == 0) return;
  if (B.r != n) throw std::runtime_error("getrs: size mismatch");
  if (B.c <= 8) {
    hipLaunchKernelGGL(getrs_vec_kernel<T>, dim3(B.c), dim3(256), 0, st, (const T*)LU.p, LU.ld, n, ipiv, B.p, B.ld);
    HIPCHK(hipGetLastError());
    return;
  }
  const int nc = B.c;
  hipLaunchKernelGGL(laswp_all_kernel<T>, dim3((nc + 127) / 128), dim3(128), 0, st, B.p, B.ld, nc, n, ipiv);
  for (int j = 0; j < n; j += LU_NB) {
    const int nb = std::min(LU_NB, n - j);
    hipLaunchKernelGGL(trsm_lower_unit_kernel<T>, dim3((nc + 63) / 64), dim3(64), 0, st, (const T*)LU.p, LU.ld, j,
                       nb, B.p, B.ld, nc);
    const int rest = n - j - nb;
    if (rest > 0) {
      std::vector<GemmOp<T>> g(1);
      g[0].A = LU.p + (size_t)j * LU.ld + j + nb; g[0].lda = LU.ld; g[0].opA = OP_N;
      g[0].B = B.p + j; g[0].ldb = B.ld; g[0].opB = OP_N;
      g[0].C = B.p + j + nb; g[0].ldc = B.ld;
      g[0].m = rest; g[0].n = nc; g[0].k = nb;
      g[0].alpha = T(-1, 0); g[0].beta = T(1, 0);
      gemm_grouped(g, cx);
    }
  }
  const int nblk = (n + LU_NB - 1) / LU_NB;
  for (int bi = nblk - 1; bi >= 0; --bi) {
    const int j = bi * LU_NB, nb = std::min(LU_NB, n - j);
    hipLaunchKernelGGL(trsm_upper_kernel<T>, dim3((nc + 63) / 64), dim3(64), 0, st, (const T*)LU.p, LU.ld, j, nb,
                       B.p, B.ld, nc);
    if (j > 0) {
      std::vector<GemmOp<T>> g(1);
      g[0].A = LU.p + (size_t)j * LU.ld; g[0].lda = LU.ld; g[0].opA = OP_N;
      g[0].B = B.p + j; g[0].ldb = B.ld; g[0].opB = OP_N;
      g[0].C = B.p; g[0].ldc = B.ld;
      g[0].m = j; g[0].n = nc; g[0].k = nb;
      g[0].alpha = T(-1, 0); g[0].beta = T(1, 0);
      gemm_grouped(g, cx);
    }
  }
  HIPCHK(hipGetLastError());
}

// ---------------------------------------------------------------- elementwise
template <typename T>
__global__ void copy2d_kernel(T* d, int ldd, const T* s, int lds, int r, int c, int adj) {
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)r * c) return;
  int i = e % r, j = e / r;
  if (!adj) d[(size_t)i + (size_t)j * ldd] = s[(size_t)i + (size_t)j * lds];
  else {
    T v = s[(size_t)j + (size_t)i * lds];
    v.im = -v.im;
    d[(size_t)i + (size_t)j * ldd] = v;
  }
}
template <typename T>
void copy2d(DMat<T> dst, DMat<T> src, hipStream_t st) {
  if (dst.r != src.r || dst.c != src.c) throw std::runtime_error("copy2d: shape mismatch");
  size_t n = dst.size();
  if (!n) return;
  hipLaunchKernelGGL(copy2d_kernel<T>, dim3((n + 255) / 256), dim3(256), 0, st, dst.p, dst.ld, (const T*)src.p,
                     src.ld, dst.r, dst.c, 0);
}
template <typename T>
void adjoint2d(DMat<T> dst, DMat<T> src, hipStream_t st) {
  if (dst.r != src.c || dst.c != src.r) throw std::runtime_error("adjoint2d: shape mismatch");
  size_t n = dst.size();
  if (!n) return;
  hipLaunchKernelGGL(copy2d_kernel<T>, dim3((n + 255) / 256), dim3(256), 0, st, dst.p, dst.ld, (const T*)src.p,
                     src.ld, dst.r, dst.c, 1);
}
template <typename T>
__global__ void fill_kernel(T* d, int ld, int r, int c, T diag, int ident) {
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)r * c) return;
  int i = e % r, j = e / r;
  d[(size_t)i + (size_t)j * ld] = (ident && i == j) ? diag : T(0, 0);
}
template <typename T>
void fill_identity(DMat<T> dst, T diag, hipStream_t st) {
  size_t n = dst.size();
  if (!n) return;
  hipLaunchKernelGGL(fill_kernel<T>, dim3((n + 255) / 256), dim3(256), 0, st, dst.p, dst.ld, dst.r, dst.c, diag, 1);
}
template <typename T>
void zero2d(DMat<T> dst, hipStream_t st) {
  size_t n = dst.size();
  if (!n) return;
  hipLaunchKernelGGL(fill_kernel<T>, dim3((n + 255) / 256), dim3(256), 0, st, dst.p, dst.ld, dst.r, dst.c, T(0, 0), 0);
}
template <typename T>
__global__ void axpy_kernel(T* y, int ldy, const T* x, int ldx, int r, int c, T a) {
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)r * c) return;
  int i = e % r, j = e / r;
  y[(size_t)i + (size_t)j * ldy] = y[(size_t)i + (size_t)j * ldy] + a * x[(size_t)i + (size_t)j * ldx];
}
template <typename T>
void axpy2d(DMat<T> y, DMat<T> x, T a, hipStream_t st) {
  if (y.r != x.r || y.c != x.c) throw std::runtime_error("axpy2d: shape mismatch");
  size_t n = y.size();
  if (!n) return;
  hipLaunchKernelGGL(axpy_kernel<T>, dim3((n + 255) / 256), dim3(256), 0, st, y.p, y.ld, (const T*)x.p, x.ld, y.r,
                     y.c, a);
}

template <typename T>
struct NormTask {
  const T* p;
  int r, c, ld;
};
template <typename T>
__global__ void __launch_bounds__(256) sqnorm_kernel(const NormTask<T>* tasks, double* out) {
  const NormTask<T> t = tasks[blockIdx.x];
  double s = 0;
  size_t n = (size_t)t.r * t.c;
  for (size_t e = threadIdx.x; e < n; e += blockDim.x) {
    int i = e % t.r, j = e / t.r;
    T v = t.p[(size_t)i + (size_t)j * t.ld];
    s += (double)v.re * v.re + (double)v.im * v.im;
  }
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}
template <typename T>
void sqnorms(const std::vector<DMat<T>>& mats, double* d_out, Ctx& cx) {
  hipStream_t st = cx.st;
  if (mats.empty()) return;
  std::vector<NormTask<T>> t(mats.size());
  for (size_t i = 0; i < mats.size(); ++i) t[i] = {mats[i].p, mats[i].r, mats[i].c, mats[i].ld};
  const NormTask<T>* d = (const NormTask<T>*)cx.push(t.data(), sizeof(NormTask<T>) * t.size());
  hipLaunchKernelGGL(sqnorm_kernel<T>, dim3(t.size()), dim3(256), 0, st, d, d_out);
  HIPCHK(hipGetLastError());
}

template <typename T>
__global__ void __launch_bounds__(256) copy_batched_kernel(const CopyOp<T>* ops, int nops) {
  const int tile = blockIdx.x;
  int lo = 0, hi = nops - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (ops[mid].tile0 <= tile) lo = mid;
    else hi = mid - 1;
  }
  const CopyOp<T> op = ops[lo];
  const size_t e = (size_t)(tile - op.tile0) * 1024;
  const size_t n = (size_t)op.r * op.c;
  for (int s = 0; s < 4; ++s) {
    size_t x = e + s * 256 + threadIdx.x;
    if (x >= n) return;
    int i = x % op.r, j = x / op.r;
    T v;
    if (!op.adj) v = op.src[(size_t)i + (size_t)j * op.lds];
    else { v = op.src[(size_t)j + (size_t)i * op.lds]; v.im = -v.im; }
    op.dst[(size_t)i + (size_t)j * op.ldd] = v;
  }
}
template <typename T>
void copy_batched(std::vector<CopyOp<T>>& ops, Ctx& cx) {
  std::vector<CopyOp<T>> live;
  int tiles = 0;
  for (auto& o : ops) {
    size_t n = (size_t)o.r * o.c;
    if (!n) continue;
    o.tile0 = tiles;
    tiles += (int)((n + 1023) / 1024);
    live.push_back(o);
  }
  if (live.empty()) return;
  const CopyOp<T>* d = (const CopyOp<T>*)cx.push(live.data(), sizeof(CopyOp<T>) * live.size());
  hipLaunchKernelGGL(copy_batched_kernel<T>, dim3(tiles), dim3(256), 0, cx.st, d, (int)live.size());
  HIPCHK(hipGetLastError());
}
template <typename T>
__global__ void gather_kernel(T* d, int ldd, const T* s, int lds, int r, int c, const int* idx, int scatter) {
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)r * c) return;
  int i = e % r, j = e / r;
  if (!scatter) d[(size_t)i + (size_t)j * ldd] = s[(size_t)idx[i] + (size_t)j * lds];
  else d[(size_t)idx[i] + (size_t)j * ldd] = s[(size_t)i + (size_t)j * lds];
}
template <typename T>
void gather_rows(DMat<T> dst, DMat<T> src, const int* d_idx, hipStream_t st) {
  size_t n = dst.size();
  if (!n) return;
  hipLaunchKernelGGL(gather_kernel<T>, dim3((n + 255) / 256), dim3(256), 0, st, dst.p, dst.ld, (const T*)src.p, src.ld,
                     dst.r, dst.c, d_idx, 0);
}
template <typename T>
void scatter_rows(DMat<T> dst, DMat<T> src, const int* d_idx, hipStream_t st) {
  size_t n = src.size();
  if (!n) return;
  hipLaunchKernelGGL(gather_kernel<T>, dim3((n + 255) / 256), dim3(256), 0, st, dst.p, dst.ld, (const T*)src.p, src.ld,
                     src.r, src.c, d_idx, 1);
}

__global__ void cast_kernel_dn(cpx<float>* d, int ldd, const cpx<double>* s, int lds, int r, int c) {
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)r * c) return;
  int i = e % r, j = e / r;
  cpx<double> v = s[(size_t)i + (size_t)j * lds];
  d[(size_t)i + (size_t)j * ldd] = cpx<float>((float)v.re, (float)v.im);
}
__global__ void cast_kernel_up(cpx<double>* d, int ldd, const cpx<float>* s, int lds, int r, int c) {
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)r * c) return;
  int i = e % r, j = e / r;
  cpx<float> v = s[(size_t)i + (size_t)j * lds];
  d[(size_t)i + (size_t)j * ldd] = cpx<double>(v.re, v.im);
}
void cast_c128_to_c64(DMat<cpx<float>> dst, DMat<cpx<double>> src, hipStream_t st) {
  size_t n = dst.size();
  if (!n) return;
  hipLaunchKernelGGL(cast_kernel_dn, dim3((n + 255) / 256), dim3(256), 0, st, dst.p, dst.ld, (const cpx<double>*)src.p,
                     src.ld, dst.r, dst.c);
}
void cast_c64_to_c128(DMat<cpx<double>> dst, DMat<cpx<float>> src, hipStream_t st) {
  size_t n = dst.size();
  if (!n) return;
  hipLaunchKernelGGL(cast_kernel_up, dim3((n + 255) / 256), dim3(256), 0, st, dst.p, dst.ld, (const cpx<float>*)src.p,
                     src.ld, dst.r, dst.c);
}

// ---------------------------------------------------------------- RRQR
struct RRQRState {
  int m, n, partition, fillInRank, ldv;
  int j, phase, stopped, rank1, first, steps, rank2, nrefl, apply, apply_j, error;
  double tol1, tol2, inputPivot, max1, maxPivot;
};

template <typename T>
__global__ void __launch_bounds__(256) rrqr_init_kernel(const T* A, int lda, int m, int n, double* ns) {
  const int c = blockIdx.x;
  double s = 0;
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    T v = A[(size_t)i + (size_t)c * lda];
    s += (double)v.re * v.re + (double)v.im * v.im;
  }
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) ns[c] = red[0];
}

__device__ inline void rrqr_setup_phase2(RRQRState& s, double* ns) {
  // ns[i] = 0 for i in [rank1, partition) is done by the caller's threads
  s.steps = s.steps - (s.partition - s.rank1);
  s.first = s.rank1;
  s.rank2 = min(s.steps - s.first, s.m);
  if (s.rank2 < 0) s.rank2 = 0;
  s.j = s.first;
  s.phase = 2;
}

// Choose the pivot column, swap it into position j, build the Householder vector and
// decide whether the step is accepted (src/rrqr.cpp:92-130 and 138-176).
template <typename T>
__global__ void __launch_bounds__(256) rrqr_pivot_kernel(T* A, int lda, double* ns, T* V, RRQRState* gs) {
  using R = typename RealOf<T>::type;
  __shared__ RRQRState s;
  __shared__ double rv[256];
  __shared__ int ri[256];
  __shared__ double rr[256], ri2[256];
  const int tid = threadIdx.x;
  if (tid == 0) { s = *gs; s.apply = 0; }
  __syncthreads();
  if (s.stopped) { if (tid == 0) gs->apply = 0; return; }
  if (s.phase == 1 && s.j >= s.rank1) {
    for (int i = max(s.rank1, 0) + tid; i < s.partition && i < s.n; i += blockDim.x) ns[i] = 0.0;
    __syncthreads();
    if (tid == 0) rrqr_setup_phase2(s, ns);
    __syncthreads();
  }
  if (s.phase == 2 && s.j >= s.steps) {
    if (tid == 0) { s.stopped = 1; *gs = s; }
    return;
  }
  const int j = s.j, m = s.m;
  const int end = (s.phase == 1) ? s.partition : s.n;
  // first maximal downdated norm over [j, end)
  double best = -1.0;
  int bi = end;
  for (int c = j + tid; c < end; c += blockDim.x) {
    double v = ns[c];
    if (v > best) { best = v; bi = c; }
  }
  rv[tid] = best;
  ri[tid] = bi;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (tid < k) {
      if (rv[tid + k] > rv[tid] || (rv[tid + k] == rv[tid] && ri[tid + k] < ri[tid])) {
        rv[tid] = rv[tid + k];
        ri[tid] = ri[tid + k];
      }
    }
    __syncthreads();
  }
  int p = ri[0];
  if (p >= end) p = j;
  if (p != j) {
    for (int i = tid; i < m; i += blockDim.x) {
      T t = A[(size_t)i + (size_t)p * lda];
      A[(size_t)i + (size_t)p * lda] = A[(size_t)i + (size_t)j * lda];
      A[(size_t)i + (size_t)j * lda] = t;
    }
    __syncthreads();
    if (tid == 0) { double t = ns[p]; ns[p] = ns[j]; ns[j] = t; }
  }
  __syncthreads();
  // Householder vector from x = A[j:m, j] (src/rrqr.cpp:46-67)
  const T* x = A + (size_t)j * lda;
  double sn = 0;
  for (int i = j + tid; i < m; i += blockDim.x) sn += (double)x[i].re * x[i].re + (double)x[i].im * x[i].im;
  rr[tid] = sn;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (tid < k) rr[tid] += rr[tid + k];
    __syncthreads();
  }
  const double xn = sqrt(rr[0]);
  const T xj = x[j];
  const double ang = atan2((double)xj.im, (double)xj.re);
  const T v0 = T((R)((double)xj.re + cos(ang) * xn), (R)((double)xj.im + sin(ang) * xn));
  // ||v||^2 = ||x||^2 - |xj|^2 + |v0|^2
  const double vn2 = rr[0] - ((double)xj.re * xj.re + (double)xj.im * xj.im) + ((double)v0.re * v0.re + (double)v0.im * v0.im);
  const double scale = vn2 > 0 ? sqrt(2.0) / sqrt(vn2) : 1.0;
  T* v = V + (size_t)s.nrefl * s.ldv;
  for (int i = tid; i < m; i += blockDim.x) {
    T vi = (i < j) ? T(0, 0) : (i == j ? v0 : x[i]);
    if (vn2 > 0) vi = vi * (R)scale;
    v[i] = vi;
  }
  __syncthreads();
  // diag after reflection: x_j - v_j (v^H x)
  double wr = 0, wi = 0;
  for (int i = j + tid; i < m; i += blockDim.x) {
    T vi = v[i], xi = x[i];
    wr += (double)vi.re * xi.re + (double)vi.im * xi.im;
    wi += (double)vi.re * xi.im - (double)vi.im * xi.re;
  }
  rr[tid] = wr;
  ri2[tid] = wi;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (tid < k) { rr[tid] += rr[tid + k]; ri2[tid] += ri2[tid + k]; }
    __syncthreads();
  }
  if (tid == 0) {
    const T vj = v[j];
    double dr = (double)xj.re - ((double)vj.re * rr[0] - (double)vj.im * ri2[0]);
    double di = (double)xj.im - ((double)vj.re * ri2[0] + (double)vj.im * rr[0]);
    double pivot = sqrt(dr * dr + di * di);
    if (s.phase == 1) {
      if (j == 0) s.max1 = pivot;
      s.maxPivot = s.max1;
      if (!(pivot / s.max1 >= s.tol1)) {  // pivot/max1 < tol1 (NaN counts as below)
        s.rank1 = j - 1;
        s.apply = 0;
        s.phase = 3;  // phase-2 set-up pending in the next launch
      } else {
        s.apply = 1;
        s.apply_j = j;
        s.nrefl++;
        s.j = j + 1;
      }
    } else {
      if (j == s.first) s.maxPivot = fmax(s.inputPivot, pivot);
      if (!(pivot / s.maxPivot >= s.tol2) || j - s.first >= s.fillInRank) {
        s.rank2 = j - s.first;
        s.stopped = 1;
        s.apply = 0;
      } else {
        s.apply = 1;
        s.apply_j = j;
        s.nrefl++;
        s.j = j + 1;
      }
    }
    *gs = s;
  }
}

// Early stop of the first partition: zero the remaining norms and restart at rank1.
__global__ void rrqr_phase3_kernel(double* ns, RRQRState* gs) {
  __shared__ RRQRState s;
  if (threadIdx.x == 0) s = *gs;
  __syncthreads();
  if (s.stopped || s.phase != 3) return;
  for (int i = max(s.rank1, 0) + threadIdx.x; i < s.partition && i < s.n; i += blockDim.x) ns[i] = 0.0;
  __syncthreads();
  if (threadIdx.x == 0) {
    if (s.rank1 < 0) s.rank1 = 0;
    rrqr_setup_phase2(s, ns);
    *gs = s;
  }
}

// Apply the accepted reflector to columns j..n-1 and downdate the column norms.
template <typename T>
__global__ void __launch_bounds__(256) rrqr_apply_kernel(T* A, int lda, double* ns, const T* V, const RRQRState* gs) {
  using R = typename RealOf<T>::type;
  const RRQRState* s = gs;
  if (!s->apply) return;
  const int j = s->apply_j, m = s->m, n = s->n;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int c = j + blockIdx.x * 4 + wave;
  if (c >= n) return;
  const T* v = V + (size_t)(s->nrefl - 1) * s->ldv;
  T* a = A + (size_t)c * lda;
  double wr = 0, wi = 0;
  for (int i = j + lane; i < m; i += 64) {
    T vi = v[i], ai = a[i];
    wr += (double)vi.re * ai.re + (double)vi.im * ai.im;
    wi += (double)vi.re * ai.im - (double)vi.im * ai.re;
  }
  for (int off = 32; off > 0; off >>= 1) {
    wr += __shfl_xor(wr, off);
    wi += __shfl_xor(wi, off);
  }
  const T w((R)wr, (R)wi);
  for (int i = j + lane; i < m; i += 64) a[i] = a[i] - v[i] * w;
  if (c == j) {
    for (int i = j + 1 + lane; i < m; i += 64) a[i] = T(0, 0);
  } else if (lane == 0) {
    double ar = a[j].re, ai = a[j].im;
    double ab = sqrt(ar * ar + ai * ai);
    ns[c] = ns[c] - ab * ab;
  }
}

// Q(:, c) = H_0 H_1 ... H_{nrefl-1} e_c, one workgroup per column.
template <typename T>
__global__ void __launch_bounds__(256) rrqr_formq_kernel(T* Q, int ldq, const T* V, const RRQRState* gs) {
  using R = typename RealOf<T>::type;
  const int c = blockIdx.x;
  const int m = gs->m, nrefl = gs->nrefl, ldv = gs->ldv;
  T* q = Q + (size_t)c * ldq;
  for (int i = threadIdx.x; i < m; i += blockDim.x) q[i] = (i == c) ? T(1, 0) : T(0, 0);
  __syncthreads();
  __shared__ double rr[256], ri[256];
  for (int sidx = nrefl - 1; sidx >= 0; --sidx) {
    const T* v = V + (size_t)sidx * ldv;
    double wr = 0, wi = 0;
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
      T vi = v[i], qi = q[i];
      wr += (double)vi.re * qi.re + (double)vi.im * qi.im;
      wi += (double)vi.re * qi.im - (double)vi.im * qi.re;
    }
    rr[threadIdx.x] = wr;
    ri[threadIdx.x] = wi;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
      if (threadIdx.x < k) { rr[threadIdx.x] += rr[threadIdx.x + k]; ri[threadIdx.x] += ri[threadIdx.x + k]; }
      __syncthreads();
    }
    const T w((R)rr[0], (R)ri[0]);
    for (int i = threadIdx.x; i < m; i += blockDim.x) q[i] = q[i] - v[i] * w;
    __syncthreads();
  }
}

size_t rrqr_work_bytes(int m, int n, int elem_bytes) {
  size_t steps = (size_t)std::min(m, n) + 2;
  return 256 + sizeof(double) * (size_t)n + (size_t)elem_bytes * (size_t)m * steps + 256;
}

template <typename T>
RRQRInfo rrqr(DMat<T> A, int partition, double tol1, double tol2, double inputPivot, int fillInRank, DMat<T> Q,
              Ctx& cx) {
  hipStream_t st = cx.st;
  const int m = A.r, n = A.c;
  RRQRInfo info{0, 0, 0, 0.0};
  if (m == 0 || n == 0) return info;
  const size_t work_bytes = rrqr_work_bytes(m, n, sizeof(T));
  char* w = (char*)cx.pool.alloc(work_bytes);
  RRQRState* d_state = (RRQRState*)w;
  double* ns = (double*)(w + 256);
  T* V = (T*)(w + 256 + ((sizeof(double) * (size_t)n + 255) / 256) * 256);
  RRQRState s{};
  s.m = m; s.n = n; s.partition = partition; s.fillInRank = fillInRank; s.ldv = m;
  s.j = 0; s.phase = 1; s.stopped = 0;
  s.rank1 = std::min(partition, m);
  s.steps = std::min(n, m);
  s.first = 0; s.rank2 = 0; s.nrefl = 0; s.apply = 0; s.apply_j = 0; s.error = 0;
  s.tol1 = tol1; s.tol2 = tol2; s.inputPivot = inputPivot; s.max1 = 0; s.maxPivot = 0;
  HIPCHK(hipMemcpyAsync(d_state, &s, sizeof(s), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(rrqr_init_kernel<T>, dim3(n), dim3(256), 0, st, (const T*)A.p, A.ld, m, n, ns);
  const int budget = std::min(partition, m) + std::min(fillInRank, std::min(m, n)) + 3;
  const int cgrid = (n + 3) / 4;
  for (int it = 0; it < budget; ++it) {
    hipLaunchKernelGGL(rrqr_pivot_kernel<T>, dim3(1), dim3(256), 0, st, A.p, A.ld, ns, V, d_state);
    hipLaunchKernelGGL(rrqr_phase3_kernel, dim3(1), dim3(256), 0, st, ns, d_state);
    hipLaunchKernelGGL(rrqr_apply_kernel<T>, dim3(cgrid), dim3(256), 0, st, A.p, A.ld, ns, (const T*)V,
                       (const RRQRState*)d_state);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(cx.h_scalars, d_state, sizeof(s), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  std::memcpy(&s, cx.h_scalars, sizeof(s));
  if (!s.stopped) {
    // the loop of the second partition ran to its end (rank2 keeps its default)
    if (s.phase != 2 || s.j < s.steps) throw std::runtime_error("rrqr: step budget exhausted");
  }
  info.rank1 = s.rank1;
  info.rank2 = s.rank2;
  info.rank = s.rank1 + s.rank2;
  info.maxPivot = s.maxPivot;
  if (info.rank > Q.c) throw std::runtime_error("rrqr: Q buffer too small");
  if (info.rank > 0)
    hipLaunchKernelGGL(rrqr_formq_kernel<T>, dim3(info.rank), dim3(256), 0, st, Q.p, Q.ld, (const T*)V,
                       (const RRQRState*)d_state);
  HIPCHK(hipGetLastError());
  cx.pool.free(w);  // stream-ordered: later users of the block are queued after formq
  return info;
}

// ---------------------------------------------------------------- instantiations
#define RCS_INST(T)                                                                                   \
  template void gemm_grouped<T>(std::vector<GemmOp<T>>&, Ctx&);                                      \
  template void getrf<T>(DMat<T>, int*, Ctx&);                                                        \
  template void getrs<T>(DMat<T>, const int*, DMat<T>, Ctx&);                                         \
  template void copy2d<T>(DMat<T>, DMat<T>, hipStream_t);                                             \
  template void adjoint2d<T>(DMat<T>, DMat<T>, hipStream_t);                                          \
  template void fill_identity<T>(DMat<T>, T, hipStream_t);                                            \
  template void zero2d<T>(DMat<T>, hipStream_t);                                                      \
  template void axpy2d<T>(DMat<T>, DMat<T>, T, hipStream_t);                                          \
  template void sqnorms<T>(const std::vector<DMat<T>>&, double*, Ctx&);                               \
  template RRQRInfo rrqr<T>(DMat<T>, int, double, double, double, int, DMat<T>, Ctx&);                \
  template void copy_batched<T>(std::vector<CopyOp<T>>&, Ctx&);                                       \
  template void gather_rows<T>(DMat<T>, DMat<T>, const int*, hipStream_t);                            \
  template void scatter_rows<T>(DMat<T>, DMat<T>, const int*, hipStream_t);
RCS_INST(cpx<double>)
RCS_INST(cpx<float>)

}  // namespace rcs
