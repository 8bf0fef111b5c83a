// Dense complex kernels for gfx950 (CDNA4, 64-wide wavefronts).
#include <algorithm>
#include <cstdlib>

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

// Small groups travel as kernel arguments (no parameter copy to HBM before the launch).
constexpr int GA_OUTS = 8, GA_SEGS = 16;
template <typename T>
struct GemmArgs {
  GemmOut<T> outs[GA_OUTS];
  GemmSeg<T> segs[GA_SEGS];
  int nouts;
};

// Products of an output block over the global inner-index range [klo, khi) (the inner
// dimensions of its segments laid end to end) into acc.
template <typename T>
__device__ __forceinline__ void gemm_acc(const GemmOut<T>& out, const GemmSeg<T>* segs, int m0, int n0, int klo,
                                         int khi, T (&acc)[4][4]) {
  __shared__ T As[GK][GT + 1];
  __shared__ T Bs[GK][GT + 1];
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  int kbase = 0;
  for (int sgi = 0; sgi < out.nseg; ++sgi) {
    const GemmSeg<T> op = segs[out.seg0 + sgi];
    const int lo = max(klo - kbase, 0), hi = min(khi - kbase, op.k);
    kbase += op.k;
    for (int k0 = lo; k0 < hi; k0 += GK) {
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        int i, kk;
        if (op.opA == OP_N) { i = tid & 63; kk = (tid >> 6) + 4 * s; }
        else { kk = tid & 15; i = (tid >> 4) + 16 * s; }
        T v(0, 0);
        if (m0 + i < out.m && k0 + kk < hi) v = ld_op(op.A, op.lda, m0 + i, k0 + kk, op.opA);
        As[kk][i] = v;
        int j2, k2;
        if (op.opB == OP_N) { k2 = tid & 15; j2 = (tid >> 4) + 16 * s; }
        else { j2 = tid & 63; k2 = (tid >> 6) + 4 * s; }
        T w(0, 0);
        if (n0 + j2 < out.n && k0 + k2 < hi) w = ld_op(op.B, op.ldb, k0 + k2, n0 + j2, op.opB);
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
}

template <typename T>
__device__ __forceinline__ int find_out(const GemmOut<T>* outs, int nouts, int tile) {
  int lo = 0, hi = nouts - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (outs[mid].tile0 <= tile) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// C = alpha * acc + beta * C for the tile at (m0, n0).
template <typename T>
__device__ __forceinline__ void gemm_store(const GemmOut<T>& out, int m0, int n0, const T (&acc)[4][4]) {
  using R = typename RealOf<T>::type;
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
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
__device__ __forceinline__ void gemm_tile(const GemmOut<T>* outs, int nouts, const GemmSeg<T>* segs) {
  const int tile = blockIdx.x;
  const GemmOut<T> out = outs[find_out(outs, nouts, tile)];
  const int t = tile - out.tile0;
  const int tm = (out.m + GT - 1) / GT;
  const int m0 = (t % tm) * GT, n0 = (t / tm) * GT;
  T acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = T(0, 0);
  gemm_acc(out, segs, m0, n0, 0, 0x7fffffff, acc);
  gemm_store(out, m0, n0, acc);
}

// ROLE only separates launch sites in kernel traces (1: Schur-complement updates of the
// elimination); the code is the same.
template <typename T, int ROLE>
__global__ void __launch_bounds__(256) gemm_grouped_kernel(const GemmOut<T>* outs, int nouts,
                                                           const GemmSeg<T>* segs) {
  gemm_tile(outs, nouts, segs);
}
template <typename T, int ROLE>
__global__ void __launch_bounds__(256) gemm_small_kernel(GemmArgs<T> a) {
  gemm_tile(a.outs, a.nouts, a.segs);
}

// Split inner dimension for groups with few output tiles: workgroup (tile, part) sums the
// products over its slice of the inner index into part's 64x64 slot of `partial`, then
// counts its arrival on the tile's counter; the workgroup that arrives last adds the parts
// of the tile in index order (the same sum whichever arrives last), applies alpha and beta,
// and resets the counter for the next launch.
template <typename T>
__global__ void __launch_bounds__(256) gemm_splitk_kernel(GemmArgs<T> a, int parts, T* partial, unsigned* counters) {
  const int tile = blockIdx.x / parts, part = blockIdx.x % parts;
  const GemmOut<T> out = a.outs[find_out(a.outs, a.nouts, tile)];
  int ktot = 0;
  for (int sgi = 0; sgi < out.nseg; ++sgi) ktot += a.segs[out.seg0 + sgi].k;
  const int kc = ((ktot + parts - 1) / parts + GK - 1) / GK * GK;
  const int t = tile - out.tile0, tm = (out.m + GT - 1) / GT;
  const int m0 = (t % tm) * GT, n0 = (t / tm) * GT;
  T acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = T(0, 0);
  gemm_acc(out, a.segs, m0, n0, part * kc, min((part + 1) * kc, ktot), acc);
  T* slot = partial + (size_t)blockIdx.x * GT * GT;
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) slot[(ty + 16 * i) + (tx + 16 * j) * GT] = acc[i][j];
  __threadfence();  // this part is visible device-wide before its arrival is counted
  __shared__ unsigned last;
  __syncthreads();
  if (tid == 0) last = (atomicAdd(&counters[tile], 1u) == (unsigned)parts - 1u) ? 1u : 0u;
  __syncthreads();
  if (!last) return;
  __threadfence();
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = T(0, 0);
  for (int p = 0; p < parts; ++p) {
    const T* sp = partial + ((size_t)tile * parts + p) * GT * GT;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = acc[i][j] + sp[(ty + 16 * i) + (tx + 16 * j) * GT];
  }
  gemm_store(out, m0, n0, acc);
  if (tid == 0) counters[tile] = 0u;
}

// Narrow products (at most NARROW_N output columns, op(A) = A, op(B) = B), as in solves with
// few right-hand sides: one output row per lane, the four wavefronts of a workgroup split
// the inner index and add their parts in LDS in a fixed order. A is read once, coalesced.
constexpr int NARROW_N = 8, NA_OUTS = 32, NA_SEGS = 32;
template <typename T>
struct NarrowArgs {
  GemmOut<T> outs[NA_OUTS];  // tile0: first 64-row block of the output
  GemmSeg<T> segs[NA_SEGS];
  int nouts;
};
template <typename T>
__global__ void __launch_bounds__(256) gemm_narrow_kernel(NarrowArgs<T> a) {
  const int blk = blockIdx.x;
  int o = 0;
  for (int q = 1; q < a.nouts; ++q)
    if (a.outs[q].tile0 <= blk) o = q;
  const GemmOut<T>& out = a.outs[o];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = (blk - out.tile0) * 64 + lane, n = out.n;
  T acc[NARROW_N];
#pragma unroll
  for (int c = 0; c < NARROW_N; ++c) acc[c] = T(0, 0);
  for (int sgi = 0; sgi < out.nseg; ++sgi) {
    const GemmSeg<T>& sg = a.segs[out.seg0 + sgi];
    if (r < out.m) {
      const T* arow = sg.A + r;
#pragma unroll 4
      for (int k = wave; k < sg.k; k += 4) {
        const T av = arow[(size_t)k * sg.lda];
#pragma unroll
        for (int c = 0; c < NARROW_N; ++c)
          if (c < n) cfma(acc[c], av, sg.B[k + (size_t)c * sg.ldb]);
      }
    }
  }
  __shared__ T part[3][NARROW_N][64];
  if (wave > 0) {
#pragma unroll
    for (int c = 0; c < NARROW_N; ++c) part[wave - 1][c][lane] = acc[c];
  }
  __syncthreads();
  if (wave == 0 && r < out.m) {
    using R = typename RealOf<T>::type;
    const bool zb = (out.beta.re == R(0) && out.beta.im == R(0));
#pragma unroll
    for (int c = 0; c < NARROW_N; ++c) {
      if (c >= n) continue;
      const T v0 = ((acc[c] + part[0][c][lane]) + part[1][c][lane]) + part[2][c][lane];
      T* cp = out.C + (size_t)r + (size_t)c * out.ldc;
      T v = out.alpha * v0;
      if (!zb) v = v + out.beta * (*cp);
      *cp = v;
    }
  }
}

constexpr int SPLITK_TILES = 96;    // groups with fewer output tiles than this split the inner index
constexpr int SPLITK_MIN_K = 64;    // inner index per part, at least

template <typename T>
void gemm_grouped(std::vector<GemmOp<T>>& ops, Ctx& cx, int role, int prof_tag) {
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
  static const bool no_narrow = std::getenv("RCS_NO_NARROW") != nullptr;  // debugging: tile kernel only
  bool narrow = !no_narrow && role == 0 && outs.size() <= (size_t)NA_OUTS && segs.size() <= (size_t)NA_SEGS;
  for (const auto& h : outs) narrow = narrow && h.n <= NARROW_N;
  for (const auto& sg : segs) narrow = narrow && sg.opA == OP_N && sg.opB == OP_N;
  if (narrow) {
    static_assert(sizeof(NarrowArgs<T>) <= 4000, "narrow GEMM arguments exceed the kernel argument segment");
    NarrowArgs<T> na;
    int blocks = 0;
    for (size_t i = 0; i < outs.size(); ++i) {
      na.outs[i] = outs[i];
      na.outs[i].tile0 = blocks;
      blocks += (outs[i].m + 63) / 64;
    }
    std::copy(segs.begin(), segs.end(), na.segs);
    na.nouts = (int)outs.size();
    const int h = prof_tag >= 0 ? cx.prof.begin(cx.st) : -1;
    hipLaunchKernelGGL((gemm_narrow_kernel<T>), dim3(blocks), dim3(256), 0, cx.st, na);
    if (h >= 0) {
      double fl = 0;
      for (const auto& o : ops) fl += 8.0 * o.m * (double)o.n * std::max(o.k, 0);
      cx.prof.end(cx.st, h, prof_tag, fl);
    }
    HIPCHK(hipGetLastError());
    return;
  }
  // (Schur updates keep one kernel name in traces: always the HBM parameter tables)
  const bool small = role == 0 && outs.size() <= (size_t)GA_OUTS && segs.size() <= (size_t)GA_SEGS;
  GemmArgs<T> args;
  const GemmOut<T>* d_outs = nullptr;
  const GemmSeg<T>* d_segs = nullptr;
  int parts = 1;
  if (small) {
    std::copy(outs.begin(), outs.end(), args.outs);
    std::copy(segs.begin(), segs.end(), args.segs);
    args.nouts = (int)outs.size();
    if (tiles < SPLITK_TILES) {
      int kmax = 0;
      for (const auto& h : outs) {
        int kt = 0;
        for (int i = 0; i < h.nseg; ++i) kt += segs[h.seg0 + i].k;
        kmax = std::max(kmax, kt);
      }
      parts = std::max(1, std::min((2 * SPLITK_TILES + tiles - 1) / tiles, kmax / SPLITK_MIN_K));
      static_assert(SPLITK_TILES <= Ctx::n_tile_counters, "split GEMM tiles exceed the counter array");
    }
  } else {
    d_outs = (const GemmOut<T>*)cx.push(outs.data(), sizeof(GemmOut<T>) * outs.size());
    d_segs = (const GemmSeg<T>*)cx.push(segs.data(), sizeof(GemmSeg<T>) * segs.size());
  }
  T* partial = parts > 1 ? (T*)cx.pool.alloc(sizeof(T) * (size_t)tiles * parts * GT * GT) : nullptr;
  // the profiled interval holds the kernel alone (the parameter copies come before it)
  const int h = prof_tag >= 0 ? cx.prof.begin(cx.st) : -1;
  if (parts > 1) {
    hipLaunchKernelGGL((gemm_splitk_kernel<T>), dim3(tiles * parts), dim3(256), 0, cx.st, args, parts, partial,
                       cx.d_tile_counters);
  } else if (small) {
    hipLaunchKernelGGL((gemm_small_kernel<T, 0>), dim3(tiles), dim3(256), 0, cx.st, args);
  } else if (role == 1) {
    hipLaunchKernelGGL((gemm_grouped_kernel<T, 1>), dim3(tiles), dim3(256), 0, cx.st, d_outs, (int)outs.size(), d_segs);
  } else {
    hipLaunchKernelGGL((gemm_grouped_kernel<T, 0>), dim3(tiles), dim3(256), 0, cx.st, d_outs, (int)outs.size(), d_segs);
  }
  if (h >= 0) {
    double fl = 0;
    for (const auto& o : ops) fl += 8.0 * o.m * (double)o.n * std::max(o.k, 0);
    cx.prof.end(cx.st, h, prof_tag, fl);
  }
  if (partial) cx.pool.free(partial);  // stream ordered: later users of the block run after the reduce
  HIPCHK(hipGetLastError());
}

// ---------------------------------------------------------------- LU factorization
constexpr int LU_NB = 32;

// Factor the panel A[j:n, j:j+nb] (one workgroup). Pivot = first row of maximal |a|
// (largest value, lowest row index), as Eigen's PartialPivLU picks it.
template <typename T>
__global__ void __launch_bounds__(1024) lu_panel_kernel(T* A, int lda, int n, int j, int nb, int* ipiv) {
  using R = typename RealOf<T>::type;
  __shared__ R sval[16];
  __shared__ int sidx[16];
  __shared__ T urow[LU_NB];
  const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wave = tid >> 6, nw = nt >> 6;
  for (int jj = 0; jj < nb; ++jj) {
    const int c = j + jj;
    T* colc = A + (size_t)c * lda;
    R best = R(-1);
    int bi = n;
    for (int r = c + tid; r < n; r += nt) {
      R v = dabs(colc[r]);
      if (v > best) { best = v; bi = r; }
    }
    for (int off = 32; off > 0; off >>= 1) {
      R ov = __shfl_xor(best, off);
      int oi = __shfl_xor(bi, off);
      if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
    }
    if (lane == 0) { sval[wave] = best; sidx[wave] = bi; }
    __syncthreads();
    best = sval[0];
    bi = sidx[0];
    for (int w = 1; w < nw; ++w)
      if (sval[w] > best || (sval[w] == best && sidx[w] < bi)) { best = sval[w]; bi = sidx[w]; }
    const int p = (bi >= n) ? c : bi;
    if (tid == 0) ipiv[c] = p;
    // swap rows c and p inside the panel; keep row c (the new U row) in LDS
    if (tid < nb) {
      T* cq = A + (size_t)(j + tid) * lda;
      T a = cq[c], b = cq[p];
      if (p != c) { cq[c] = b; cq[p] = a; }
      urow[tid] = b;  // b is row p's entry, which is row c's when p == c (no select: a select
                      // of the two complex values was built in scratch memory)
    }
    __syncthreads();
    const T piv = urow[jj];
    const bool nz = !(piv.re == R(0) && piv.im == R(0));
    // scale column c below the diagonal and update the remaining panel columns: thread
    // owns rows r = c+1+tid+nt*k; the l factor is computed once per row
    const int cols = nb - jj - 1;
    for (int r = c + 1 + tid; r < n; r += nt) {
      T l = colc[r];
      if (nz) { l = cdiv(l, piv); colc[r] = l; }
      for (int q = 0; q < cols; ++q) {
        T* cq = A + (size_t)(c + 1 + q) * lda;
        cq[r] = cq[r] - l * urow[jj + 1 + q];
      }
    }
    __syncthreads();
  }
}

// The same panel factorization with the panel in registers: thread t owns row j + t (all nb
// <= LU_NB panel entries of it), so the pivot steps touch no global memory. The pivot row
// and row c are exchanged through LDS, which also broadcasts the new U row. Needs
// n - j <= NT; complex128 rows take 128 VGPRs (NT = 512), complex64 rows 64 (NT = 1024).
// The steps are instantiated one by one (JJ a template argument) so that every index into
// the row array is a constant and the array stays in registers.
template <typename T, int NT>
struct LuRegPanel {
  using R = typename RealOf<T>::type;
  R* sval;
  int* sidx;
  T* urow;
  T* crow;
  int* ipiv;
  int n, j, nb, r, tid;
  bool own;
  template <int JJ>
  __device__ __forceinline__ void step(T (&a)[LU_NB]) {
    if (JJ < nb) {
      const int lane = tid & 63, wave = tid >> 6, c = j + JJ;
      R best = (own && r >= c) ? dabs(a[JJ]) : R(-1);
      int bi = (own && r >= c) ? r : n;
      for (int off = 32; off > 0; off >>= 1) {
        R ov = __shfl_xor(best, off);
        int oi = __shfl_xor(bi, off);
        if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
      }
      if (lane == 0) { sval[wave] = best; sidx[wave] = bi; }
      __syncthreads();
      best = sval[0];
      bi = sidx[0];
#pragma unroll
      for (int w = 1; w < NT / 64; ++w)
        if (sval[w] > best || (sval[w] == best && sidx[w] < bi)) { best = sval[w]; bi = sidx[w]; }
      const int p = (bi >= n) ? c : bi;
      if (tid == 0) ipiv[c] = p;
      if (r == p) {
#pragma unroll
        for (int q = 0; q < LU_NB; ++q) urow[q] = a[q];
      }
      if (r == c) {
#pragma unroll
        for (int q = 0; q < LU_NB; ++q) crow[q] = a[q];
      }
      __syncthreads();
      if (p != c) {
        if (r == c) {
#pragma unroll
          for (int q = 0; q < LU_NB; ++q) a[q] = urow[q];
        } else if (r == p) {
#pragma unroll
          for (int q = 0; q < LU_NB; ++q) a[q] = crow[q];
        }
      }
      const T piv = urow[JJ];
      const bool nz = !(piv.re == R(0) && piv.im == R(0));
      if (own && r > c) {
        T l = a[JJ];
        if (nz) l = cdiv(l, piv);
        a[JJ] = l;
#pragma unroll
        for (int q = JJ + 1; q < LU_NB; ++q) a[q] = a[q] - l * urow[q];
      }
      __syncthreads();
    }
    if constexpr (JJ + 1 < LU_NB) step<JJ + 1>(a);
  }
};

template <typename T, int NT>
__global__ void __launch_bounds__(NT) lu_panel_reg_kernel(T* A, int lda, int n, int j, int nb, int* ipiv) {
  using R = typename RealOf<T>::type;
  __shared__ R sval[NT / 64];
  __shared__ int sidx[NT / 64];
  __shared__ T urow[LU_NB];  // the pivot row: row c after the exchange (its U part)
  __shared__ T crow[LU_NB];  // row c before the exchange
  LuRegPanel<T, NT> P{sval, sidx, urow, crow, ipiv, n, j, nb, j + (int)threadIdx.x, (int)threadIdx.x,
                      j + (int)threadIdx.x < n};
  T a[LU_NB];
#pragma unroll
  for (int q = 0; q < LU_NB; ++q) a[q] = (P.own && q < nb) ? A[(size_t)(j + q) * lda + P.r] : T(0, 0);
  P.template step<0>(a);
#pragma unroll
  for (int q = 0; q < LU_NB; ++q)
    if (P.own && q < nb) A[(size_t)(j + q) * lda + P.r] = a[q];
}
// register-panel thread count per element type
template <typename T> struct LuRegNT;
template <> struct LuRegNT<cpx<double>> { static constexpr int v = 512; };
template <> struct LuRegNT<cpx<float>> { static constexpr int v = 1024; };

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

// perm = the row swaps ipiv[0..n) applied in order to the identity (one workgroup; the
// swaps are sequential by definition, done in LDS when they fit).
constexpr int PERM_LDS = 8192;
__global__ void __launch_bounds__(256) perm_from_ipiv_kernel(const int* ipiv, int* perm, int n) {
  __shared__ int pl[PERM_LDS];
  __shared__ int il[PERM_LDS];
  if (n <= PERM_LDS) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) { pl[i] = i; il[i] = ipiv[i]; }
    __syncthreads();
    if (threadIdx.x == 0)
      for (int r = 0; r < n; ++r) {
        int p = il[r];
        if (p != r) { int t = pl[r]; pl[r] = pl[p]; pl[p] = t; }
      }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) perm[i] = pl[i];
  } else {
    for (int i = threadIdx.x; i < n; i += blockDim.x) perm[i] = i;
    __syncthreads();
    if (threadIdx.x == 0)
      for (int r = 0; r < n; ++r) {
        int p = ipiv[r];
        if (p != r) { int t = perm[r]; perm[r] = perm[p]; perm[p] = t; }
      }
  }
}

// Triangular solves with the nb x nb diagonal block (nb <= 32) of an LU factor, for many
// right-hand-side columns: 32 lanes per column (lane i owns row i), 8 columns per
// 256-thread workgroup; the solved component of step k reaches the other lanes by a
// shuffle, and the diagonal block is staged in LDS once per workgroup.
constexpr int TRSM_COLS = 8;

template <typename T>
__device__ inline T shfl32(T v, int src) {
  T r;
  r.re = __shfl(v.re, src, 32);
  r.im = __shfl(v.im, src, 32);
  return r;
}

// B[j:j+nb, c] <- L11^{-1} B[j:j+nb, c] (unit lower L11 = A[j:j+nb, j:j+nb]).
template <typename T>
__global__ void __launch_bounds__(256) trsm_lower_unit_kernel(const T* L, int ldl, int j, int nb, T* B, int ldb,
                                                              int ncols) {
  __shared__ T Ls[LU_NB][LU_NB + 1];  // Ls[k][i] = L(j+i, j+k)
  const int tid = threadIdx.x, i = tid & 31;
  for (int e = tid; e < nb * nb; e += 256) Ls[e / nb][e % nb] = L[(size_t)(j + e / nb) * ldl + j + e % nb];
  __syncthreads();
  const int c = blockIdx.x * TRSM_COLS + (tid >> 5);
  const bool live = c < ncols && i < nb;
  T* b = B + (size_t)c * ldb + j;
  T x = live ? b[i] : T(0, 0);
  for (int k = 0; k < nb - 1; ++k) {
    const T xk = shfl32(x, k);
    if (i > k) x = x - Ls[k][i] * xk;
  }
  if (live) b[i] = x;
}

// B[j:j+nb, c] <- U11^{-1} B[j:j+nb, c] (upper, non-unit).
template <typename T>
__global__ void __launch_bounds__(256) trsm_upper_kernel(const T* U, int ldu, int j, int nb, T* B, int ldb,
                                                         int ncols) {
  __shared__ T Us[LU_NB][LU_NB + 1];  // Us[k][i] = U(j+i, j+k)
  const int tid = threadIdx.x, i = tid & 31;
  for (int e = tid; e < nb * nb; e += 256) Us[e / nb][e % nb] = U[(size_t)(j + e / nb) * ldu + j + e % nb];
  __syncthreads();
  const int c = blockIdx.x * TRSM_COLS + (tid >> 5);
  const bool live = c < ncols && i < nb;
  T* b = B + (size_t)c * ldb + j;
  T x = live ? b[i] : T(0, 0);
  for (int k = nb - 1; k >= 0; --k) {
    if (i == k && live) x = cdiv(x, Us[k][k]);
    const T xk = shfl32(x, k);
    if (i < k) x = x - Us[k][i] * xk;
  }
  if (live) b[i] = x;
}

template <typename T>
void getrf(DMat<T> A, int* ipiv, Ctx& cx) {
  hipStream_t st = cx.st;
  const int n = A.r;
  if (n == 0) return;
  if (A.c != n) throw std::runtime_error("getrf: matrix is not square");
  for (int j = 0; j < n; j += LU_NB) {
    const int nb = std::min(LU_NB, n - j);
    constexpr int NT = LuRegNT<T>::v;
    if (n - j <= NT) {
      hipLaunchKernelGGL((lu_panel_reg_kernel<T, NT>), dim3(1), dim3(NT), 0, st, A.p, A.ld, n, j, nb, ipiv);
    } else {
      int threads = 1024;
      if (n - j <= 256) threads = 256;
      else if (n - j <= 512) threads = 512;
      hipLaunchKernelGGL(lu_panel_kernel<T>, dim3(1), dim3(threads), 0, st, A.p, A.ld, n, j, nb, ipiv);
    }
    hipLaunchKernelGGL(laswp_kernel<T>, dim3((n + 127) / 128), dim3(128), 0, st, A.p, A.ld, n, j, nb,
                       (const int*)ipiv, j, j + nb);
    const int rest = n - j - nb;
    if (rest > 0) {
      hipLaunchKernelGGL(trsm_lower_unit_kernel<T>, dim3((rest + TRSM_COLS - 1) / TRSM_COLS), dim3(256), 0, st,
                         (const T*)A.p, A.ld, j, nb, A.p + (size_t)(j + nb) * A.ld, A.ld, rest);
      std::vector<GemmOp<T>> g(1);
      g[0].A = A.p + (size_t)j * A.ld + j + nb; g[0].lda = A.ld; g[0].opA = OP_N;
      g[0].B = A.p + (size_t)(j + nb) * A.ld + j; g[0].ldb = A.ld; g[0].opB = OP_N;
      g[0].C = A.p + (size_t)(j + nb) * A.ld + j + nb; g[0].ldc = A.ld;
      g[0].m = rest; g[0].n = rest; g[0].k = nb;
      g[0].alpha = T(-1, 0); g[0].beta = T(1, 0);
      gemm_grouped(g, cx);
    }
  }
  hipLaunchKernelGGL(perm_from_ipiv_kernel, dim3(1), dim3(256), 0, st, (const int*)ipiv, ipiv + n, n);
  HIPCHK(hipGetLastError());
}

// Few right-hand sides of a system of at most GETRS_LDS unknowns: one workgroup per
// column, the column gathered through the permutation into LDS and solved there.
constexpr int GETRS_LDS = 1024;

// x_r - sum_{k < nb} LU(r, j0 + k) x_{j0 + k}, with sixteen loads of the row issued together.
template <typename T>
__device__ __forceinline__ T getrs_update(const T* LU, int ld, int r, int j0, int nb, const T* xl) {
  T s = xl[r];
  int k = 0;
  for (; k + 16 <= nb; k += 16) {
    T a[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) a[q] = LU[(size_t)(j0 + k + q) * ld + r];
#pragma unroll
    for (int q = 0; q < 16; ++q) s = s - a[q] * xl[j0 + k + q];
  }
  for (; k < nb; ++k) s = s - LU[(size_t)(j0 + k) * ld + r] * xl[j0 + k];
  return s;
}

// Solve with the LU factors for a few right-hand sides (one workgroup per column), in blocks
// of 64 unknowns: the block's diagonal part of L (or U) is staged in LDS by all threads, one
// wavefront solves it (lane i owns unknown i, the solved value is broadcast by a shuffle),
// then all threads update the remaining unknowns with the block's columns (coalesced reads).
template <typename T>
__global__ void __launch_bounds__(256) getrs_vec_kernel(const T* LU, int ld, int n, const int* perm, T* B, int ldb) {
  __shared__ T xl[GETRS_LDS];
  __shared__ T dg[64][65];  // dg[k][i] = LU(j0 + i, j0 + k)
  T* b = B + (size_t)blockIdx.x * ldb;
  const int tid = threadIdx.x;
  for (int i = tid; i < n; i += blockDim.x) xl[i] = b[perm[i]];
  // forward (unit lower)
  for (int j0 = 0; j0 < n; j0 += 64) {
    const int nb = min(64, n - j0);
#pragma unroll 16
    for (int e = tid; e < nb * 64; e += blockDim.x) {
      const int k = e >> 6, i = e & 63;
      dg[k][i] = i < nb ? LU[(size_t)(j0 + k) * ld + j0 + i] : T(0, 0);
    }
    __syncthreads();
    if (tid < 64) {
      T x = tid < nb ? xl[j0 + tid] : T(0, 0);
      for (int k = 0; k < nb; ++k) {
        T xk;
        xk.re = __shfl(x.re, k);
        xk.im = __shfl(x.im, k);
        if (tid > k && tid < nb) x = x - dg[k][tid] * xk;
      }
      if (tid < nb) xl[j0 + tid] = x;
    }
    __syncthreads();
    for (int r = j0 + nb + tid; r < n; r += blockDim.x) xl[r] = getrs_update(LU, ld, r, j0, nb, xl);
    __syncthreads();
  }
  // backward (upper)
  for (int j1 = n; j1 > 0; j1 -= 64) {
    const int j0 = max(0, j1 - 64), nb = j1 - j0;
#pragma unroll 16
    for (int e = tid; e < nb * 64; e += blockDim.x) {
      const int k = e >> 6, i = e & 63;
      dg[k][i] = i < nb ? LU[(size_t)(j0 + k) * ld + j0 + i] : T(0, 0);
    }
    __syncthreads();
    if (tid < 64) {
      T x = tid < nb ? xl[j0 + tid] : T(0, 0);
      for (int k = nb - 1; k >= 0; --k) {
        if (tid == k) x = cdiv(x, dg[k][k]);
        T xk;
        xk.re = __shfl(x.re, k);
        xk.im = __shfl(x.im, k);
        if (tid < k) x = x - dg[k][tid] * xk;
      }
      if (tid < nb) xl[j0 + tid] = x;
    }
    __syncthreads();
    for (int r = tid; r < j0; r += blockDim.x) xl[r] = getrs_update(LU, ld, r, j0, nb, xl);
    __syncthreads();
  }
  for (int i = tid; i < n; i += blockDim.x) b[i] = xl[i];
}

template <typename T>
__global__ void permute_rows_kernel(T* d, int ldd, const T* s, int lds, int r, int c, const int* perm) {
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)r * c) return;
  int i = e % r, j = e / r;
  d[(size_t)i + (size_t)j * ldd] = s[(size_t)perm[i] + (size_t)j * lds];
}

template <typename T>
void getrs(DMat<T> LU, const int* ipiv, DMat<T> B, Ctx& cx) {
  hipStream_t st = cx.st;
  const int n = LU.r;
  if (n == 0 || B.c == 0) return;
  if (B.r != n) throw std::runtime_error("getrs: size mismatch");
  const int* perm = ipiv + n;
  // a few right-hand sides of a small system: one workgroup per column is fastest; large
  // systems take the blocked path below, whose updates spread over the whole chip
  if (B.c <= 8 && n <= GETRS_LDS) {
    hipLaunchKernelGGL(getrs_vec_kernel<T>, dim3(B.c), dim3(256), 0, st, (const T*)LU.p, LU.ld, n, perm, B.p, B.ld);
    HIPCHK(hipGetLastError());
    return;
  }
  const int nc = B.c;
  {  // B <- P B through a gathered copy
    T* tmp = (T*)cx.pool.alloc(sizeof(T) * (size_t)n * nc);
    DMat<T> Tm{tmp, n, nc, n};
    const size_t cnt = (size_t)n * nc;
    hipLaunchKernelGGL(permute_rows_kernel<T>, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, st, tmp, n,
                       (const T*)B.p, B.ld, n, nc, perm);
    copy2d(B, Tm, st);
    cx.pool.free(tmp);
  }
  for (int j = 0; j < n; j += LU_NB) {
    const int nb = std::min(LU_NB, n - j);
    hipLaunchKernelGGL(trsm_lower_unit_kernel<T>, dim3((nc + TRSM_COLS - 1) / TRSM_COLS), dim3(256), 0, st,
                       (const T*)LU.p, LU.ld, j, nb, B.p, B.ld, nc);
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
    hipLaunchKernelGGL(trsm_upper_kernel<T>, dim3((nc + TRSM_COLS - 1) / TRSM_COLS), dim3(256), 0, st,
                       (const T*)LU.p, LU.ld, j, nb, B.p, B.ld, nc);
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

// ---------------------------------------------------------------- instantiations
#define RCS_INST(T)                                                                                   \
  template void gemm_grouped<T>(std::vector<GemmOp<T>>&, Ctx&, int, int);                            \
  template void getrf<T>(DMat<T>, int*, Ctx&);                                                        \
  template void getrs<T>(DMat<T>, const int*, DMat<T>, Ctx&);                                         \
  template void copy2d<T>(DMat<T>, DMat<T>, hipStream_t);                                             \
  template void adjoint2d<T>(DMat<T>, DMat<T>, hipStream_t);                                          \
  template void fill_identity<T>(DMat<T>, T, hipStream_t);                                            \
  template void zero2d<T>(DMat<T>, hipStream_t);                                                      \
  template void axpy2d<T>(DMat<T>, DMat<T>, T, hipStream_t);                                          \
  template void sqnorms<T>(const std::vector<DMat<T>>&, double*, Ctx&);                               \
  template void copy_batched<T>(std::vector<CopyOp<T>>&, Ctx&);                                       \
  template void gather_rows<T>(DMat<T>, DMat<T>, const int*, hipStream_t);                            \
  template void scatter_rows<T>(DMat<T>, DMat<T>, const int*, hipStream_t);
RCS_INST(cpx<double>)
RCS_INST(cpx<float>)

}  // namespace rcs
