// Batched ACA on gfx950: see aca.hpp. The per-step semantics follow src/ACA.cpp:639-859:
// residual long vector at the current short index, first-max pivot over the remaining
// long indices, v = residual / pivot, residual short vector at the pivot, u = that
// vector, Frobenius estimate update (A:750-755), next short pivot = first max of |u|
// over the remaining short indices, stop when |u||v| <= tol*|S| or rank = min(N1, N2).
#include <algorithm>

#include "aca.hpp"
#include "entries_dev.hpp"

namespace rcs {

constexpr int ACA_CHUNK = 1024;  // long entries per workgroup (4 per thread)

template <typename T>
struct AcaBox {
  const int* rows;
  const int* cols;
  int N1, N2, branch;  // branch 1: rows are the short side (N1 < N2)
  int ns, nl, cap, nchunks;
  T* U;  // ns x cap residual short vectors
  T* V;  // nl x cap normalised long vectors
  T* W;  // nl current long vector
  unsigned char* remS;
  unsigned char* remL;
  int* pivS;
  int* pivL;
  double* pmax;  // nchunks
  int* pidx;     // nchunks
  double* pdot;  // nchunks x (cap+1) x 2
  int rank, active, init_done, s_cur, l_cur, retry;
  double normS, uv;
  T piv;
};

template <typename R>
__device__ inline R cabs_(cpx<R> a) { return sizeof(R) == 8 ? (R)hypot((double)a.re, (double)a.im) : (R)hypotf((float)a.re, (float)a.im); }

__device__ inline void argmax_combine(double& v, int& i, double v2, int i2) {
  if (v2 > v || (v2 == v && i2 < i)) { v = v2; i = i2; }
}

// K1: residual long vector + partial argmax over the remaining long indices.
template <typename R, typename T>
__global__ void __launch_bounds__(256) aca_long_kernel(DevGeometry g, int kind, AcaBox<T>* boxes, const int2* work) {
  const int b = work[blockIdx.x].x, chunk = work[blockIdx.x].y;
  AcaBox<T>& bx = boxes[b];
  // a converged box, or one whose U/V capacity is full, takes no step (bursts of steps run
  // between host checks, so the launch lists can be stale)
  if (!bx.active || bx.rank >= bx.cap) return;
  const int rank = bx.rank, s = bx.s_cur, nl = bx.nl;
  __shared__ T us[256];
  __shared__ double rv[256];
  __shared__ int ri[256];
  const int tid = threadIdx.x;
  T val[4];
  int tt[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    tt[e] = chunk * ACA_CHUNK + e * 256 + tid;
    val[e] = T(0, 0);
    if (tt[e] < nl) {
      int i, j;
      if (bx.branch == 1) { i = bx.rows[s]; j = bx.cols[tt[e]]; }
      else { i = bx.rows[tt[e]]; j = bx.cols[s]; }
      cpx<R> v = kernel_entry<R>(g, kind, i, j);
      val[e] = T(v.re, v.im);
    }
  }
  for (int q0 = 0; q0 < rank; q0 += 256) {
    const int nq = min(256, rank - q0);
    __syncthreads();
    if (tid < nq) us[tid] = bx.U[(size_t)(q0 + tid) * bx.ns + s];
    __syncthreads();
    for (int q = 0; q < nq; ++q) {
      const T uq = us[q];
      const T* vq = bx.V + (size_t)(q0 + q) * nl;
#pragma unroll
      for (int e = 0; e < 4; ++e)
        if (tt[e] < nl) val[e] = val[e] - uq * vq[tt[e]];
    }
  }
  double best = -1.0;
  int bi = 0x7fffffff;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    if (tt[e] < nl) {
      bx.W[tt[e]] = val[e];
      if (bx.remL[tt[e]]) {
        double a = (double)cabs_(val[e]);
        argmax_combine(best, bi, a, tt[e]);
      }
    }
  }
  rv[tid] = best;
  ri[tid] = bi;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (tid < k) {
      double v2 = rv[tid + k];
      int i2 = ri[tid + k];
      double v1 = rv[tid];
      int i1 = ri[tid];
      argmax_combine(v1, i1, v2, i2);
      rv[tid] = v1;
      ri[tid] = i1;
    }
    __syncthreads();
  }
  if (tid == 0) {
    bx.pmax[chunk] = rv[0];
    bx.pidx[chunk] = ri[0];
  }
}

// K2: pivot of the long vector (one workgroup per box).
template <typename T>
__global__ void __launch_bounds__(256) aca_pivot_kernel(AcaBox<T>* boxes, const int* ids) {
  AcaBox<T>& bx = boxes[ids[blockIdx.x]];
  if (!bx.active) return;
  if (bx.rank >= bx.cap) {  // capacity full: skip the step (the host grows U/V, then resumes)
    if (threadIdx.x == 0) bx.retry = 1;
    return;
  }
  __shared__ double rv[256];
  __shared__ int ri[256];
  const int tid = threadIdx.x;
  double best = -1.0;
  int bi = 0x7fffffff;
  for (int c = tid; c < bx.nchunks; c += 256) argmax_combine(best, bi, bx.pmax[c], bx.pidx[c]);
  rv[tid] = best;
  ri[tid] = bi;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (tid < k) {
      double v1 = rv[tid];
      int i1 = ri[tid];
      argmax_combine(v1, i1, rv[tid + k], ri[tid + k]);
      rv[tid] = v1;
      ri[tid] = i1;
    }
    __syncthreads();
  }
  if (tid == 0) {
    const int l = ri[0];
    const double pv = rv[0];
    if (!bx.init_done && !(pv > 1e-16)) {  // A:663-669: skip an all-zero first vector
      bx.retry = 1;
      bx.s_cur += 1;
      if (bx.s_cur >= bx.ns) bx.active = 0;
      return;
    }
    if (l < 0 || l >= bx.nl) {  // no remaining long index
      bx.retry = 1;
      bx.active = 0;
      return;
    }
    bx.retry = 0;
    bx.l_cur = l;
    bx.piv = bx.W[l];
    bx.pivS[bx.rank] = bx.s_cur;
    bx.pivL[bx.rank] = l;
    bx.remL[l] = 0;
    bx.remS[bx.s_cur] = 0;
  }
}

// K3: v = w / pivot, stored as column `rank` of V, with partial dots V(:,q)^H v.
template <typename T>
__global__ void __launch_bounds__(256) aca_norm_kernel(AcaBox<T>* boxes, const int2* work) {
  const int b = work[blockIdx.x].x, chunk = work[blockIdx.x].y;
  AcaBox<T>& bx = boxes[b];
  if (bx.retry || !bx.active) return;
  const int rank = bx.rank, nl = bx.nl;
  const T piv = bx.piv;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  __shared__ double wr[4], wi[4];
  T v[4];
  int tt[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    tt[e] = chunk * ACA_CHUNK + e * 256 + tid;
    v[e] = T(0, 0);
    if (tt[e] < nl) {
      v[e] = cdiv(bx.W[tt[e]], piv);
      bx.V[(size_t)rank * nl + tt[e]] = v[e];
    }
  }
  for (int q = 0; q <= rank; ++q) {
    double sr = 0, si = 0;
    if (q < rank) {
      const T* vq = bx.V + (size_t)q * nl;
#pragma unroll
      for (int e = 0; e < 4; ++e)
        if (tt[e] < nl) {
          T a = vq[tt[e]];
          sr += (double)a.re * v[e].re + (double)a.im * v[e].im;
          si += (double)a.re * v[e].im - (double)a.im * v[e].re;
        }
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e)
        if (tt[e] < nl) sr += (double)v[e].re * v[e].re + (double)v[e].im * v[e].im;
    }
    for (int off = 32; off > 0; off >>= 1) {
      sr += __shfl_xor(sr, off);
      si += __shfl_xor(si, off);
    }
    if (lane == 0) { wr[wave] = sr; wi[wave] = si; }
    __syncthreads();
    if (tid == 0) {
      double* pd = bx.pdot + ((size_t)chunk * (bx.cap + 1) + q) * 2;
      pd[0] = (wr[0] + wr[1]) + (wr[2] + wr[3]);
      pd[1] = (wi[0] + wi[1]) + (wi[2] + wi[3]);
    }
    __syncthreads();
  }
}

// K4: residual short vector, dots, Frobenius update, stopping test, next short pivot.
template <typename R, typename T>
__global__ void __launch_bounds__(256) aca_short_kernel(DevGeometry g, int kind, AcaBox<T>* boxes, const int* ids,
                                                        double tol) {
  AcaBox<T>& bx = boxes[ids[blockIdx.x]];
  if (bx.retry || !bx.active) return;
  const int rank = bx.rank, ns = bx.ns, l = bx.l_cur;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  __shared__ T vs[256];
  __shared__ double wr[4], wi[4];
  __shared__ double sdu_r[1], sdu_i[1];
  __shared__ double rv[256];
  __shared__ int ri[256];
  T* u = bx.U + (size_t)rank * ns;
  // residual short vector (kept in HBM: ns can exceed the workgroup)
  for (int t = tid; t < ns; t += 256) {
    int i, j;
    if (bx.branch == 1) { i = bx.rows[t]; j = bx.cols[l]; }
    else { i = bx.rows[l]; j = bx.cols[t]; }
    cpx<R> e = kernel_entry<R>(g, kind, i, j);
    u[t] = T(e.re, e.im);
  }
  __syncthreads();
  for (int q0 = 0; q0 < rank; q0 += 256) {
    const int nq = min(256, rank - q0);
    if (tid < nq) vs[tid] = bx.V[(size_t)(q0 + tid) * bx.nl + l];
    __syncthreads();
    for (int t = tid; t < ns; t += 256) {
      T a = u[t];
      for (int q = 0; q < nq; ++q) a = a - vs[q] * bx.U[(size_t)(q0 + q) * ns + t];
      u[t] = a;
    }
    __syncthreads();
  }
  // dots U(:,q)^H u, q <= rank (q = rank gives |u|^2), combined with the V dots
  double acc = 0;  // sum_q 2|dU_q dV_q|
  double un2 = 0, vn2 = 0;
  for (int q = 0; q <= rank; ++q) {
    double sr = 0, si = 0;
    const T* uq = bx.U + (size_t)q * ns;
    for (int t = tid; t < ns; t += 256) {
      T a = uq[t], c = u[t];
      sr += (double)a.re * c.re + (double)a.im * c.im;
      si += (double)a.re * c.im - (double)a.im * c.re;
    }
    for (int off = 32; off > 0; off >>= 1) {
      sr += __shfl_xor(sr, off);
      si += __shfl_xor(si, off);
    }
    if (lane == 0) { wr[wave] = sr; wi[wave] = si; }
    __syncthreads();
    if (tid == 0) {
      double dur = (wr[0] + wr[1]) + (wr[2] + wr[3]);
      double dui = (wi[0] + wi[1]) + (wi[2] + wi[3]);
      double dvr = 0, dvi = 0;
      for (int c = 0; c < bx.nchunks; ++c) {
        const double* pd = bx.pdot + ((size_t)c * (bx.cap + 1) + q) * 2;
        dvr += pd[0];
        dvi += pd[1];
      }
      if (q == rank) { un2 = dur; vn2 = dvr; }
      double pr = dur * dvr - dui * dvi, pi = dur * dvi + dui * dvr;
      acc += 2.0 * sqrt(pr * pr + pi * pi);
      sdu_r[0] = dur;
      sdu_i[0] = dui;
    }
    __syncthreads();
  }
  // next short pivot: first max |u| over the remaining short indices
  double best = -1.0;
  int bi = 0x7fffffff;
  for (int t = tid; t < ns; t += 256)
    if (bx.remS[t]) argmax_combine(best, bi, (double)cabs_(u[t]), t);
  rv[tid] = best;
  ri[tid] = bi;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (tid < k) {
      double v1 = rv[tid];
      int i1 = ri[tid];
      argmax_combine(v1, i1, rv[tid + k], ri[tid + k]);
      rv[tid] = v1;
      ri[tid] = i1;
    }
    __syncthreads();
  }
  if (tid == 0) {
    const double uv = sqrt(un2) * sqrt(vn2);
    if (bx.init_done) {
      double ns2 = bx.normS * bx.normS + uv * uv + acc;
      bx.normS = sqrt(ns2);
    }
    bx.init_done = 1;
    bx.uv = uv;
    bx.rank = rank + 1;
    const int mn = min(bx.N1, bx.N2);
    const int snext = ri[0];
    const bool more = (uv > tol * bx.normS) && (bx.rank < mn) && (snext >= 0 && snext < ns);
    bx.active = more ? 1 : 0;
    if (more) bx.s_cur = snext;
  }
}

__global__ void fill_u8_kernel(unsigned char* p, size_t n, unsigned char v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

template <typename T>
static void aca_group(const DevGeometry& g, int kind, double tol, const std::vector<AcaProblem>& probs,
                      std::vector<AcaResult>& out, Ctx& cx, AcaStats& stats);

// The residual vectors of every box of a group stay in HBM until the group converges
// (long side x rank each), so the boxes of a level are processed in groups whose expected
// footprint stays under a budget.
template <typename T>
void aca_batched(const DevGeometry& g, int kind, double tol, const std::vector<AcaProblem>& probs,
                 std::vector<AcaResult>& out, Ctx& cx, AcaStats& stats) {
  const size_t budget = size_t(24) << 30;
  out.assign(probs.size(), AcaResult());
  std::vector<size_t> grp;
  size_t used = 0;
  auto flush = [&]() {
    if (grp.empty()) return;
    std::vector<AcaProblem> sub(grp.size());
    for (size_t i = 0; i < grp.size(); ++i) sub[i] = probs[grp[i]];
    std::vector<AcaResult> r;
    aca_group<T>(g, kind, tol, sub, r, cx, stats);
    for (size_t i = 0; i < grp.size(); ++i) out[grp[i]] = std::move(r[i]);
    grp.clear();
    used = 0;
  };
  for (size_t b = 0; b < probs.size(); ++b) {
    const size_t n1 = probs[b].rows.size(), n2 = probs[b].cols.size();
    const size_t ns = std::min(n1, n2), nl = std::max(n1, n2);
    const size_t cap = std::min<size_t>(ns, probs[b].pred > 0 ? probs[b].pred + probs[b].pred / 4 + 8 : 256);
    const size_t bytes = (nl + ns) * cap * sizeof(T);
    if (!grp.empty() && used + bytes > budget) flush();
    grp.push_back(b);
    used += bytes;
  }
  flush();
}

template <typename T>
static void aca_group(const DevGeometry& g, int kind, double tol, const std::vector<AcaProblem>& probs,
                      std::vector<AcaResult>& out, Ctx& cx, AcaStats& stats) {
  using R = decltype(T().re);
  const int nb = (int)probs.size();
  out.assign(nb, AcaResult());
  if (nb == 0) return;
  std::vector<AcaBox<T>> hb(nb);
  std::vector<void*> owned;
  auto alloc = [&](size_t bytes) {
    void* p = cx.pool.alloc(bytes);
    owned.push_back(p);
    return p;
  };
  std::vector<int> live;
  for (int b = 0; b < nb; ++b) {
    const AcaProblem& P = probs[b];
    AcaBox<T>& x = hb[b];
    std::memset(&x, 0, sizeof(x));
    x.N1 = (int)P.rows.size();
    x.N2 = (int)P.cols.size();
    if (x.N1 == 0 || x.N2 == 0) continue;
    x.branch = (x.N1 < x.N2) ? 1 : 2;
    x.ns = std::min(x.N1, x.N2);
    x.nl = std::max(x.N1, x.N2);
    // U/V capacity: the predicted rank with a margin when there is one
    x.cap = P.pred > 0 ? std::min(x.ns, P.pred + P.pred / 4 + 8) : std::min(x.ns, 64);
    x.nchunks = (x.nl + ACA_CHUNK - 1) / ACA_CHUNK;
    int* dr = (int*)alloc(sizeof(int) * x.N1);
    int* dc = (int*)alloc(sizeof(int) * x.N2);
    cx.upload(dr, P.rows.data(), sizeof(int) * x.N1);
    cx.upload(dc, P.cols.data(), sizeof(int) * x.N2);
    x.rows = dr;
    x.cols = dc;
    x.U = (T*)cx.pool.alloc(sizeof(T) * (size_t)x.ns * x.cap);
    x.V = (T*)cx.pool.alloc(sizeof(T) * (size_t)x.nl * x.cap);
    x.W = (T*)alloc(sizeof(T) * (size_t)x.nl);
    x.remS = (unsigned char*)alloc(x.ns);
    x.remL = (unsigned char*)alloc(x.nl);
    x.pivS = (int*)alloc(sizeof(int) * x.ns);
    x.pivL = (int*)alloc(sizeof(int) * x.ns);
    x.pmax = (double*)alloc(sizeof(double) * x.nchunks);
    x.pidx = (int*)alloc(sizeof(int) * x.nchunks);
    x.pdot = (double*)cx.pool.alloc(sizeof(double) * 2 * (size_t)x.nchunks * (x.cap + 1));
    hipLaunchKernelGGL(fill_u8_kernel, dim3((x.ns + 255) / 256), dim3(256), 0, cx.st, x.remS, (size_t)x.ns, (unsigned char)1);
    hipLaunchKernelGGL(fill_u8_kernel, dim3((x.nl + 255) / 256), dim3(256), 0, cx.st, x.remL, (size_t)x.nl, (unsigned char)1);
    x.active = 1;
    x.s_cur = 0;
    live.push_back(b);
  }
  AcaBox<T>* d_boxes = (AcaBox<T>*)cx.pool.alloc(sizeof(AcaBox<T>) * nb);
  cx.upload(d_boxes, hb.data(), sizeof(AcaBox<T>) * nb);
  std::vector<int> active = live;
  bool predicted = false;
  for (int b : live) predicted = predicted || probs[b].pred > 0;
  bool first = true;
  while (!active.empty()) {
    std::vector<int2> work;
    for (int b : active)
      for (int c = 0; c < hb[b].nchunks; ++c) work.push_back(make_int2(b, c));
    // steps between host checks: the predicted remaining rank on the first pass, then the
    // fixed burst (converged boxes skip the extra steps inside the kernels, so a burst
    // never changes the skeletons, only how often the host looks)
    int burst = std::max(1, stats.burst);
    if (predicted && first) {  // the prediction plus an eighth: extra steps of a converged box are no-ops
      int need = 1;
      for (int b : active) {
        const int p = probs[b].pred + probs[b].pred / 8 + 2;
        need = std::max(need, std::min(p, hb[b].cap) - hb[b].rank);
      }
      burst = std::min(need, 256);
    }
    first = false;
    const int2* d_work = (const int2*)cx.push(work.data(), sizeof(int2) * work.size());
    const int* d_ids = (const int*)cx.push(active.data(), sizeof(int) * active.size());
    for (int st = 0; st < burst; ++st) {
      hipLaunchKernelGGL((aca_long_kernel<R, T>), dim3(work.size()), dim3(256), 0, cx.st, g, kind, d_boxes, d_work);
      hipLaunchKernelGGL(aca_pivot_kernel<T>, dim3(active.size()), dim3(256), 0, cx.st, d_boxes, d_ids);
      hipLaunchKernelGGL(aca_norm_kernel<T>, dim3(work.size()), dim3(256), 0, cx.st, d_boxes, d_work);
      hipLaunchKernelGGL((aca_short_kernel<R, T>), dim3(active.size()), dim3(256), 0, cx.st, g, kind, d_boxes, d_ids,
                         tol);
    }
    stats.steps += burst;
    stats.syncs += 1;
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(hb.data(), d_boxes, sizeof(AcaBox<T>) * nb, hipMemcpyDeviceToHost, cx.st));
    cx.sync();
    std::vector<int> next;
    bool grown = false;
    for (int b : active) {
      AcaBox<T>& x = hb[b];
      if (!x.active) continue;
      next.push_back(b);
      if (x.rank >= x.cap) {  // grow U, V and the dot partials
        int ncap = std::min(x.ns, x.cap * 2);
        T* nU = (T*)cx.pool.alloc(sizeof(T) * (size_t)x.ns * ncap);
        T* nV = (T*)cx.pool.alloc(sizeof(T) * (size_t)x.nl * ncap);
        HIPCHK(hipMemcpyAsync(nU, x.U, sizeof(T) * (size_t)x.ns * x.cap, hipMemcpyDeviceToDevice, cx.st));
        HIPCHK(hipMemcpyAsync(nV, x.V, sizeof(T) * (size_t)x.nl * x.cap, hipMemcpyDeviceToDevice, cx.st));
        cx.pool.free(x.U);
        cx.pool.free(x.V);
        cx.pool.free(x.pdot);
        x.U = nU;
        x.V = nV;
        x.cap = ncap;
        x.pdot = (double*)cx.pool.alloc(sizeof(double) * 2 * (size_t)x.nchunks * (x.cap + 1));
        grown = true;
      }
    }
    if (grown) cx.upload(d_boxes, hb.data(), sizeof(AcaBox<T>) * nb);
    active.swap(next);
  }
  for (int b : live) {
    AcaBox<T>& x = hb[b];
    AcaResult& r = out[b];
    r.rank = x.rank;
    std::vector<int> ps(x.rank), pl(x.rank);
    if (x.rank > 0) {
      HIPCHK(hipMemcpyAsync(ps.data(), x.pivS, sizeof(int) * x.rank, hipMemcpyDeviceToHost, cx.st));
      HIPCHK(hipMemcpyAsync(pl.data(), x.pivL, sizeof(int) * x.rank, hipMemcpyDeviceToHost, cx.st));
    }
    cx.sync();
    if (x.branch == 1) { r.row_bases = ps; r.col_bases = pl; }
    else { r.row_bases = pl; r.col_bases = ps; }
    cx.pool.free(x.U);
    cx.pool.free(x.V);
    cx.pool.free(x.pdot);
  }
  for (void* p : owned) cx.pool.free(p);
  cx.pool.free(d_boxes);
}

template <typename T>
__global__ void __launch_bounds__(256) maxcol_kernel(const DMat<T>* mats, double* out) {
  const DMat<T> m = mats[blockIdx.x];
  __shared__ double rv[256];
  double best = 0.0;
  for (int c = 0; c < m.c; ++c) {
    double s = 0;
    for (int i = threadIdx.x; i < m.r; i += 256) {
      T a = m.p[(size_t)i + (size_t)c * m.ld];
      s += (double)a.re * a.re + (double)a.im * a.im;
    }
    rv[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
      if (threadIdx.x < k) rv[threadIdx.x] += rv[threadIdx.x + k];
      __syncthreads();
    }
    best = fmax(best, sqrt(rv[0]));
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = best;
}

template <typename T>
void max_col_norms(const std::vector<DMat<T>>& mats, std::vector<double>& out, Ctx& cx) {
  out.assign(mats.size(), 0.0);
  if (mats.empty()) return;
  const DMat<T>* d = (const DMat<T>*)cx.push(mats.data(), sizeof(DMat<T>) * mats.size());
  hipLaunchKernelGGL(maxcol_kernel<T>, dim3(mats.size()), dim3(256), 0, cx.st, d, cx.d_scalars);
  const double* h = cx.fetch_scalars(mats.size());
  for (size_t i = 0; i < mats.size(); ++i) out[i] = h[i];
}

template void aca_batched<cpx<double>>(const DevGeometry&, int, double, const std::vector<AcaProblem>&,
                                       std::vector<AcaResult>&, Ctx&, AcaStats&);
template void aca_batched<cpx<float>>(const DevGeometry&, int, double, const std::vector<AcaProblem>&,
                                      std::vector<AcaResult>&, Ctx&, AcaStats&);
template void max_col_norms<cpx<double>>(const std::vector<DMat<cpx<double>>>&, std::vector<double>&, Ctx&);
template void max_col_norms<cpx<float>>(const std::vector<DMat<cpx<float>>>&, std::vector<double>&, Ctx&);

}  // namespace rcs
