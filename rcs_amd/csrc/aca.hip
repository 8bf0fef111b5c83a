// Batched ACA on gfx950: see aca.hpp. The per-step semantics follow src/ACA.cpp:639-859:
// residual long vector at the current short index, first-max pivot over the remaining
// long indices, v = residual / pivot, residual short vector at the pivot, u = that
// vector, Frobenius estimate update (A:750-755, the sum over l includes the new pair),
// next short pivot = first max of |u| over the remaining short indices, continue while
// |u||v| > tol*|S| and rank < min(N1, N2).
//
// The stopping test of step k needs V(:,q)^H v_k for every q <= k, i.e. a pass over all of
// V, and step k+1's residual long vector needs the same pass. Both are fused: step k+1's
// entry kernel writes v_k = w_k / pivot into V, its long kernel forms the dots and the new
// residual in one read of V, and its pivot kernel finishes step k's test before it pivots.
// A box that stops has computed one residual long vector too many; its skeleton is
// unchanged.
#include <algorithm>
#include <chrono>

#include "aca.hpp"
#include "entries_dev.hpp"

namespace rcs {

constexpr int ACA_CHUNK = 1024;  // long entries per workgroup (4 per thread)
constexpr int ACA_SLOTS = 4;     // dot partials per chunk (one per wavefront)
constexpr int ACA_PANEL = 64;    // V columns per separately allocated panel

template <typename T>
struct AcaBox {
  const int* rows;
  const int* cols;
  int N1, N2, branch;  // branch 1: rows are the short side (N1 < N2)
  int ns, nl, cap, nchunks;
  T* U;  // ns x cap residual short vectors
  // V, the normalised long vectors, in panels of ACA_PANEL columns, each stored chunk-major
  // ([chunk][column][ACA_CHUNK]) so that one workgroup streams a contiguous block of it
  T** Vp;
  T* W;  // nl: residual long vector of the latest step
  unsigned char* remS;
  unsigned char* remL;
  int* pivS;
  int* pivL;
  double* pmax;  // nchunks
  int* pidx;     // nchunks
  double* pdot;  // (nchunks * ACA_SLOTS) x cap x 2: V(:,q)^H v of the pending step, per slot
  double* du;    // cap x 2: U(:,q)^H u of the pending step
  int rank;      // steps taken
  int active, s_cur, l_cur, retry;
  int pend;      // the latest step's v is not yet in V and its stopping test is not done
  double normS, uv;
  T piv;  // pivot of the latest step
};

// The ACA_CHUNK entries of V column q that belong to long chunk c.
template <typename T>
__device__ inline T* vcol(const AcaBox<T>& bx, int q, int c) {
  return bx.Vp[q / ACA_PANEL] + ((size_t)c * ACA_PANEL + q % ACA_PANEL) * ACA_CHUNK;
}

template <typename R>
__device__ inline R cabs_(cpx<R> a) { return sizeof(R) == 8 ? (R)hypot((double)a.re, (double)a.im) : (R)hypotf((float)a.re, (float)a.im); }

__device__ inline void argmax_combine(double& v, int& i, double v2, int i2) {
  if (v2 > v || (v2 == v && i2 < i)) { v = v2; i = i2; }
}

__device__ inline void wave_sum2(double& a, double& b) {
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off);
    b += __shfl_xor(b, off);
  }
}

// First maximum over a workgroup of 256 (values in rv, indices in ri); result in rv[0], ri[0].
__device__ inline void block_argmax(double* rv, int* ri, int tid) {
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
}

// K0 (one workgroup per chunk of a box's long side): the pending step's v = w / pivot into
// V(:,rank-1), then this step's long entries (row or column of the box x interaction-list
// block at short index s_cur) into W. Entry generation lives apart from the streaming
// kernel below, whose occupancy the quadrature's registers would otherwise cap.
template <typename R, typename T>
__global__ void __launch_bounds__(256) aca_long_entries_kernel(DevGeometry g, int kind, AcaBox<T>* boxes,
                                                               const int2* work) {
  const int b = work[blockIdx.x].x, chunk = work[blockIdx.x].y;
  AcaBox<T>& bx = boxes[b];
  // bursts of steps run between host checks, so the launch lists can be stale
  if (!bx.active) return;
  const int rank = bx.rank, nl = bx.nl, s = bx.s_cur;
  const bool pend = bx.pend != 0, step = rank < bx.cap;
  const T piv = bx.piv;
  T* vc = pend ? vcol(bx, rank - 1, chunk) : nullptr;
  for (int e = 0; e < ACA_CHUNK / 256; ++e) {
    const int t = chunk * ACA_CHUNK + e * 256 + (int)threadIdx.x;
    if (t >= nl) break;
    if (pend) vc[e * 256 + threadIdx.x] = cdiv(bx.W[t], piv);
    if (step) {
      int i, j;
      if (bx.branch == 1) { i = bx.rows[s]; j = bx.cols[t]; }
      else { i = bx.rows[t]; j = bx.cols[s]; }
      cpx<R> v = kernel_entry<R>(g, kind, i, j);
      bx.W[t] = T(v.re, v.im);
    }
  }
}

// Sums x[0..7] over the 64 lanes of a wavefront with 10 exchanges instead of 48: at each of
// the first three levels a lane keeps half of its values and trades the other half with its
// partner. Lane l ends with the total of x[(l >> 3) & 7].
__device__ inline double wave_sum8(const double (&x)[8], int lane) {
  const int b5 = (lane >> 5) & 1, b4 = (lane >> 4) & 1, b3 = (lane >> 3) & 1;
  double y[4], z[2];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const double keep = b5 ? x[i + 4] : x[i], send = b5 ? x[i] : x[i + 4];
    y[i] = keep + __shfl_xor(send, 32);
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const double keep = b4 ? y[i + 2] : y[i], send = b4 ? y[i] : y[i + 2];
    z[i] = keep + __shfl_xor(send, 16);
  }
  double w = (b3 ? z[1] : z[0]) + __shfl_xor(b3 ? z[0] : z[1], 8);
  w += __shfl_xor(w, 4);
  w += __shfl_xor(w, 2);
  w += __shfl_xor(w, 1);
  return w;
}

constexpr int ACA_QB = 4;  // V columns per iteration of the streaming loop (loads issued together)

// K1 (one workgroup per chunk of a box's long side), one pass over V(:, 0..rank-1):
//  pending step: partial dots V(:,q)^H v, q <= rank-1, with v = V(:,rank-1);
//  this step (when U/V have room): residual long vector w = W - V U(s,:)^T -> W, partial
//  first max of |w| over the remaining long indices.
template <typename T>
__global__ void __launch_bounds__(256) aca_long_kernel(AcaBox<T>* boxes, const int2* work) {
  const int b = work[blockIdx.x].x, chunk = work[blockIdx.x].y;
  AcaBox<T>& bx = boxes[b];
  if (!bx.active) return;
  const int rank = bx.rank, nl = bx.nl, cap = bx.cap, s = bx.s_cur;
  const bool pend = bx.pend != 0;
  const bool step = rank < cap;
  if (!pend && !step) return;
  __shared__ T us[ACA_PANEL];
  __shared__ double rv[256];
  __shared__ int ri[256];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int tt[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) tt[e] = chunk * ACA_CHUNK + e * 256 + tid;
  T vn[4], val[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const bool in = tt[e] < nl;
    vn[e] = (pend && in) ? vcol(bx, rank - 1, chunk)[e * 256 + tid] : T(0, 0);
    val[e] = (step && in) ? bx.W[tt[e]] : T(0, 0);
  }
  double* pd = bx.pdot + (size_t)(chunk * ACA_SLOTS + wave) * cap * 2;
  for (int q0 = 0; q0 < rank; q0 += ACA_PANEL) {
    const int nq = min(ACA_PANEL, rank - q0);
    const T* vp = vcol(bx, q0, chunk);  // this panel's columns of the chunk, contiguous
    if (step) {
      __syncthreads();
      if (tid < nq) us[tid] = bx.U[(size_t)(q0 + tid) * bx.ns + s];
      __syncthreads();
    }
    for (int qq = 0; qq < nq; qq += ACA_QB) {
      T a[ACA_QB][4];
#pragma unroll
      for (int i = 0; i < ACA_QB; ++i) {
        const T* vq = vp + (size_t)(qq + i) * ACA_CHUNK;
#pragma unroll
        for (int e = 0; e < 4; ++e) a[i][e] = (qq + i < nq && tt[e] < nl) ? vq[e * 256 + tid] : T(0, 0);
      }
      if (step) {
#pragma unroll
        for (int i = 0; i < ACA_QB; ++i) {
          const T uq = qq + i < nq ? us[qq + i] : T(0, 0);
#pragma unroll
          for (int e = 0; e < 4; ++e) val[e] = val[e] - uq * a[i][e];
        }
      }
      if (pend) {
        double d[2 * ACA_QB];
#pragma unroll
        for (int i = 0; i < ACA_QB; ++i) {
          double sr = 0, si = 0;
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            sr += (double)a[i][e].re * vn[e].re + (double)a[i][e].im * vn[e].im;
            si += (double)a[i][e].re * vn[e].im - (double)a[i][e].im * vn[e].re;
          }
          d[2 * i] = sr;
          d[2 * i + 1] = si;
        }
        const double tot = wave_sum8(d, lane);
        const int idx = lane >> 3;  // re/im of column qq + idx / 2
        if ((lane & 7) == 0 && qq + (idx >> 1) < nq) pd[2 * (q0 + qq) + idx] = tot;
      }
    }
  }
  if (!step) return;
  double best = -1.0;
  int bi = 0x7fffffff;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    if (tt[e] < nl) {
      bx.W[tt[e]] = val[e];
      if (bx.remL[tt[e]]) argmax_combine(best, bi, (double)cabs_(val[e]), tt[e]);
    }
  }
  rv[tid] = best;
  ri[tid] = bi;
  block_argmax(rv, ri, tid);
  if (tid == 0) {
    bx.pmax[chunk] = rv[0];
    bx.pidx[chunk] = ri[0];
  }
}

// K2 (one workgroup per box): the pending step's Frobenius update and stopping test, then
// this step's long pivot.
template <typename T>
__global__ void __launch_bounds__(256) aca_pivot_kernel(AcaBox<T>* boxes, const int* ids, double tol) {
  AcaBox<T>& bx = boxes[ids[blockIdx.x]];
  if (!bx.active) return;
  __shared__ double rv[256];
  __shared__ int ri[256];
  __shared__ double nrm[2];
  __shared__ int stop;
  const int tid = threadIdx.x;
  const int rank = bx.rank;
  if (bx.pend) {
    // |S|^2 += (|u||v|)^2 + sum_{q<=k} 2|U(:,q)^H u . V(:,q)^H v|  (A:750-755)
    const int k = rank - 1, cap = bx.cap, nslots = bx.nchunks * ACA_SLOTS;
    double acc = 0;
    for (int q = tid; q <= k; q += 256) {
      double dvr = 0, dvi = 0;
      for (int sl = 0; sl < nslots; ++sl) {
        const double* pd = bx.pdot + ((size_t)sl * cap + q) * 2;
        dvr += pd[0];
        dvi += pd[1];
      }
      const double dur = bx.du[2 * q], dui = bx.du[2 * q + 1];
      const double pr = dur * dvr - dui * dvi, pi = dur * dvi + dui * dvr;
      acc += 2.0 * sqrt(pr * pr + pi * pi);
      if (q == k) { nrm[0] = dur; nrm[1] = dvr; }
    }
    rv[tid] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
      if (tid < h) rv[tid] += rv[tid + h];
      __syncthreads();
    }
    if (tid == 0) {
      const double uv = sqrt(nrm[0]) * sqrt(nrm[1]);
      if (k > 0) bx.normS = sqrt(bx.normS * bx.normS + uv * uv + rv[0]);  // |S| is 0 after step 0
      bx.uv = uv;
      bx.pend = 0;
      stop = (uv > tol * bx.normS) ? 0 : 1;
      if (stop) bx.active = 0;
    }
    __syncthreads();
    if (stop) return;
  }
  if (rank >= bx.cap) {  // U/V full: the host grows them, then the step runs
    if (tid == 0) bx.retry = 1;
    return;
  }
  double best = -1.0;
  int bi = 0x7fffffff;
  for (int c = tid; c < bx.nchunks; c += 256) argmax_combine(best, bi, bx.pmax[c], bx.pidx[c]);
  rv[tid] = best;
  ri[tid] = bi;
  block_argmax(rv, ri, tid);
  if (tid == 0) {
    const int l = ri[0];
    const double pv = rv[0];
    if (rank == 0 && !(pv > 1e-16)) {  // A:663-669: skip an all-zero first vector
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
    bx.pivS[rank] = bx.s_cur;
    bx.pivL[rank] = l;
    bx.remL[l] = 0;
    bx.remS[bx.s_cur] = 0;
  }
}

// K2b (one workgroup per 256 short indices of a box): this step's short entries (column or
// row of the block at the long pivot) into U(:,rank).
template <typename R, typename T>
__global__ void __launch_bounds__(256) aca_short_entries_kernel(DevGeometry g, int kind, AcaBox<T>* boxes,
                                                                const int2* work) {
  const int b = work[blockIdx.x].x, t = work[blockIdx.x].y * 256 + (int)threadIdx.x;
  AcaBox<T>& bx = boxes[b];
  if (bx.retry || !bx.active || t >= bx.ns) return;
  const int l = bx.l_cur;
  int i, j;
  if (bx.branch == 1) { i = bx.rows[t]; j = bx.cols[l]; }
  else { i = bx.rows[l]; j = bx.cols[t]; }
  cpx<R> e = kernel_entry<R>(g, kind, i, j);
  bx.U[(size_t)bx.rank * bx.ns + t] = T(e.re, e.im);
}

// K3 (one workgroup per box): residual short vector u = U(:,rank) - U V(l,:)^T at the long
// pivot, its dots U(:,q)^H u for q <= rank (one wavefront per column), next short pivot.
template <typename T>
__global__ void __launch_bounds__(256) aca_short_kernel(AcaBox<T>* boxes, const int* ids) {
  AcaBox<T>& bx = boxes[ids[blockIdx.x]];
  if (bx.retry || !bx.active) return;
  const int rank = bx.rank, ns = bx.ns, nl = bx.nl, l = bx.l_cur;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  __shared__ T vs[ACA_PANEL];
  __shared__ double rv[256];
  __shared__ int ri[256];
  T* u = bx.U + (size_t)rank * ns;
  for (int q0 = 0; q0 < rank; q0 += ACA_PANEL) {
    const int nq = min(ACA_PANEL, rank - q0);
    __syncthreads();
    if (tid < nq) vs[tid] = vcol(bx, q0 + tid, l / ACA_CHUNK)[l % ACA_CHUNK];
    __syncthreads();
    for (int t = tid; t < ns; t += 256) {
      T a = u[t];
      for (int q = 0; q < nq; ++q) a = a - vs[q] * bx.U[(size_t)(q0 + q) * ns + t];
      u[t] = a;
    }
  }
  __syncthreads();
  for (int q = wave; q <= rank; q += 4) {
    const T* uq = bx.U + (size_t)q * ns;
    double sr = 0, si = 0;
    for (int t = lane; t < ns; t += 64) {
      const T a = uq[t], c = u[t];
      sr += (double)a.re * c.re + (double)a.im * c.im;
      si += (double)a.re * c.im - (double)a.im * c.re;
    }
    wave_sum2(sr, si);
    if (lane == 0) {
      bx.du[2 * q] = sr;
      bx.du[2 * q + 1] = si;
    }
  }
  double best = -1.0;
  int bi = 0x7fffffff;
  for (int t = tid; t < ns; t += 256)
    if (bx.remS[t]) argmax_combine(best, bi, (double)cabs_(u[t]), t);
  rv[tid] = best;
  ri[tid] = bi;
  block_argmax(rv, ri, tid);
  if (tid == 0) {
    const int nr = rank + 1, snext = ri[0];
    bx.rank = nr;
    bx.pend = 1;
    // rank = min(N1, N2) ends the loop without the Frobenius test (A:709, A:802)
    if (nr >= min(bx.N1, bx.N2) || snext < 0 || snext >= ns) bx.active = 0;
    else bx.s_cur = snext;
  }
}

__global__ void fill_u8_kernel(unsigned char* p, size_t n, unsigned char v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

template <typename T>
static void aca_group(const DevGeometry& g, int kind, double tol, const std::vector<AcaProblem>& probs,
                      std::vector<AcaResult>& out, Ctx& cx, AcaStats& stats);

// U/V capacity in whole V panels: the predicted rank with a margin when there is one.
static int aca_cap0(int ns, int pred) {
  const int want = pred > 0 ? pred + pred / 4 + 8 : ACA_PANEL;
  return std::min(ns, (want + ACA_PANEL - 1) / ACA_PANEL * ACA_PANEL);
}

// The vectors of every box of a group stay in HBM until the group converges (long + short
// side x rank each), so the boxes of a level run in groups whose expected footprint (at a
// rank of 256, or the predicted capacity) fits in half of the device memory that is free
// or cached by the pool. The pool keeps what the ACA frees: V panels have one size per
// long side, so the next direction and level reuse them without going back to the driver.
template <typename T>
void aca_batched(const DevGeometry& g, int kind, double tol, const std::vector<AcaProblem>& probs,
                 std::vector<AcaResult>& out, Ctx& cx, AcaStats& stats) {
  size_t fr = 0, tot = 0;
  HIPCHK(hipMemGetInfo(&fr, &tot));
  const size_t avail = fr + (cx.pool.total() - cx.pool.in_use());
  const size_t budget = std::max<size_t>(avail / 2, size_t(1) << 30);
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
    const size_t cap = probs[b].pred > 0 ? (size_t)aca_cap0((int)ns, probs[b].pred) : std::min<size_t>(ns, 256);
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
  std::vector<std::vector<T*>> panels(nb);  // V panels of each box (host copy of its table)
  std::vector<void*> owned;
  auto alloc = [&](size_t bytes) {
    void* p = cx.pool.alloc(bytes);
    owned.push_back(p);
    return p;
  };
  auto pdot_bytes = [](const AcaBox<T>& x) { return sizeof(double) * 2 * (size_t)x.nchunks * ACA_SLOTS * x.cap; };
  // V panels for columns [0, cap): allocate the missing ones and publish their pointers
  auto add_panels = [&](int b) {
    AcaBox<T>& x = hb[b];
    const int have = (int)panels[b].size(), want = (x.cap + ACA_PANEL - 1) / ACA_PANEL;
    for (int p = have; p < want; ++p)
      panels[b].push_back((T*)cx.pool.alloc(sizeof(T) * (size_t)x.nchunks * ACA_CHUNK * ACA_PANEL));
    if (want > have) cx.upload(x.Vp + have, panels[b].data() + have, sizeof(T*) * (want - have));
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
    x.cap = aca_cap0(x.ns, P.pred);
    x.nchunks = (x.nl + ACA_CHUNK - 1) / ACA_CHUNK;
    int* dr = (int*)alloc(sizeof(int) * x.N1);
    int* dc = (int*)alloc(sizeof(int) * x.N2);
    cx.upload(dr, P.rows.data(), sizeof(int) * x.N1);
    cx.upload(dc, P.cols.data(), sizeof(int) * x.N2);
    x.rows = dr;
    x.cols = dc;
    x.Vp = (T**)alloc(sizeof(T*) * ((x.ns + ACA_PANEL - 1) / ACA_PANEL));
    add_panels(b);
    x.U = (T*)cx.pool.alloc(sizeof(T) * (size_t)x.ns * x.cap);
    x.W = (T*)alloc(sizeof(T) * (size_t)x.nl);
    x.remS = (unsigned char*)alloc(x.ns);
    x.remL = (unsigned char*)alloc(x.nl);
    x.pivS = (int*)alloc(sizeof(int) * x.ns);
    x.pivL = (int*)alloc(sizeof(int) * x.ns);
    x.pmax = (double*)alloc(sizeof(double) * x.nchunks);
    x.pidx = (int*)alloc(sizeof(int) * x.nchunks);
    x.pdot = (double*)cx.pool.alloc(pdot_bytes(x));
    x.du = (double*)cx.pool.alloc(sizeof(double) * 2 * (size_t)x.cap);
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
    std::vector<int2> work, swork;
    for (int b : active) {
      for (int c = 0; c < hb[b].nchunks; ++c) work.push_back(make_int2(b, c));
      for (int c = 0; c < (hb[b].ns + 255) / 256; ++c) swork.push_back(make_int2(b, c));
    }
    // steps between host checks: the predicted remaining rank on the first pass, then the
    // fixed burst (stopped boxes skip the extra steps inside the kernels, so a burst never
    // changes the skeletons, only how often the host looks)
    int burst = std::max(1, stats.burst);
    if (predicted && first) {
      int need = 1;
      for (int b : active) {
        const int p = probs[b].pred + probs[b].pred / 8 + 2;
        need = std::max(need, std::min(p, hb[b].cap) - hb[b].rank);
      }
      burst = std::min(need + 1, 257);  // + 1: the last step's test runs in the next step's launches
    }
    first = false;
    const int2* d_work = (const int2*)cx.push(work.data(), sizeof(int2) * work.size());
    const int2* d_swork = (const int2*)cx.push(swork.data(), sizeof(int2) * swork.size());
    const int* d_ids = (const int*)cx.push(active.data(), sizeof(int) * active.size());
    for (int st = 0; st < burst; ++st) {
      hipLaunchKernelGGL((aca_long_entries_kernel<R, T>), dim3(work.size()), dim3(256), 0, cx.st, g, kind, d_boxes,
                         d_work);
      hipLaunchKernelGGL(aca_long_kernel<T>, dim3(work.size()), dim3(256), 0, cx.st, d_boxes, d_work);
      hipLaunchKernelGGL(aca_pivot_kernel<T>, dim3(active.size()), dim3(256), 0, cx.st, d_boxes, d_ids, tol);
      hipLaunchKernelGGL((aca_short_entries_kernel<R, T>), dim3(swork.size()), dim3(256), 0, cx.st, g, kind, d_boxes,
                         d_swork);
      hipLaunchKernelGGL(aca_short_kernel<T>, dim3(active.size()), dim3(256), 0, cx.st, d_boxes, d_ids);
    }
    stats.steps += burst;
    stats.syncs += 1;
    HIPCHK(hipGetLastError());
    const auto w0 = std::chrono::steady_clock::now();
    HIPCHK(hipMemcpyAsync(hb.data(), d_boxes, sizeof(AcaBox<T>) * nb, hipMemcpyDeviceToHost, cx.st));
    cx.sync();
    const auto w1 = std::chrono::steady_clock::now();
    stats.wait_s += std::chrono::duration<double>(w1 - w0).count();
    // grow U/V ahead of the next burst, so that a box rarely waits at a full capacity: more
    // V panels (half as many again), U and the pending U dots copied into larger blocks, new
    // dot partials. Every launch is on the one stream, so blocks given back to the pool
    // here are not reused before the copies out of them have run.
    const int margin = std::max(1, stats.burst) + 1;
    std::vector<int> next;
    bool grown = false;
    for (int b : active) {
      AcaBox<T>& x = hb[b];
      if (!x.active) continue;
      next.push_back(b);
      if (x.cap < x.ns && x.rank + margin > x.cap) {
        const int np = (x.cap + ACA_PANEL - 1) / ACA_PANEL;
        const int ncap = std::min(x.ns, (np + std::max(1, np / 2)) * ACA_PANEL);
        T* nU = (T*)cx.pool.alloc(sizeof(T) * (size_t)x.ns * ncap);
        double* ndu = (double*)cx.pool.alloc(sizeof(double) * 2 * (size_t)ncap);
        HIPCHK(hipMemcpyAsync(nU, x.U, sizeof(T) * (size_t)x.ns * x.cap, hipMemcpyDeviceToDevice, cx.st));
        HIPCHK(hipMemcpyAsync(ndu, x.du, sizeof(double) * 2 * (size_t)x.cap, hipMemcpyDeviceToDevice, cx.st));
        cx.pool.free(x.U);
        cx.pool.free(x.du);
        cx.pool.free(x.pdot);
        x.U = nU;
        x.du = ndu;
        x.cap = ncap;
        x.pdot = (double*)cx.pool.alloc(pdot_bytes(x));
        add_panels(b);
        x.retry = 0;
        grown = true;
      }
    }
    if (grown) cx.upload(d_boxes, hb.data(), sizeof(AcaBox<T>) * nb);
    stats.grow_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - w1).count();
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
    for (T* p : panels[b]) cx.pool.free(p);
    cx.pool.free(x.U);
    cx.pool.free(x.pdot);
    cx.pool.free(x.du);
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
