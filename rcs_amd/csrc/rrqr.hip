// Partitioned column-pivoted Householder QR (RRQR::colPivHouseholderQRWithPartition,
// src/rrqr.cpp:69-182) for several independent matrices at once, one fused launch per
// pivot step.
//
// A pivot step needs the downdated norms of every remaining column (pivot choice) and
// the pivot column (Householder vector) before any column can be updated. Instead of a
// pivot kernel, a state kernel and an update kernel per step, every workgroup of one
// launch recomputes the pivot choice and the Householder vector from the same data with
// the same arithmetic (so all workgroups agree bit for bit) and then updates its own
// four columns, one wavefront per column. Matrices and norms are double-buffered: a
// step reads buffer b and writes buffer 1-b, so no workgroup reads what another writes
// in the same launch; the step state is double-buffered the same way. Columns are never
// swapped in memory: logical position j takes its data from the pivot column p and
// position p from column j.
//
// Semantics kept from src/rrqr.cpp: first maximal downdated norm wins (Q:98-99,
// 143-144 resolved to the lowest index); the first partition stops at
// pivot/max1 < tol1 with rank1 = j-1 (Q:115-117) and restarts the second partition at
// rank1 with the norms of [rank1, partition) zeroed; the second partition stops at
// pivot/max(inputPivot, first pivot) < tol2 or after fillInRank columns (Q:162-166);
// norms are downdated, never recomputed (Q:124-127).
#include "dense.hpp"

namespace rcs {

template <typename T> struct RqReal;
template <> struct RqReal<cpx<double>> { using type = double; };
template <> struct RqReal<cpx<float>> { using type = float; };

struct RRQRState {
  int m, n, partition, fillInRank;
  int j, phase, stopped, rank1, first, steps, rank2, nrefl, buf, pad;
  double tol1, tol2, inputPivot, max1, maxPivot;
};

template <typename T>
struct RRQRJob {
  T* A[2];         // m x n, ld m
  double* ns[2];   // n squared column norms
  T* V;            // m x (min(m,n)+2) Householder vectors, ld m
  RRQRState* st;   // two state slots
  int m, n, wg0;   // wg0: first workgroup of this job in a launch
};

// Job tables travel as kernel arguments (no parameter copies to HBM).
constexpr int RQ_MAXJ = 8;
template <typename T>
struct RRQRJobs {
  RRQRJob<T> j[RQ_MAXJ];
  int n;
};
struct RRQRInits {
  RRQRState s[RQ_MAXJ];
};
template <typename T>
struct RRQRQJob {
  const T* V;
  const RRQRState* st;
  T* Q;
  int ldq, m, wg0;
};
template <typename T>
struct RRQRQJobs {
  RRQRQJob<T> j[RQ_MAXJ];
  int n;
};

// the job owning workgroup wg (static indices only: the table stays in the kernarg segment)
template <typename J>
__device__ inline J rq_pick(const J (&jobs)[RQ_MAXJ], int n, int wg) {
  J jb = jobs[0];
#pragma unroll
  for (int q = 1; q < RQ_MAXJ; ++q)
    if (q < n && jobs[q].wg0 <= wg) jb = jobs[q];
  return jb;
}

__device__ inline void rq_setup_phase2(RRQRState& s) {
  s.steps = s.steps - (s.partition - s.rank1);
  s.first = s.rank1;
  s.rank2 = min(s.steps - s.first, s.m);
  if (s.rank2 < 0) s.rank2 = 0;
  s.j = s.first;
  s.phase = 2;
}

// Acceptance test of step j alone (every thread evaluates it; no state copy).
__device__ inline int rq_accept(const RRQRState& s, int j, double pivot) {
  if (s.phase == 1) {
    const double max1 = (j == 0) ? pivot : s.max1;
    return (pivot / max1 >= s.tol1) ? 1 : 0;
  }
  const double mp = (j == s.first) ? fmax(s.inputPivot, pivot) : s.maxPivot;
  return (pivot / mp >= s.tol2 && !(j - s.first >= s.fillInRank)) ? 1 : 0;
}

// Step decision (src/rrqr.cpp:108-121 and 155-170): returns 1 if column j is accepted;
// u receives the next state.
__device__ inline int rq_decide(const RRQRState& s, int j, double pivot, RRQRState& u) {
  u = s;
  if (s.phase == 1) {
    if (j == 0) u.max1 = pivot;
    u.maxPivot = u.max1;
    if (!(pivot / u.max1 >= u.tol1)) {  // pivot/max1 < tol1 (NaN counts as below)
      u.rank1 = j - 1;
      u.phase = 3;  // second-partition set-up at the next step
      return 0;
    }
  } else {
    if (j == s.first) u.maxPivot = fmax(s.inputPivot, pivot);
    if (!(pivot / u.maxPivot >= u.tol2) || j - s.first >= s.fillInRank) {
      u.rank2 = j - s.first;
      u.stopped = 1;
      return 0;
    }
  }
  u.nrefl = s.nrefl + 1;
  u.j = j + 1;
  u.buf = s.buf ^ 1;
  return 1;
}

// Squared column norms of the input (one workgroup per column of every job); the first
// workgroup of a job also writes its initial state.
template <typename T>
__global__ void __launch_bounds__(256) rrqr_norms_kernel(RRQRJobs<T> jobs, RRQRInits init) {
  const RRQRJob<T> jb = rq_pick(jobs.j, jobs.n, blockIdx.x);
  const int c = blockIdx.x - jb.wg0;
  if (c == 0 && threadIdx.x == 0) {
    int q = 0;
#pragma unroll
    for (int t = 1; t < RQ_MAXJ; ++t)
      if (t < jobs.n && jobs.j[t].wg0 <= (int)blockIdx.x) q = t;
    // stored straight from the argument block: a selected aggregate would be built in
    // scratch memory
#pragma unroll
    for (int t = 0; t < RQ_MAXJ; ++t)
      if (t == q) jb.st[0] = init.s[t];
  }
  const T* a = jb.A[0] + (size_t)c * jb.m;
  double s = 0;
  for (int i = threadIdx.x; i < jb.m; i += blockDim.x) s += (double)a[i].re * a[i].re + (double)a[i].im * a[i].im;
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) jb.ns[0][c] = red[0];
}

constexpr int RQ_XS = 2048;  // pivot-column rows staged in LDS (longer columns read HBM)
constexpr int RQ_RC = 16;    // column rows per lane kept in registers (columns of <= 1024 rows)

// One pivot step of every job. Launch it reads state slot it&1 and writes (it+1)&1.
template <typename T>
__global__ void __launch_bounds__(256) rrqr_step_kernel(RRQRJobs<T> jobs, int it) {
  using R = typename RqReal<T>::type;
  const RRQRJob<T> jb = rq_pick(jobs.j, jobs.n, blockIdx.x);
  const int wgl = blockIdx.x - jb.wg0;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  __shared__ RRQRState s;
  __shared__ int sh_mlo, sh_mhi;
  __shared__ double rv[4], rn[4];
  __shared__ int ri[4];
  __shared__ T xs[RQ_XS];
  if (tid == 0) {
    s = jb.st[it & 1];
    int mlo = 0, mhi = 0;
    if (!s.stopped) {
      if (s.phase == 1 && s.j >= s.rank1) {  // first partition ran to its end
        mlo = max(s.rank1, 0);
        mhi = min(s.partition, s.n);
        rq_setup_phase2(s);
      } else if (s.phase == 3) {  // first partition stopped early at the previous step
        mlo = max(s.rank1, 0);
        mhi = min(s.partition, s.n);
        if (s.rank1 < 0) s.rank1 = 0;
        rq_setup_phase2(s);
      }
      if (s.phase == 2 && s.j >= s.steps) s.stopped = 1;
    }
    sh_mlo = mlo;
    sh_mhi = mhi;
  }
  __syncthreads();
  RRQRState* out = jb.st + ((it + 1) & 1);
  if (s.stopped) {
    if (wgl == 0 && tid == 0) *out = s;
    return;
  }
  const int j = s.j, m = s.m, n = s.n, buf = s.buf;
  const int mlo = sh_mlo, mhi = sh_mhi;  // norms of [mlo, mhi) read as zero in this step
  const int end = (s.phase == 1) ? s.partition : n;
  // buf is 0 or 1; selecting keeps the job's pointer pairs in registers (indexing them with
  // a run-time value would place the whole job in scratch memory)
  const double* nsi = buf ? jb.ns[1] : jb.ns[0];
  const T* Ain = buf ? jb.A[1] : jb.A[0];
  T* const Aout = buf ? jb.A[0] : jb.A[1];
  double* const nsout = buf ? jb.ns[0] : jb.ns[1];
  // this wavefront's column, rows j + lane + 64 t, loaded into registers now so that the
  // loads overlap the pivot search (positions j and p are reloaded once p is known)
  const int c = j + wgl * 4 + wave;
  const bool cached = m - j <= 64 * RQ_RC;
  T ar[RQ_RC];
  if (cached) {
#pragma unroll
    for (int t = 0; t < RQ_RC; ++t) {
      const int i = j + lane + 64 * t;
      ar[t] = (c < n && i < m) ? Ain[(size_t)c * m + i] : T(0, 0);
    }
  }
  // first maximal downdated norm over [j, end): (largest value, lowest index)
  double best = -1.0;
  int bi = end;
  for (int c0 = j; c0 < end; c0 += 1024) {  // four loads per thread in flight together
    double nv[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int c = c0 + tid + 256 * t;
      nv[t] = c < end ? ((c >= mlo && c < mhi) ? 0.0 : nsi[c]) : -2.0;
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int c = c0 + tid + 256 * t;
      if (c < end && nv[t] > best) { best = nv[t]; bi = c; }
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    double ob = __shfl_xor(best, off);
    int oi = __shfl_xor(bi, off);
    if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
  }
  if (lane == 0) { rv[wave] = best; ri[wave] = bi; }
  __syncthreads();
  best = rv[0];
  bi = ri[0];
#pragma unroll
  for (int w = 1; w < 4; ++w)
    if (rv[w] > best || (rv[w] == best && ri[w] < bi)) { best = rv[w]; bi = ri[w]; }
  const int p = (bi >= end) ? j : bi;
  // pivot column p, rows j..m-1, staged in LDS, and its norm
  const T* x = Ain + (size_t)p * m;
  const bool inl = m <= RQ_XS;
  double sn = 0;
#pragma unroll 4
  for (int i = j + tid; i < m; i += 256) {
    T xv = x[i];
    if (inl) xs[i] = xv;
    sn += (double)xv.re * xv.re + (double)xv.im * xv.im;
  }
  for (int off = 32; off > 0; off >>= 1) sn += __shfl_xor(sn, off);
  if (lane == 0) rn[wave] = sn;
  __syncthreads();
  const double xx = (rn[0] + rn[1]) + (rn[2] + rn[3]);
  // Householder vector v = scale*(x + e^{i arg x_j} ||x|| e_j) (src/rrqr.cpp:46-67)
  const double xn = sqrt(xx);
  const T xj = inl ? xs[j] : x[j];
  const double xj2 = (double)xj.re * xj.re + (double)xj.im * xj.im;
  const double ang = atan2((double)xj.im, (double)xj.re);
  const T v0 = T((R)((double)xj.re + cos(ang) * xn), (R)((double)xj.im + sin(ang) * xn));
  const double vn2 = xx - xj2 + ((double)v0.re * v0.re + (double)v0.im * v0.im);
  const double scale = vn2 > 0 ? sqrt(2.0) / sqrt(vn2) : 1.0;
  auto vel = [&](int i) -> T {  // element i >= j of v
    T vi = (i == j) ? v0 : (inl ? xs[i] : x[i]);
    if (vn2 > 0) vi = vi * (R)scale;
    return vi;
  };
  // diagonal after reflection x_j - v_j (v^H x), with v^H x = scale (conj(v0) x_j + ||x||^2 - |x_j|^2)
  const double hr = scale * ((double)v0.re * xj.re + (double)v0.im * xj.im + (xx - xj2));
  const double hi = scale * ((double)v0.re * xj.im - (double)v0.im * xj.re);
  const T vj = vel(j);
  const double dr = (double)xj.re - ((double)vj.re * hr - (double)vj.im * hi);
  const double di = (double)xj.im - ((double)vj.re * hi + (double)vj.im * hr);
  const double pivot = sqrt(dr * dr + di * di);
  const int acc = rq_accept(s, j, pivot);
  if (wgl == 0 && tid == 0) {  // the next state, written once per job
    RRQRState u;
    rq_decide(s, j, pivot, u);
    *out = u;
  }
  if (!acc) return;
  if (wgl == 0) {  // store the accepted reflector
    T* v = jb.V + (size_t)s.nrefl * m;
    for (int i = tid; i < m; i += 256) v[i] = (i < j) ? T(0, 0) : vel(i);
  }
  // apply to logical columns j..n-1, one wavefront per column, and downdate the norms
  if (c < n && cached) {
    const int src = (c == j) ? p : (c == p ? j : c);
    if (src != c) {
#pragma unroll
      for (int t = 0; t < RQ_RC; ++t) {
        const int i = j + lane + 64 * t;
        ar[t] = i < m ? Ain[(size_t)src * m + i] : T(0, 0);
      }
    }
    T vr[RQ_RC];
    double wr = 0, wi = 0;
#pragma unroll
    for (int t = 0; t < RQ_RC; ++t) {
      const int i = j + lane + 64 * t;
      vr[t] = i < m ? vel(i) : T(0, 0);
      wr += (double)vr[t].re * ar[t].re + (double)vr[t].im * ar[t].im;
      wi += (double)vr[t].re * ar[t].im - (double)vr[t].im * ar[t].re;
    }
    for (int off = 32; off > 0; off >>= 1) {
      wr += __shfl_xor(wr, off);
      wi += __shfl_xor(wi, off);
    }
    const T w((R)wr, (R)wi);
    T* ao = Aout + (size_t)c * m;
#pragma unroll
    for (int t = 0; t < RQ_RC; ++t) {
      const int i = j + lane + 64 * t;
      if (i < m) ao[i] = (c == j && i > j) ? T(0, 0) : ar[t] - vr[t] * w;
    }
    if (c != j && lane == 0) {
      const T d = ar[0] - vr[0] * w;
      const double dd = sqrt((double)d.re * d.re + (double)d.im * d.im);
      const double nsrc = (src >= mlo && src < mhi) ? 0.0 : nsi[src];
      nsout[c] = nsrc - dd * dd;
    }
  } else if (c < n) {
    const int src = (c == j) ? p : (c == p ? j : c);
    const T* a = Ain + (size_t)src * m;
    T* ao = Aout + (size_t)c * m;
    double ar = 0, ai = 0;
#pragma unroll 4
    for (int i = j + lane; i < m; i += 64) {
      T vi = vel(i), av = a[i];
      ar += (double)vi.re * av.re + (double)vi.im * av.im;
      ai += (double)vi.re * av.im - (double)vi.im * av.re;
    }
    for (int off = 32; off > 0; off >>= 1) {
      ar += __shfl_xor(ar, off);
      ai += __shfl_xor(ai, off);
    }
    const T w((R)ar, (R)ai);
    if (c == j) {
      for (int i = j + lane; i < m; i += 64) ao[i] = (i == j) ? a[i] - vel(i) * w : T(0, 0);
    } else {
#pragma unroll 4
      for (int i = j + lane; i < m; i += 64) ao[i] = a[i] - vel(i) * w;
      if (lane == 0) {
        const T d = a[j] - vel(j) * w;
        const double dd = sqrt((double)d.re * d.re + (double)d.im * d.im);
        const double nsrc = (src >= mlo && src < mhi) ? 0.0 : nsi[src];
        nsout[c] = nsrc - dd * dd;
      }
    }
  }
}

// Q(:, c) = H_0 H_1 ... H_{nrefl-1} e_c, one workgroup per output column of every job.
template <typename T>
__global__ void __launch_bounds__(256) rrqr_formq_kernel(RRQRQJobs<T> jobs) {
  using R = typename RqReal<T>::type;
  const RRQRQJob<T> jb = rq_pick(jobs.j, jobs.n, blockIdx.x);
  const int c = blockIdx.x - jb.wg0;
  const int m = jb.m, nrefl = jb.st->nrefl;
  T* q = jb.Q + (size_t)c * jb.ldq;
  for (int i = threadIdx.x; i < m; i += blockDim.x) q[i] = (i == c) ? T(1, 0) : T(0, 0);
  __syncthreads();
  __shared__ double rr[256], ri[256];
  for (int sidx = nrefl - 1; sidx >= 0; --sidx) {
    const T* v = jb.V + (size_t)sidx * m;
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

static size_t rq_align(size_t b) { return (b + 255) & ~size_t(255); }

template <typename T>
static void rrqr_chunk(std::vector<RRQRTask<T>>& tasks, const std::vector<int>& ids, double tol1, double tol2,
                       Ctx& cx) {
  RRQRJobs<T> jobs{};
  RRQRInits init{};
  std::vector<void*> work;
  int wgs = 0, budget = 0, ncols = 0, nj = 0;
  for (int t : ids) {
    RRQRTask<T>& tk = tasks[t];
    const int m = tk.A.r, n = tk.A.c;
    const int steps = std::min(m, n) + 2;
    const size_t bA = rq_align(sizeof(T) * (size_t)m * n), bN = rq_align(sizeof(double) * (size_t)n);
    const size_t bV = rq_align(sizeof(T) * (size_t)m * steps), bS = rq_align(2 * sizeof(RRQRState));
    char* w = (char*)cx.pool.alloc(2 * bA + 2 * bN + bV + bS);
    work.push_back(w);
    RRQRJob<T>& jb = jobs.j[nj];
    jb.A[0] = (T*)w;
    jb.A[1] = (T*)(w + bA);
    jb.ns[0] = (double*)(w + 2 * bA);
    jb.ns[1] = (double*)(w + 2 * bA + bN);
    jb.V = (T*)(w + 2 * bA + 2 * bN);
    jb.st = (RRQRState*)(w + 2 * bA + 2 * bN + bV);
    jb.m = m;
    jb.n = n;
    jb.wg0 = wgs;
    wgs += (n + 3) / 4;
    ncols += n;
    DMat<T> a0{jb.A[0], m, n, m};
    copy2d(a0, tk.A, cx.st);
    RRQRState& s = init.s[nj];
    s = RRQRState{};
    s.m = m; s.n = n; s.partition = tk.partition; s.fillInRank = tk.fillInRank;
    s.j = 0; s.phase = 1; s.stopped = 0;
    s.rank1 = std::min(tk.partition, m);
    s.steps = std::min(n, m);
    s.first = 0; s.rank2 = 0; s.nrefl = 0; s.buf = 0;
    s.tol1 = tol1; s.tol2 = tol2; s.inputPivot = tk.inputPivot; s.max1 = 0; s.maxPivot = 0;
    budget = std::max(budget, std::min(tk.partition, m) + std::min(tk.fillInRank, std::min(m, n)) + 3);
    ++nj;
  }
  jobs.n = nj;
  {  // norms (one workgroup per column) and initial states
    RRQRJobs<T> nrm = jobs;
    int w0 = 0;
    for (int i = 0; i < nj; ++i) { nrm.j[i].wg0 = w0; w0 += nrm.j[i].n; }
    hipLaunchKernelGGL(rrqr_norms_kernel<T>, dim3(ncols), dim3(256), 0, cx.st, nrm, init);
  }
  for (int it = 0; it < budget; ++it)
    hipLaunchKernelGGL(rrqr_step_kernel<T>, dim3(wgs), dim3(256), 0, cx.st, jobs, it);
  HIPCHK(hipGetLastError());
  const int slot = budget & 1;
  for (int i = 0; i < nj; ++i)
    HIPCHK(hipMemcpyAsync((char*)cx.h_scalars + i * sizeof(RRQRState), jobs.j[i].st + slot, sizeof(RRQRState),
                          hipMemcpyDeviceToHost, cx.st));
  cx.sync();
  RRQRState fin[RQ_MAXJ];
  std::memcpy(fin, cx.h_scalars, sizeof(RRQRState) * nj);
  RRQRQJobs<T> qj{};
  int qwgs = 0;
  for (int i = 0; i < nj; ++i) {
    const RRQRState& s = fin[i];
    RRQRTask<T>& tk = tasks[ids[i]];
    if (!s.stopped && (s.phase != 2 || s.j < s.steps)) throw std::runtime_error("rrqr: step budget exhausted");
    tk.info.rank1 = s.rank1;
    tk.info.rank2 = s.rank2;
    tk.info.rank = s.rank1 + s.rank2;
    tk.info.maxPivot = s.maxPivot;
    if (tk.info.rank > tk.Q.c) throw std::runtime_error("rrqr: Q buffer too small");
    if (tk.info.rank > 0) {
      qj.j[qj.n++] = RRQRQJob<T>{jobs.j[i].V, jobs.j[i].st + slot, tk.Q.p, tk.Q.ld, jobs.j[i].m, qwgs};
      qwgs += tk.info.rank;
    }
  }
  if (qj.n > 0) {
    hipLaunchKernelGGL(rrqr_formq_kernel<T>, dim3(qwgs), dim3(256), 0, cx.st, qj);
    HIPCHK(hipGetLastError());
  }
  for (void* w : work) cx.pool.free(w);  // stream-ordered: later users queue after formQ
}

template <typename T>
void rrqr_batched(std::vector<RRQRTask<T>>& tasks, double tol1, double tol2, Ctx& cx) {
  static_assert(sizeof(RRQRJobs<T>) + sizeof(RRQRInits) < 3072, "kernel argument block too large");
  if (RQ_MAXJ * sizeof(RRQRState) > sizeof(double) * cx.n_scalars) throw std::runtime_error("rrqr: scalar area");
  std::vector<int> ids;
  for (size_t t = 0; t < tasks.size(); ++t) {
    tasks[t].info = RRQRInfo{0, 0, 0, 0.0};
    if (tasks[t].A.r == 0 || tasks[t].A.c == 0) continue;
    ids.push_back((int)t);
    if ((int)ids.size() == RQ_MAXJ) { rrqr_chunk(tasks, ids, tol1, tol2, cx); ids.clear(); }
  }
  if (!ids.empty()) rrqr_chunk(tasks, ids, tol1, tol2, cx);
}

template <typename T>
RRQRInfo rrqr(DMat<T> A, int partition, double tol1, double tol2, double inputPivot, int fillInRank, DMat<T> Q,
              Ctx& cx) {
  std::vector<RRQRTask<T>> t(1);
  t[0].A = A;
  t[0].partition = partition;
  t[0].inputPivot = inputPivot;
  t[0].fillInRank = fillInRank;
  t[0].Q = Q;
  rrqr_batched(t, tol1, tol2, cx);
  return t[0].info;
}

template void rrqr_batched<cpx<double>>(std::vector<RRQRTask<cpx<double>>>&, double, double, Ctx&);
template void rrqr_batched<cpx<float>>(std::vector<RRQRTask<cpx<float>>>&, double, double, Ctx&);
template RRQRInfo rrqr<cpx<double>>(DMat<cpx<double>>, int, double, double, double, int, DMat<cpx<double>>, Ctx&);
template RRQRInfo rrqr<cpx<float>>(DMat<cpx<float>>, int, double, double, double, int, DMat<cpx<float>>, Ctx&);

}  // namespace rcs
