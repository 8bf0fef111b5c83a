// extern "C" boundary of librcs_amd.so (declarations and reference citations in
// include/rcs_amd.h).
#include <cstring>
#include <memory>
#include <mutex>

#include "../../include/rcs_amd.h"
#include "aca.hpp"
#include "ctx.hpp"
#include "dense.hpp"
#include "entries.hpp"
#include "geometry.hpp"
#include "mlp.hpp"
#include "refine.hpp"
#include "tree.hpp"

using namespace rcs;

struct rcs_mesh_s {
  std::shared_ptr<DeviceGeometry> g;
  bool on_device = false;  // false for rcs_mesh_create_host: host tables only
};
struct rcs_solver_s {
  std::unique_ptr<SolverBase> s;
  std::shared_ptr<DeviceGeometry> g;
};
struct rcs_operator_s {
  std::unique_ptr<SolverBase> s;  // assembled, never factorized
  std::shared_ptr<DeviceGeometry> g;
};
struct rcs_mlp_s {
  std::unique_ptr<RankMLP> m;
};
namespace {
rcs_mlp_t g_rank_model = nullptr;  // used by solvers created afterwards
int g_aca_burst = 1;
}  // namespace

namespace {
thread_local std::string g_err;
// Library contexts (stream, HBM pool, staging ring, profiler), one per lane; a host thread
// works in the lane it selected (0 by default). Objects keep the context they were created
// in, and calls on an object run on that context's stream.
constexpr int MAX_LANES = 64;
std::vector<std::unique_ptr<Ctx>> g_ctxs;
std::mutex g_mu;
thread_local int g_lane = 0;

Ctx& ctx() {
  std::lock_guard<std::mutex> lk(g_mu);
  if ((int)g_ctxs.size() <= g_lane) g_ctxs.resize(g_lane + 1);
  if (!g_ctxs[g_lane]) {
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (n <= 0) throw std::runtime_error("no HIP device");
    g_ctxs[g_lane].reset(new Ctx());
  }
  return *g_ctxs[g_lane];
}
int fail(const std::exception& e) {
  g_err = e.what();
  return -1;
}
// Argument checks, all on the host and before any device work: a bad argument becomes an
// error code and a message, never a wild pointer or an out-of-range index in a kernel.
void need(const void* p, const char* what) {
  if (!p) throw std::runtime_error(std::string(what) + " is null");
}
void check_indices(const int* idx, int n, int N, const char* what) {
  if (n > 0) need(idx, what);
  for (int i = 0; i < n; ++i)
    if (idx[i] < 0 || idx[i] >= N) throw std::runtime_error(std::string(what) + " index out of range");
}
void check_bbox(const double* bb) {
  need(bb, "bbox");
  for (int d = 0; d < 3; ++d)
    if (!(bb[2 * d + 1] >= bb[2 * d])) throw std::runtime_error("bbox must list min <= max for x, y and z");
  if (!(std::max({bb[1] - bb[0], bb[3] - bb[2], bb[5] - bb[4]}) > 0)) throw std::runtime_error("bbox is empty");
}
void need_device(rcs_mesh_t m) {
  need(m, "mesh");
  if (!m->on_device) throw std::runtime_error("mesh was created by rcs_mesh_create_host and has no device tables");
}

// Surface current at the patch centroids, getJAtCentroids (E:1100-1130): every RWG basis adds
// coef * rho * L / (2 A) to its T+ patch and the negative of it, with its own rho, to its T-
// patch (rho = centroid - free vertex). J is F x 3 complex, column-major (interleaved re, im).
void j_at_centroids(const HostGeometry& h, const double* coef, double* J) {
  const size_t F = h.v1.size();
  std::fill(J, J + 6 * F, 0.0);
  for (size_t i = 0; i < h.n1.size(); ++i) {
    const double cr = coef[2 * i], ci = coef[2 * i + 1];
    const int P = h.tp[i], M = h.tm[i];
    const double sp = h.len[i] / 2 / h.area[P], sm = -h.len[i] / 2 / h.area[M];
    const double rp[3] = {h.cx[P] - h.nx[h.fp[i]], h.cy[P] - h.ny[h.fp[i]], h.cz[P] - h.nz[h.fp[i]]};
    const double rm[3] = {h.cx[M] - h.nx[h.fm[i]], h.cy[M] - h.ny[h.fm[i]], h.cz[M] - h.nz[h.fm[i]]};
    for (int d = 0; d < 3; ++d) {
      double* jp = J + 2 * (P + d * F);
      double* jm = J + 2 * (M + d * F);
      jp[0] += cr * rp[d] * sp;
      jp[1] += ci * rp[d] * sp;
      jm[0] += cr * rm[d] * sm;
      jm[1] += ci * rm[d] * sm;
    }
  }
}

// Monostatic RCS from the centroid currents, getRCS_OnePointQuadrature (E:916-930), m^2:
// k^2 / (4 pi) * eta^2 * (|sum J_theta e^{i k.c} A|^2 + |sum J_phi e^{i k.c} A|^2), eta = 377.
void rcs_from_j(const HostGeometry& h, const double* J, const double* theta, const double* phi, int n, double* out) {
  const size_t F = h.v1.size();
  const double PI = 4.0 * std::atan(1.0);
  const double kap = 2 * PI / (3.0e8 / h.freq);
  const int neta = 377;  // E:917
  for (int a = 0; a < n; ++a) {
    const double th = theta[a], ph = phi[a];
    const double kv[3] = {kap * std::sin(th) * std::cos(ph), kap * std::sin(th) * std::sin(ph), kap * std::cos(th)};
    double str = 0, sti = 0, spr = 0, spi = 0;
    for (size_t t = 0; t < F; ++t) {
      const double ph_ = kv[0] * h.cx[t] + kv[1] * h.cy[t] + kv[2] * h.cz[t];
      const double c = std::cos(ph_), s = std::sin(ph_), A = h.area[t];
      const double* jx = J + 2 * t;
      const double* jy = J + 2 * (t + F);
      const double* jz = J + 2 * (t + 2 * F);
      const double ar = jx[0] * std::cos(th) * std::cos(ph) + jy[0] * std::cos(th) * std::sin(ph) - jz[0] * std::sin(th);
      const double ai = jx[1] * std::cos(th) * std::cos(ph) + jy[1] * std::cos(th) * std::sin(ph) - jz[1] * std::sin(th);
      str += (ar * c - ai * s) * A;
      sti += (ar * s + ai * c) * A;
      const double br = -jx[0] * std::sin(ph) + jy[0] * std::cos(ph), bi = -jx[1] * std::sin(ph) + jy[1] * std::cos(ph);
      spr += (br * c - bi * s) * A;
      spi += (br * s + bi * c) * A;
    }
    const double at = std::sqrt(str * str + sti * sti), ap = std::sqrt(spr * spr + spi * spi);
    out[a] = kap * kap / (4.0 * PI) * neta * at * neta * at + kap * kap / (4.0 * PI) * neta * ap * neta * ap;
  }
}

std::unique_ptr<rcs_mesh_s> build_mesh(const double* nodes, int n_nodes, const int* tris, int n_tris,
                                       const int* conductor, int kernel_kind, double freq, rcs_mesh_t* out) {
  if (kernel_kind != KIND_EFIE && kernel_kind != KIND_LAPLACE) throw std::runtime_error("unknown kernel kind");
  if (!(freq > 0)) throw std::runtime_error("frequency must be positive");
  if (n_nodes < 3 || n_tris < 1) throw std::runtime_error("a mesh needs at least 3 nodes and 1 triangle");
  need(nodes, "nodes");
  need(tris, "tris");
  need(out, "out");
  std::unique_ptr<rcs_mesh_s> m(new rcs_mesh_s);
  m->g = std::make_shared<DeviceGeometry>();
  m->g->h.kind = kernel_kind;
  m->g->h.freq = freq;
  m->g->h.build(nodes, n_nodes, tris, n_tris, conductor);
  return m;
}

// Pool blocks of one C-ABI call, returned to the pool when the call ends, also when it fails
// half way. Their reuse is ordered after their last use because all work runs on the
// library's one stream.
struct Blocks {
  DevPool& pool;
  std::vector<void*> held;
  explicit Blocks(DevPool& p) : pool(p) {}
  Blocks(const Blocks&) = delete;
  Blocks& operator=(const Blocks&) = delete;
  void* alloc(size_t bytes) {
    void* q = pool.alloc(bytes);
    held.push_back(q);
    return q;
  }
  ~Blocks() {
    for (void* q : held) {
      try {
        pool.free(q);
      } catch (...) {
      }
    }
  }
};

template <typename T>
DMat<T> dalloc(Blocks& b, int r, int c) {
  DMat<T> m;
  m.r = r;
  m.c = c;
  m.ld = std::max(r, 1);
  m.p = (T*)b.alloc(sizeof(T) * std::max<size_t>((size_t)r * c, 1));
  return m;
}
// host complex128 -> device T
template <typename T>
DMat<T> upload(Ctx& cx, Blocks& b, const double* h, int r, int c, int ldh);
template <>
DMat<cpx<double>> upload(Ctx& cx, Blocks& b, const double* h, int r, int c, int ldh) {
  DMat<cpx<double>> m = dalloc<cpx<double>>(b, r, c);
  if (ldh == m.ld) {
    cx.upload(m.p, h, sizeof(cpx<double>) * (size_t)r * c);
  } else {
    for (int j = 0; j < c; ++j)
      cx.upload(m.p + (size_t)j * m.ld, h + 2 * (size_t)j * ldh, sizeof(cpx<double>) * (size_t)r);
  }
  return m;
}
template <>
DMat<cpx<float>> upload(Ctx& cx, Blocks& b, const double* h, int r, int c, int ldh) {
  DMat<cpx<double>> t = upload<cpx<double>>(cx, b, h, r, c, ldh);
  DMat<cpx<float>> m = dalloc<cpx<float>>(b, r, c);
  cast_c128_to_c64(m, t, cx.st);
  return m;
}
template <typename T>
void download(Ctx& cx, Blocks& b, DMat<T> m, double* h);
template <>
void download(Ctx& cx, Blocks&, DMat<cpx<double>> m, double* h) {
  HIPCHK(hipMemcpy2DAsync(h, sizeof(cpx<double>) * m.r, m.p, sizeof(cpx<double>) * m.ld, sizeof(cpx<double>) * m.r,
                          m.c, hipMemcpyDeviceToHost, cx.st));
  cx.sync();
}
template <>
void download(Ctx& cx, Blocks& b, DMat<cpx<float>> m, double* h) {
  DMat<cpx<double>> t = dalloc<cpx<double>>(b, m.r, m.c);
  cast_c64_to_c128(t, m, cx.st);
  download<cpx<double>>(cx, b, t, h);
}

template <typename T>
void entries_t(rcs_mesh_t m, const int* rows, int nr, const int* cols, int nc, double* out) {
  Ctx& cx = ctx();
  Blocks b(cx.pool);
  DMat<T> Z = dalloc<T>(b, nr, nc);
  if (nr > 0 && nc > 0) {
    EntryTask<T> t{};
    t.rows = (const int*)cx.push(rows, sizeof(int) * nr);
    t.cols = (const int*)cx.push(cols, sizeof(int) * nc);
    t.nr = nr;
    t.nc = nc;
    t.out = Z.p;
    t.ld = Z.ld;
    t.tile0 = 0;
    const EntryTask<T>* d = (const EntryTask<T>*)cx.push(&t, sizeof(t));
    launch_entries<T>(m->g->d, m->g->h.kind, d, 1, entry_tiles(nr, nc), cx.st);
    download<T>(cx, b, Z, out);
  }
}

template <typename T>
void aca_t(rcs_mesh_t m, double tol, const int* rows, int nr, const int* cols, int nc, int* rank, int* rb, int* cb) {
  Ctx& cx = ctx();
  std::vector<AcaProblem> probs(1);
  probs[0].rows.assign(rows, rows + nr);
  probs[0].cols.assign(cols, cols + nc);
  std::vector<AcaResult> res;
  AcaStats st;
  aca_batched<T>(m->g->d, m->g->h.kind, tol, probs, res, cx, st);
  *rank = res[0].rank;
  for (int i = 0; i < res[0].rank; ++i) {
    rb[i] = probs[0].rows.empty() ? 0 : res[0].row_bases[i];
    cb[i] = res[0].col_bases[i];
  }
}

template <typename T>
void rrqr_t(const double* A, int m, int n, int tp1, int tp2, int partition, double ip, int fr, double* Q, int* rank,
            double* mp) {
  Ctx& cx = ctx();
  Blocks b(cx.pool);
  DMat<T> dA = upload<T>(cx, b, A, m, n, m);
  int mn = std::max(1, std::min(m, n));
  DMat<T> dQ = dalloc<T>(b, m, mn);
  RRQRInfo inf = rrqr(dA, partition, std::pow(10.0, -tp1), std::pow(10.0, -tp2), ip, fr, dQ, cx);
  *rank = inf.rank;
  *mp = inf.maxPivot;
  if (inf.rank > 0) download<T>(cx, b, dQ.cols(0, inf.rank), Q);
}

template <typename T>
void lu_t(const double* A, int n, const double* B, int nrhs, double* X) {
  Ctx& cx = ctx();
  Blocks b(cx.pool);
  DMat<T> dA = upload<T>(cx, b, A, n, n, n);
  DMat<T> dB = upload<T>(cx, b, B, n, nrhs, n);
  int* ip = (int*)b.alloc(sizeof(int) * 2 * std::max(n, 1));  // pivots + permutation
  getrf(dA, ip, cx);
  getrs(dA, ip, dB, cx);
  download<T>(cx, b, dB, X);
}

template <typename T>
void gemm_t(int opA, int opB, int m, int n, int k, const double* A, int lda, const double* B, int ldb, double* C) {
  Ctx& cx = ctx();
  Blocks b(cx.pool);
  int ar = opA == OP_N ? m : k, ac = opA == OP_N ? k : m;
  int br = opB == OP_N ? k : n, bc = opB == OP_N ? n : k;
  DMat<T> dA = upload<T>(cx, b, A, ar, ac, lda);
  DMat<T> dB = upload<T>(cx, b, B, br, bc, ldb);
  DMat<T> dC = dalloc<T>(b, m, n);
  std::vector<GemmOp<T>> g(1);
  g[0] = GemmOp<T>{};
  g[0].A = dA.p; g[0].lda = dA.ld; g[0].opA = opA;
  g[0].B = dB.p; g[0].ldb = dB.ld; g[0].opB = opB;
  g[0].C = dC.p; g[0].ldc = dC.ld;
  g[0].m = m; g[0].n = n; g[0].k = k;
  g[0].alpha = T(1, 0); g[0].beta = T(0, 0);
  if (k == 0) zero2d(dC, cx.st);
  else gemm_grouped(g, cx);
  download<T>(cx, b, dC, C);
}
}  // namespace

#define TRY try {
#define CATCH                              \
  }                                        \
  catch (const std::exception& e) {        \
    return fail(e);                        \
  }

extern "C" {

int rcs_init(void) {
  TRY ctx();
  return 0;
  CATCH
}
const char* rcs_last_error(void) { return g_err.c_str(); }
int rcs_set_lane(int lane) {
  TRY if (lane < 0 || lane >= MAX_LANES) throw std::runtime_error("lane out of range [0, 64)");
  // Switched off: a run of 64 scatterers in 8 lanes went quiet past its time limit and left
  // its process inside the GPU driver when it was stopped; until a multi-lane run has been
  // verified, every thread works in lane 0.
  if (lane != 0) throw std::runtime_error("concurrent lanes are switched off (lane 0 only)");
  g_lane = lane;
  return 0;
  CATCH
}

int rcs_mesh_create(const double* nodes, int n_nodes, const int* tris, int n_tris, const int* conductor,
                    int kernel_kind, double freq, rcs_mesh_t* out) {
  TRY std::unique_ptr<rcs_mesh_s> m = build_mesh(nodes, n_nodes, tris, n_tris, conductor, kernel_kind, freq, out);
  m->g->upload(ctx());
  m->on_device = true;
  *out = m.release();
  return 0;
  CATCH
}
int rcs_mesh_create_host(const double* nodes, int n_nodes, const int* tris, int n_tris, const int* conductor,
                         int kernel_kind, double freq, rcs_mesh_t* out) {
  TRY *out = build_mesh(nodes, n_nodes, tris, n_tris, conductor, kernel_kind, freq, out).release();
  return 0;
  CATCH
}
void rcs_mesh_destroy(rcs_mesh_t m) { delete m; }
int rcs_mesh_size(rcs_mesh_t m) { return m ? m->g->h.n_unknowns() : -1; }
int rcs_mesh_locations(rcs_mesh_t m, double* xyz) {
  TRY need(m, "mesh");
  need(xyz, "xyz");
  std::vector<double> v;
  m->g->h.locations(v);
  std::memcpy(xyz, v.data(), sizeof(double) * v.size());
  return 0;
  CATCH
}
int rcs_mesh_rwg(rcs_mesh_t m, int* out6, double* length) {
  TRY need(m, "mesh");
  need(out6, "out6");
  need(length, "length");
  const HostGeometry& h = m->g->h;
  for (size_t i = 0; i < h.n1.size(); ++i) {
    int* o = out6 + 6 * i;
    o[0] = h.n1[i]; o[1] = h.n2[i]; o[2] = h.tp[i]; o[3] = h.tm[i]; o[4] = h.fp[i]; o[5] = h.fm[i];
    length[i] = h.len[i];
  }
  return 0;
  CATCH
}

int rcs_entries(rcs_mesh_t m, int precision, const int* rows, int nr, const int* cols, int nc, double* out) {
  TRY need_device(m);
  const int N = m->g->h.n_unknowns();
  check_indices(rows, nr, N, "row");
  check_indices(cols, nc, N, "column");
  if (nr > 0 && nc > 0) need(out, "out");
  if (precision == PREC_C128) entries_t<cpx<double>>(m, rows, nr, cols, nc, out);
  else if (precision == PREC_C64) entries_t<cpx<float>>(m, rows, nr, cols, nc, out);
  else throw std::runtime_error("unknown precision");
  return 0;
  CATCH
}

int rcs_rhs(rcs_mesh_t m, const double* E, const double* kvec, double* out) {
  TRY need(m, "mesh");
  need(E, "E");
  need(kvec, "kvec");
  need(out, "out");
  const HostGeometry& h = m->g->h;
  if (h.kind != KIND_EFIE) throw std::runtime_error("rcs_rhs needs an EFIE mesh");
  for (size_t i = 0; i < h.n1.size(); ++i) {
    int P = h.tp[i], M = h.tm[i];
    double rp[3] = {h.cx[P] - h.nx[h.fp[i]], h.cy[P] - h.ny[h.fp[i]], h.cz[P] - h.nz[h.fp[i]]};
    double rm[3] = {h.cx[M] - h.nx[h.fm[i]], h.cy[M] - h.ny[h.fm[i]], h.cz[M] - h.nz[h.fm[i]]};
    double kp = kvec[0] * h.cx[P] + kvec[1] * h.cy[P] + kvec[2] * h.cz[P];
    double km = kvec[0] * h.cx[M] + kvec[1] * h.cy[M] + kvec[2] * h.cz[M];
    double ep = E[0] * rp[0] + E[1] * rp[1] + E[2] * rp[2];
    double em = E[0] * rm[0] + E[1] * rm[1] + E[2] * rm[2];
    double re = ep * std::cos(kp) - em * std::cos(km), im = ep * std::sin(kp) - em * std::sin(km);
    out[2 * i] = h.len[i] / 2.0 * re;
    out[2 * i + 1] = h.len[i] / 2.0 * im;
  }
  return 0;
  CATCH
}

int rcs_rcs(rcs_mesh_t m, const double* coef, const double* theta, const double* phi, int n, double* out) {
  TRY need(m, "mesh");
  need(coef, "coef");
  if (n < 0) throw std::runtime_error("n must not be negative");
  if (n > 0) { need(theta, "theta"); need(phi, "phi"); need(out, "out"); }
  const HostGeometry& h = m->g->h;
  if (h.kind != KIND_EFIE) throw std::runtime_error("rcs_rcs needs an EFIE mesh");
  std::vector<double> J(6 * h.v1.size());
  j_at_centroids(h, coef, J.data());
  rcs_from_j(h, J.data(), theta, phi, n, out);
  return 0;
  CATCH
}
int rcs_j_at_centroids(rcs_mesh_t m, const double* coef, double* J) {
  TRY need(m, "mesh");
  need(coef, "coef");
  need(J, "J");
  if (m->g->h.kind != KIND_EFIE) throw std::runtime_error("rcs_j_at_centroids needs an EFIE mesh");
  j_at_centroids(m->g->h, coef, J);
  return 0;
  CATCH
}
int rcs_rcs_from_j(rcs_mesh_t m, const double* J, const double* theta, const double* phi, int n, double* out) {
  TRY need(m, "mesh");
  need(J, "J");
  if (n < 0) throw std::runtime_error("n must not be negative");
  if (n > 0) { need(theta, "theta"); need(phi, "phi"); need(out, "out"); }
  if (m->g->h.kind != KIND_EFIE) throw std::runtime_error("rcs_rcs_from_j needs an EFIE mesh");
  rcs_from_j(m->g->h, J, theta, phi, n, out);
  return 0;
  CATCH
}
int rcs_mesh_patches(rcs_mesh_t m, double* out5, int* conductor) {
  TRY need(m, "mesh");
  need(out5, "out5");
  const HostGeometry& h = m->g->h;
  for (size_t t = 0; t < h.v1.size(); ++t) {
    double* o = out5 + 5 * t;
    o[0] = h.cx[t]; o[1] = h.cy[t]; o[2] = h.cz[t]; o[3] = h.area[t]; o[4] = h.le[t];
    if (conductor) conductor[t] = h.cond[t];
  }
  return 0;
  CATCH
}
int rcs_mesh_npatch(rcs_mesh_t m) { return m ? (int)m->g->h.v1.size() : -1; }

int rcs_aca(rcs_mesh_t m, int precision, double tol, const int* rows, int nr, const int* cols, int nc, int* rank,
            int* row_bases, int* col_bases) {
  TRY need_device(m);
  need(rank, "rank");
  if (nr <= 0 || nc <= 0) { *rank = 0; return 0; }
  const int N = m->g->h.n_unknowns();
  check_indices(rows, nr, N, "row");
  check_indices(cols, nc, N, "column");
  need(row_bases, "row_bases");
  need(col_bases, "col_bases");
  if (precision == PREC_C128) aca_t<cpx<double>>(m, tol, rows, nr, cols, nc, rank, row_bases, col_bases);
  else if (precision == PREC_C64) aca_t<cpx<float>>(m, tol, rows, nr, cols, nc, rank, row_bases, col_bases);
  else throw std::runtime_error("unknown precision");
  return 0;
  CATCH
}

int rcs_rrqr(int precision, const double* A, int m, int n, int tol_pow1, int tol_pow2, int partition,
             double input_pivot, int fillin_rank, double* Q, int* rank, double* max_pivot) {
  TRY need(rank, "rank");
  need(max_pivot, "max_pivot");
  if (m < 0 || n < 0) throw std::runtime_error("m and n must not be negative");
  if (partition < 0 || partition > n) throw std::runtime_error("partition out of range");
  if (m == 0 || n == 0) { *rank = 0; *max_pivot = 0; return 0; }
  need(A, "A");
  need(Q, "Q");
  if (precision == PREC_C128) rrqr_t<cpx<double>>(A, m, n, tol_pow1, tol_pow2, partition, input_pivot, fillin_rank, Q, rank, max_pivot);
  else if (precision == PREC_C64) rrqr_t<cpx<float>>(A, m, n, tol_pow1, tol_pow2, partition, input_pivot, fillin_rank, Q, rank, max_pivot);
  else throw std::runtime_error("unknown precision");
  return 0;
  CATCH
}

int rcs_lu_solve(int precision, const double* A, int n, const double* B, int nrhs, double* X) {
  TRY if (n < 0 || nrhs < 0) throw std::runtime_error("n and nrhs must not be negative");
  if (n == 0 || nrhs == 0) return 0;
  need(A, "A");
  need(B, "B");
  need(X, "X");
  if (precision == PREC_C128) lu_t<cpx<double>>(A, n, B, nrhs, X);
  else if (precision == PREC_C64) lu_t<cpx<float>>(A, n, B, nrhs, X);
  else throw std::runtime_error("unknown precision");
  return 0;
  CATCH
}

int rcs_gemm(int precision, int opA, int opB, int m, int n, int k, const double* A, int lda, const double* B,
             int ldb, double* C) {
  TRY if (m < 0 || n < 0 || k < 0) throw std::runtime_error("m, n and k must not be negative");
  if ((opA != 0 && opA != 1) || (opB != 0 && opB != 1)) throw std::runtime_error("op must be 0 or 1");
  if (m == 0 || n == 0) return 0;
  if (k > 0) {  // with k = 0 neither operand is read
    if (lda < (opA == OP_N ? m : k) || ldb < (opB == OP_N ? k : n))
      throw std::runtime_error("leading dimension smaller than the stored rows");
    need(A, "A");
    need(B, "B");
  }
  need(C, "C");
  if (precision == PREC_C128) gemm_t<cpx<double>>(opA, opB, m, n, k, A, lda, B, ldb, C);
  else if (precision == PREC_C64) gemm_t<cpx<float>>(opA, opB, m, n, k, A, lda, B, ldb, C);
  else throw std::runtime_error("unknown precision");
  return 0;
  CATCH
}

int rcs_solver_create(rcs_mesh_t m, const double* bbox, int leaf_size, int tol_pow, int fillin_tol, int fillin_rank,
                      int precision, int variant, rcs_solver_t* out) {
  TRY if (leaf_size <= 0) throw std::runtime_error("leaf size must be positive");
  if (variant != 0 && variant != 1) throw std::runtime_error("variant must be 0 or 1");
  if (precision != PREC_C128 && precision != PREC_C64) throw std::runtime_error("unknown precision");
  check_bbox(bbox);
  need_device(m);
  need(out, "out");
  SolverParams p;
  for (int i = 0; i < 6; ++i) p.bbox[i] = bbox[i];
  p.leafSize = leaf_size;
  p.tol_pow = tol_pow;
  p.fillin_tol = fillin_tol;
  p.fillin_rank = fillin_rank;
  p.variant = variant;
  p.rank_model = g_rank_model ? g_rank_model->m.get() : nullptr;
  p.aca_burst = g_aca_burst;
  std::unique_ptr<rcs_solver_s> s(new rcs_solver_s);
  s->g = m->g;
  s->s = make_solver(m->g, p, precision, ctx());
  s->s->assemble();
  *out = s.release();
  return 0;
  CATCH
}
int rcs_solver_rank_log(rcs_solver_t s, float* out, int cap_rows) {
  TRY need(s, "solver");
  const std::vector<float>& v = s->s->rank_log();
  const int rows = (int)(v.size() / RANK_LOG_COLS);
  if (out) std::memcpy(out, v.data(), sizeof(float) * RANK_LOG_COLS * std::min(rows, std::max(cap_rows, 0)));
  return rows;
  CATCH
}
int rcs_solver_aca_stats(rcs_solver_t s, long long* steps, long long* syncs, double* seconds) {
  TRY need(s, "solver");
  SolverStats st = s->s->stats();
  if (steps) *steps = st.aca_steps;
  if (syncs) *syncs = st.aca_syncs;
  if (seconds) *seconds = st.t_aca;
  return 0;
  CATCH
}
int rcs_mlp_create(unsigned seed, rcs_mlp_t* out) {
  TRY need(out, "out");
  std::unique_ptr<rcs_mlp_s> m(new rcs_mlp_s);
  m->m.reset(new RankMLP(ctx(), seed));
  *out = m.release();
  return 0;
  CATCH
}
void rcs_mlp_destroy(rcs_mlp_t m) {
  if (g_rank_model == m) g_rank_model = nullptr;
  delete m;
}
int rcs_mlp_get(rcs_mlp_t m, float* params, float* mean, float* scale) {
  TRY need(m, "model");
  m->m->get(params, mean, scale);
  return 0;
  CATCH
}
int rcs_mlp_set(rcs_mlp_t m, const float* params, const float* mean, const float* scale) {
  TRY need(m, "model");
  need(params, "params");
  need(mean, "mean");
  need(scale, "scale");
  m->m->set(params, mean, scale);
  return 0;
  CATCH
}
int rcs_mlp_forward(rcs_mlp_t m, const float* X, int n, float* y) {
  TRY need(m, "model");
  if (n < 0) throw std::runtime_error("n must not be negative");
  if (n > 0) { need(X, "X"); need(y, "y"); }
  m->m->forward_host(X, n, y);
  return 0;
  CATCH
}
int rcs_mlp_gradient(rcs_mlp_t m, const float* X, const float* t, int n, float* grad, double* loss) {
  TRY need(m, "model");
  if (n < 0) throw std::runtime_error("n must not be negative");
  if (n > 0) { need(X, "X"); need(t, "t"); }
  need(grad, "grad");
  double l = m->m->gradient_host(X, t, n, grad);
  if (loss) *loss = l;
  return 0;
  CATCH
}
int rcs_mlp_train(rcs_mlp_t m, const float* X, const float* t, int n, int iters, float lr, double* loss) {
  TRY need(m, "model");
  if (n < 0 || iters < 0) throw std::runtime_error("n and iters must not be negative");
  if (n > 0) { need(X, "X"); need(t, "t"); }
  double l = m->m->train_host(X, t, n, iters, lr);
  if (loss) *loss = l;
  return 0;
  CATCH
}
int rcs_set_rank_model(rcs_mlp_t m) {
  TRY g_rank_model = m;
  return 0;
  CATCH
}
int rcs_set_aca_burst(int steps) {
  TRY if (steps < 1) throw std::runtime_error("burst must be at least 1");
  g_aca_burst = steps;
  return 0;
  CATCH
}
void rcs_solver_destroy(rcs_solver_t s) { delete s; }
int rcs_solver_factorize(rcs_solver_t s) {
  TRY need(s, "solver");
  s->s->factorize();
  return 0;
  CATCH
}
int rcs_solver_solve(rcs_solver_t s, const double* rhs, int nrhs, double* x, double* x_box) {
  TRY need(s, "solver");
  if (nrhs <= 0) throw std::runtime_error("nrhs must be positive");
  need(rhs, "rhs");
  need(x, "x");
  s->s->solve(rhs, nrhs, x, x_box);
  return 0;
  CATCH
}
int rcs_operator_create(rcs_mesh_t m, const double* bbox, int leaf_size, int tol_pow, rcs_operator_t* out) {
  TRY if (leaf_size <= 0) throw std::runtime_error("leaf size must be positive");
  check_bbox(bbox);
  need_device(m);
  need(out, "out");
  SolverParams p;
  for (int i = 0; i < 6; ++i) p.bbox[i] = bbox[i];
  p.leafSize = leaf_size;
  p.tol_pow = tol_pow;
  p.fillin_tol = tol_pow;
  std::unique_ptr<rcs_operator_s> o(new rcs_operator_s);
  o->g = m->g;
  o->s = make_solver(m->g, p, RCS_PREC_C128, ctx());
  o->s->assemble();
  *out = o.release();
  return 0;
  CATCH
}
void rcs_operator_destroy(rcs_operator_t o) { delete o; }
int rcs_operator_apply(rcs_operator_t o, const double* x, int nrhs, double* y) {
  TRY need(o, "operator");
  if (nrhs <= 0) throw std::runtime_error("nrhs must be positive");
  need(x, "x");
  need(y, "y");
  Ctx& cx = o->s->context();
  Blocks b(cx.pool);
  const size_t bytes = sizeof(cpx<double>) * (size_t)o->s->size() * nrhs;
  void* dx = b.alloc(bytes);
  void* dy = b.alloc(bytes);
  cx.upload(dx, x, bytes);
  o->s->apply_device(dx, nrhs, dy);
  HIPCHK(hipMemcpyAsync(y, dy, bytes, hipMemcpyDeviceToHost, cx.st));
  cx.sync();
  return 0;
  CATCH
}
int rcs_solver_refine(rcs_solver_t s, rcs_operator_t o, const double* rhs, int nrhs, double* x, double tol, int maxit,
                      int restart, int* iters, double* relres) {
  TRY need(s, "solver");
  need(o, "operator");
  if (nrhs <= 0) throw std::runtime_error("nrhs must be positive");
  if (maxit < 0) throw std::runtime_error("maxit must not be negative");
  need(rhs, "rhs");
  need(x, "x");
  if (s->s->size() != o->s->size()) throw std::runtime_error("refine: operator and solver sizes differ");
  Ctx& cx = s->s->context();
  if (&o->s->context() != &cx) throw std::runtime_error("refine: operator and solver were created in different lanes");
  Blocks b(cx.pool);
  const size_t bytes = sizeof(cpx<double>) * (size_t)s->s->size() * nrhs;
  cpx<double>* db = (cpx<double>*)b.alloc(bytes);
  cpx<double>* dx = (cpx<double>*)b.alloc(bytes);
  cx.upload(db, rhs, bytes);
  RefineResult r = gmres_refine(*o->s, *s->s, db, nrhs, dx, tol, maxit, restart, cx);
  HIPCHK(hipMemcpyAsync(x, dx, bytes, hipMemcpyDeviceToHost, cx.st));
  cx.sync();
  if (iters) *iters = r.iters;
  if (relres) *relres = r.relres;
  return 0;
  CATCH
}
int rcs_solver_stats(rcs_solver_t s, int* info, double* vals) {
  TRY need(s, "solver");
  need(info, "info");
  need(vals, "vals");
  SolverStats st = s->s->stats();
  info[0] = st.nLevels; info[1] = st.startLevel; info[2] = (int)st.n_rrqr; info[3] = (int)st.n_compress_p2p;
  info[4] = (int)st.n_compress_p2l;
  vals[0] = st.t_assemble; vals[1] = st.t_factor; vals[2] = st.t_solve; vals[3] = st.gemm_flops;
  vals[4] = st.getrf_flops; vals[5] = st.getrs_flops; vals[6] = st.qr_flops; vals[7] = st.t_getrf;
  vals[8] = st.t_schur; vals[9] = st.t_compress; vals[10] = st.factor_bytes; vals[11] = (double)st.peak_bytes;
  return 0;
  CATCH
}
int rcs_solver_block(rcs_solver_t s, int level, int box, int which, int key, int* shape, double* out, long long cap) {
  TRY need(s, "solver");
  need(shape, "shape");
  int r = 0, c = 0;
  std::vector<double> v;
  if (!s->s->block(level, box, which, key, r, c, v)) return 1;
  shape[0] = r;
  shape[1] = c;
  if ((long long)v.size() > 2 * cap) throw std::runtime_error("block buffer too small");
  if (!v.empty()) need(out, "out");
  std::memcpy(out, v.data(), sizeof(double) * v.size());
  return 0;
  CATCH
}
int rcs_profile_enable(int on) {
  TRY ctx().prof.on = (on != 0);
  return 0;
  CATCH
}
int rcs_profile_reset(void) {
  TRY Ctx& cx = ctx();
  cx.prof.reset(cx.st);
  return 0;
  CATCH
}
int rcs_profile_get(int tag, double* work, double* ms, long long* launches) {
  TRY if (tag < 0 || tag >= PROF_NTAGS) throw std::runtime_error("unknown profile tag");
  need(work, "work");
  need(ms, "ms");
  need(launches, "launches");
  Ctx& cx = ctx();
  Profiler::Agg a = cx.prof.get(cx.st, tag);
  *work = a.work;
  *ms = a.ms;
  *launches = a.launches;
  return 0;
  CATCH
}
int rcs_solver_tree_summary(rcs_solver_t s, double* out4) {
  TRY need(s, "solver");
  need(out4, "out4");
  s->s->tree_summary(out4);
  return 0;
  CATCH
}
int rcs_solver_ranks(rcs_solver_t s, int level, int* out, int cap) {
  TRY need(s, "solver");
  std::vector<int> v;
  s->s->ranks(level, v);
  if ((int)v.size() > cap) throw std::runtime_error("rank buffer too small");
  if (!v.empty()) need(out, "out");
  std::memcpy(out, v.data(), sizeof(int) * v.size());
  return (int)v.size() / 3;
  CATCH
}

}  // extern "C"
