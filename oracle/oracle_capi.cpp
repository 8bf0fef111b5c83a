// ORACLE — test infrastructure only (see oracle_core.hpp).
//
// Plain C entry points over the FP64 restatement, loaded by the tests (ctypes) as the
// reference answer for the HIP library and timed by bench.py as the CPU baseline. Complex
// arrays are interleaved (re, im) doubles, matrices column-major.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include "oracle_tree.hpp"

using namespace oracle;

namespace {
struct OMesh {
  Mesh mesh;
};
struct OSolver {
  std::unique_ptr<Kernel> K;
  std::unique_ptr<AIFMM> S;
  double t_assemble = 0, t_factor = 0, t_solve = 0;
};
std::string g_err;
double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
// kind 0: EFIE as the reference codes it; 1: Laplace; 2: EFIE with the singular-term
// projection on the source plane (test-only, isolates the origin-plane quirk).
Kernel* makeKernel(const OMesh* m, int kind, double freq) {
  if (kind == 0 || kind == 2) {
    EFIE* e = new EFIE(&m->mesh, freq);
    e->physicalProjection = (kind == 2);
    return e;
  }
  return new Laplace(&m->mesh);
}

// Entries served from a caller's column-major m x n matrix, so that the ACA can be run on
// matrices whose cross approximation is known in closed form.
struct DenseMatrix : Kernel {
  const cd* a;
  int m, n;
  DenseMatrix(const cd* data, int rows, int cols) : a(data), m(rows), n(cols) {}
  int size() const override { return std::max(m, n); }
  cd entry(int i, int j) const override { return a[(size_t)i + (size_t)j * m]; }
};
}  // namespace

#define ORC_TRY try {
#define ORC_CATCH                       \
  }                                     \
  catch (const std::exception& e) {     \
    g_err = e.what();                   \
    return -1;                          \
  }

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }
// Threads of the oracle's dense products, LU and solves (the elimination and the solve);
// assembly uses OpenMP's own default. Returns the previous count.
int orc_set_dense_threads(int n) {
  int old = dense_threads();
  dense_threads() = n < 1 ? 1 : n;
  return old;
}

void* orc_mesh_create(const double* nodes, int nNodes, const int* tris, int nTri, const int* conductor) {
  OMesh* m = new OMesh;
  m->mesh.build(nodes, nNodes, tris, nTri, conductor);
  return m;
}
void orc_mesh_destroy(void* h) { delete (OMesh*)h; }
int orc_mesh_nrwg(void* h) { return (int)((OMesh*)h)->mesh.rwg.size(); }
int orc_mesh_npatch(void* h) { return (int)((OMesh*)h)->mesh.patches.size(); }
// per RWG: n1 n2 tp tm fp fm (node indices 1-based as in the files, patches 0-based)
void orc_mesh_rwg(void* h, int* out6, double* length) {
  const Mesh& m = ((OMesh*)h)->mesh;
  for (size_t i = 0; i < m.rwg.size(); ++i) {
    const RWG& r = m.rwg[i];
    int* o = out6 + 6 * i;
    o[0] = r.n1; o[1] = r.n2; o[2] = r.tp; o[3] = r.tm; o[4] = r.fp; o[5] = r.fm;
    length[i] = r.length;
  }
}
// per patch: centroid xyz, area, largestEdge
void orc_mesh_patches(void* h, double* out5) {
  const Mesh& m = ((OMesh*)h)->mesh;
  for (size_t i = 0; i < m.patches.size(); ++i) {
    const Patch& p = m.patches[i];
    double* o = out5 + 5 * i;
    o[0] = p.centroid.x; o[1] = p.centroid.y; o[2] = p.centroid.z; o[3] = p.area; o[4] = p.largestEdge;
  }
}

int orc_entries(void* h, int kind, double freq, const int* rows, int nr, const int* cols, int nc, double* out) {
  ORC_TRY
  std::unique_ptr<Kernel> K(makeKernel((OMesh*)h, kind, freq));
  std::vector<int> r(rows, rows + nr), c(cols, cols + nc);
  Mat M = K->getMatrix(r, c);
  std::memcpy(out, M.a.data(), sizeof(cd) * M.a.size());
  return 0;
  ORC_CATCH
}

int orc_rhs(void* h, double freq, const double* E, const double* kv, double* out) {
  ORC_TRY
  EFIE K(&((OMesh*)h)->mesh, freq);
  V3 e(E[0], E[1], E[2]), k(kv[0], kv[1], kv[2]);
  for (int i = 0; i < K.size(); ++i) {
    cd v = K.rhs(e, k, i);
    out[2 * i] = v.real();
    out[2 * i + 1] = v.imag();
  }
  return 0;
  ORC_CATCH
}

int orc_rcs(void* h, double freq, const double* coef, const double* theta, const double* phi, int n, double* out) {
  ORC_TRY
  EFIE K(&((OMesh*)h)->mesh, freq);
  std::vector<cd> c(K.size());
  std::memcpy(c.data(), coef, sizeof(cd) * c.size());
  std::vector<C3> J;
  K.jAtCentroids(c, J);
  for (int i = 0; i < n; ++i) out[i] = K.rcs(theta[i], phi[i], J);
  return 0;
  ORC_CATCH
}

// The EFIE integration primitives over source patch `tri` seen from the point tc:
// getScalarIntegrationGreensFunction (E:670-685) -> s[2] and
// getVectorIntegrationGreensFunction (E:386-446), the integral weighted by rho = r' - node fv,
// -> v[6] (x, y, z complex). kind 2 projects onto the patch's own plane (test-only).
int orc_efie_integrals(void* h, int kind, double freq, const double* tc, int tri, int fv, double* s, double* v) {
  ORC_TRY
  std::unique_ptr<Kernel> K(makeKernel((OMesh*)h, kind, freq));
  const EFIE* e = dynamic_cast<const EFIE*>(K.get());
  if (!e) throw std::runtime_error("EFIE kernel expected");
  const Mesh& m = ((OMesh*)h)->mesh;
  if (tri < 0 || tri >= (int)m.patches.size()) throw std::runtime_error("patch out of range");
  V3 t(tc[0], tc[1], tc[2]);
  cd a = e->scalarIntegration(t, tri, fv);
  C3 b = e->vectorIntegration(t, tri, fv);
  s[0] = a.real(); s[1] = a.imag();
  v[0] = b.x.real(); v[1] = b.x.imag(); v[2] = b.y.real(); v[3] = b.y.imag(); v[4] = b.z.real(); v[5] = b.z.imag();
  return 0;
  ORC_CATCH
}

// getJAtCentroids (E:1100-1130): J is n_patch x 3 complex, column-major.
int orc_j_at_centroids(void* h, double freq, const double* coef, double* J) {
  ORC_TRY
  EFIE K(&((OMesh*)h)->mesh, freq);
  std::vector<cd> c(K.size());
  std::memcpy(c.data(), coef, sizeof(cd) * c.size());
  std::vector<C3> j;
  K.jAtCentroids(c, j);
  const size_t F = j.size();
  cd* o = reinterpret_cast<cd*>(J);
  for (size_t t = 0; t < F; ++t) { o[t] = j[t].x; o[t + F] = j[t].y; o[t + 2 * F] = j[t].z; }
  return 0;
  ORC_CATCH
}

// Closed-form panel integrals of 1/R and (r'-p)/R over triangle (a, b, c) seen from tc, p
// the projection of tc on the triangle's plane
// (E:478-668), for the quadrature cross-checks.
int orc_panel(const double* tc, const double* a, const double* b, const double* c, double* lap, double* rho3) {
  V3 t(tc[0], tc[1], tc[2]), A(a[0], a[1], a[2]), Bv(b[0], b[1], b[2]), Cv(c[0], c[1], c[2]);
  *lap = analyticLaplaceTri(t, A, Bv, Cv);
  V3 r = analyticRhoLaplaceTri(t, A, Bv, Cv);
  rho3[0] = r.x; rho3[1] = r.y; rho3[2] = r.z;
  return 0;
}

// bases arrays must hold min(nr, nc) entries
int orc_aca(void* h, int kind, double freq, const int* rows, int nr, const int* cols, int nc, double tol,
            int* rank, int* row_bases, int* col_bases) {
  ORC_TRY
  std::unique_ptr<Kernel> K(makeKernel((OMesh*)h, kind, freq));
  std::vector<int> r(rows, rows + nr), c(cols, cols + nc);
  ACAResult res = aca_only_nodes2(*K, tol, r, c);
  *rank = res.rank;
  for (int i = 0; i < res.rank; ++i) { row_bases[i] = res.row_bases[i]; col_bases[i] = res.col_bases[i]; }
  return 0;
  ORC_CATCH
}

// Cost of the reference's basis-function lookup, which the restatement replaces by a table.
// EFIEMatrix keeps the edges in a std::map keyed by the sorted node pair, filled patch by
// patch (E:127-203); getMatrixEntry finds basis i and j with std::next(map.begin(), i) and
// std::next(map.begin(), j) and then looks both edges up four times (E:698-713). This times
// that lookup alone, and the restatement's whole EFIE entry, on the same `samples` random
// (i, j) pairs on one thread: seconds per entry in *walk_s and *entry_s.
int orc_map_walk_cost(void* h, double freq, int samples, unsigned seed, double* walk_s, double* entry_s) {
  ORC_TRY
  const Mesh& mesh = ((OMesh*)h)->mesh;
  struct EdgeProps {
    double length;
    std::vector<int> patches;
  };
  std::map<std::pair<int, int>, EdgeProps> edges;
  for (size_t p = 0; p < mesh.patches.size(); ++p) {
    const int v[3] = {mesh.patches[p].v1, mesh.patches[p].v2, mesh.patches[p].v3};
    for (int e = 0; e < 3; ++e) {
      const int a = v[e], b = v[(e + 1) % 3];
      auto key = std::make_pair(std::min(a, b), std::max(a, b));
      auto it = edges.find(key);
      if (it == edges.end()) {
        const V3 d = mesh.nodes[a] - mesh.nodes[b];
        EdgeProps ep;
        ep.length = std::sqrt(d.dot(d));
        ep.patches.push_back((int)p);
        edges[key] = ep;
      } else {
        it->second.patches.push_back((int)p);
      }
    }
  }
  for (auto it = edges.begin(); it != edges.end();) it = it->second.patches.size() == 1 ? edges.erase(it) : std::next(it);
  const int N = (int)edges.size();
  if (N != (int)mesh.rwg.size()) throw std::runtime_error("edge map and RWG table differ in size");
  std::vector<int> I(samples), J(samples);
  unsigned s = seed ? seed : 1u;
  for (int k = 0; k < samples; ++k) {  // xorshift32
    s ^= s << 13; s ^= s >> 17; s ^= s << 5; I[k] = (int)(s % (unsigned)N);
    s ^= s << 13; s ^= s >> 17; s ^= s << 5; J[k] = (int)(s % (unsigned)N);
  }
  double sink = 0;
  double t0 = now();
  for (int k = 0; k < samples; ++k) {
    auto ti = std::next(edges.begin(), I[k]);
    auto sj = std::next(edges.begin(), J[k]);
    const std::pair<int, int> te = ti->first, se = sj->first;
    sink += edges[te].patches[0] + edges[te].patches[1] + edges[se].patches[0] + edges[se].patches[1];
    sink += ti->second.length + sj->second.length + te.first + se.second;
  }
  *walk_s = (now() - t0) / std::max(samples, 1);
  std::unique_ptr<Kernel> K(makeKernel((OMesh*)h, 0, freq));
  cd acc = 0;
  t0 = now();
  for (int k = 0; k < samples; ++k) acc += K->entry(I[k], J[k]);
  *entry_s = (now() - t0) / std::max(samples, 1);
  if (sink == -1.0 && acc == cd(-1.0)) throw std::runtime_error("unreachable");  // keeps both loops
  return 0;
  ORC_CATCH
}

// The same ACA over all rows and columns of a column-major m x n complex matrix A.
int orc_aca_dense(const double* A, int m, int n, double tol, int* rank, int* row_bases, int* col_bases) {
  ORC_TRY
  DenseMatrix K(reinterpret_cast<const cd*>(A), m, n);
  std::vector<int> r(m), c(n);
  for (int i = 0; i < m; ++i) r[i] = i;
  for (int j = 0; j < n; ++j) c[j] = j;
  ACAResult res = aca_only_nodes2(K, tol, r, c);
  *rank = res.rank;
  for (int i = 0; i < res.rank; ++i) { row_bases[i] = res.row_bases[i]; col_bases[i] = res.col_bases[i]; }
  return 0;
  ORC_CATCH
}

// Q must hold m*min(m,n) complex values; R must hold min(m,n)*n; perm n ints.
int orc_rrqr(const double* A, int m, int n, int tolPow1, int tolPow2, int partition, double inputPivot,
             int fillInRank, double* Q, double* R, int* perm, int* rank, double* maxPivot) {
  ORC_TRY
  Mat M(m, n);
  std::memcpy(M.a.data(), A, sizeof(cd) * M.a.size());
  RRQRResult r = rrqr(M, tolPow1, tolPow2, partition, inputPivot, fillInRank);
  *rank = r.rank;
  *maxPivot = r.maxPivot;
  std::memcpy(Q, r.Q.a.data(), sizeof(cd) * r.Q.a.size());
  if (R) std::memcpy(R, r.R.a.data(), sizeof(cd) * r.R.a.size());
  if (perm)
    for (int i = 0; i < n; ++i) perm[i] = r.perm[i];
  return 0;
  ORC_CATCH
}

// Dense partial-pivoting LU solve of a caller-supplied matrix (n x n) with nrhs columns.
int orc_lu_solve(const double* A, int n, const double* B, int nrhs, double* X) {
  ORC_TRY
  Mat M(n, n), Bm(n, nrhs);
  std::memcpy(M.a.data(), A, sizeof(cd) * M.a.size());
  std::memcpy(Bm.a.data(), B, sizeof(cd) * Bm.a.size());
  LU lu(M);
  Mat Xm = lu.solve(Bm);
  std::memcpy(X, Xm.a.data(), sizeof(cd) * Xm.a.size());
  return 0;
  ORC_CATCH
}

// E:831-837: dense assembly + PartialPivLU solve.
int orc_dense_solve(void* h, int kind, double freq, const double* rhs, double* x) {
  ORC_TRY
  std::unique_ptr<Kernel> K(makeKernel((OMesh*)h, kind, freq));
  int N = K->size();
  std::vector<int> idx(N);
  for (int i = 0; i < N; ++i) idx[i] = i;
  Mat Z = K->getMatrix(idx, idx);
  Mat b(N, 1);
  std::memcpy(b.a.data(), rhs, sizeof(cd) * N);
  LU lu(Z);
  Mat xx = lu.solve(b);
  std::memcpy(x, xx.a.data(), sizeof(cd) * N);
  return 0;
  ORC_CATCH
}

// y = Z x with entries generated on the fly (S:331-338 backward-error operator).
int orc_matvec(void* h, int kind, double freq, const double* x, double* y) {
  ORC_TRY
  std::unique_ptr<Kernel> K(makeKernel((OMesh*)h, kind, freq));
  int N = K->size();
  const cd* xx = (const cd*)x;
  cd* yy = (cd*)y;
#pragma omp parallel for schedule(dynamic, 8)
  for (int i = 0; i < N; ++i) {
    cd s = 0;
    for (int j = 0; j < N; ++j) s += K->entry(i, j) * xx[j];
    yy[i] = s;
  }
  return 0;
  ORC_CATCH
}

void* orc_aifmm_create(void* h, int kind, double freq, const double* loc, int N, const double* bbox, int leafSize,
                       int tol_pow, int fillin_tol, int fillin_rank, int variant) {
  try {
    OSolver* s = new OSolver;
    s->K.reset(makeKernel((OMesh*)h, kind, freq));
    SolverInputs in;
    in.xmin = bbox[0]; in.xmax = bbox[1]; in.ymin = bbox[2]; in.ymax = bbox[3]; in.zmin = bbox[4]; in.zmax = bbox[5];
    in.leafSize = leafSize;
    in.tol_pow = tol_pow;
    in.fillin_tol = fillin_tol;
    in.fillin_rank = fillin_rank;
    in.variant = variant;
    for (int i = 0; i < N; ++i) in.locations.push_back(V3(loc[3 * i], loc[3 * i + 1], loc[3 * i + 2]));
    double t0 = now();
    s->S.reset(new AIFMM(in, s->K.get()));
    s->t_assemble = now() - t0;
    return s;
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}
void orc_aifmm_destroy(void* h) { delete (OSolver*)h; }
int orc_aifmm_factorize(void* h) {
  ORC_TRY
  OSolver* s = (OSolver*)h;
  s->S->A->ctr = Counters();
  double t0 = now();
  s->S->factorize();
  s->t_factor = now() - t0;
  return 0;
  ORC_CATCH
}
// x in the input order (getPhi); x_box (optional) is the solve() result in box order.
int orc_aifmm_solve(void* h, const double* rhs, double* x, double* x_box) {
  ORC_TRY
  OSolver* s = (OSolver*)h;
  int N = s->S->A->N;
  std::vector<cd> b(N);
  std::memcpy(b.data(), rhs, sizeof(cd) * N);
  s->S->backSubstitute1(b);
  double t0 = now();
  s->S->backSubstitute2();
  s->S->backSubstitute3();
  s->t_solve = now() - t0;
  std::vector<cd> xb = s->S->A->getx();
  if (x_box) std::memcpy(x_box, xb.data(), sizeof(cd) * N);
  std::vector<cd> p = s->S->getPhi();
  std::memcpy(x, p.data(), sizeof(cd) * N);
  return 0;
  ORC_CATCH
}
// info: nLevels, startLevel, n_rrqr, n_compress_p2p, n_compress_p2l
// times: assemble, factor, solve; flops: gemm, getrf, getrs, qr
int orc_aifmm_info(void* h, int* info, double* times, double* flops) {
  OSolver* s = (OSolver*)h;
  Tree* t = s->S->A;
  info[0] = t->nLevels;
  info[1] = t->startLevel;
  info[2] = (int)t->ctr.n_rrqr;
  info[3] = (int)t->ctr.n_compress_p2p;
  info[4] = (int)t->ctr.n_compress_p2l;
  times[0] = s->t_assemble; times[1] = s->t_factor; times[2] = s->t_solve;
  flops[0] = t->ctr.gemm_flops; flops[1] = t->ctr.getrf_flops; flops[2] = t->ctr.getrs_flops; flops[3] = t->ctr.qr_flops;
  return 0;
}
// getAvgLeafSize, getAvgMultipoles, findMemory, findMemoryAfterFactorize of the current tree.
int orc_aifmm_tree_summary(void* h, double* out4) {
  Tree* t = ((OSolver*)h)->S->A;
  out4[0] = t->getAvgLeafSize();
  out4[1] = t->getAvgMultipoles();
  out4[2] = t->findMemory();
  out4[3] = t->findMemoryAfterFactorize();
  return 0;
}
// Per-box ranks at level j: NumMultipoles, NumLocals, particle count (3*nBoxes ints).
int orc_aifmm_ranks(void* h, int j, int* out) {
  Tree* t = ((OSolver*)h)->S->A;
  if (j < 0 || j > t->nLevels) return -1;
  for (int k = 0; k < t->nBoxesPerLevel[j]; ++k) {
    out[3 * k] = t->B(j, k).NumMultipoles;
    out[3 * k + 1] = t->B(j, k).NumLocals;
    out[3 * k + 2] = (int)t->B(j, k).chargeLocations.size();
  }
  return t->nBoxesPerLevel[j];
}

}  // extern "C"

extern "C" int orc_aifmm_block(void* h, int level, int box, int which, int key, int* shape, double* out,
                               long long cap) {
  Tree* t = ((OSolver*)h)->S->A;
  if (level < 0 || level > t->nLevels || box < 0 || box >= t->nBoxesPerLevel[level]) return 1;
  Box& b = t->B(level, box);
  const Mat* m = nullptr;
  auto find = [&](std::map<int, Mat>& mp) -> const Mat* {
    auto it = mp.find(key);
    return it == mp.end() ? nullptr : &it->second;
  };
  switch (which) {
    case 0: m = &b.L2P; break;
    case 1: m = &b.P2M; break;
    case 2: m = find(b.M2L); break;
    case 3: m = find(b.P2P); break;
    case 4: m = find(b.M2P); break;
    case 5: m = find(b.P2L); break;
    case 6: m = &b.L2M_f; break;
    case 7: m = &b.particle_rhs; break;
    case 8: m = &b.multipole_rhs; break;
    case 9: m = &b.local_rhs; break;
    case 10: m = &b.particles; break;
    case 11: m = &b.multipoles; break;
    case 12: m = &b.locals; break;
    case 13: m = find(b.L2P_f); break;
    default: return 1;
  }
  if (!m) return 1;
  shape[0] = m->r;
  shape[1] = m->c;
  if ((long long)m->a.size() > cap) return -1;
  std::memcpy(out, m->a.data(), sizeof(cd) * m->a.size());
  return 0;
}
