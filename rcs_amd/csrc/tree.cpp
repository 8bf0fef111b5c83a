// GPU AIFMM tree: see tree.hpp. Function names and the order of operations follow
// include/AIFMMTree2D_3D_memory_optimized.hpp (cited as T:line) so that the elimination
// takes the same decisions (ranks, fill-in pairs, compression order) as the reference.
#include "tree.hpp"

#include "mlp.hpp"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <unordered_map>

namespace rcs {

static double wall() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void DeviceGeometry::upload(Ctx& c) {
  cx = &c;
  auto up = [&](const void* src, size_t bytes) -> void* {
    void* p = nullptr;
    HIPCHK(hipMalloc(&p, bytes ? bytes : 16));
    if (bytes) HIPCHK(hipMemcpy(p, src, bytes, hipMemcpyHostToDevice));
    bufs.push_back(p);
    return p;
  };
  d.nx = (const double*)up(h.nx.data(), sizeof(double) * h.nx.size());
  d.ny = (const double*)up(h.ny.data(), sizeof(double) * h.ny.size());
  d.nz = (const double*)up(h.nz.data(), sizeof(double) * h.nz.size());
  d.v1 = (const int*)up(h.v1.data(), sizeof(int) * h.v1.size());
  d.v2 = (const int*)up(h.v2.data(), sizeof(int) * h.v2.size());
  d.v3 = (const int*)up(h.v3.data(), sizeof(int) * h.v3.size());
  d.cx = (const double*)up(h.cx.data(), sizeof(double) * h.cx.size());
  d.cy = (const double*)up(h.cy.data(), sizeof(double) * h.cy.size());
  d.cz = (const double*)up(h.cz.data(), sizeof(double) * h.cz.size());
  d.area = (const double*)up(h.area.data(), sizeof(double) * h.area.size());
  d.le = (const double*)up(h.le.data(), sizeof(double) * h.le.size());
  d.tp = (const int*)up(h.tp.data(), sizeof(int) * h.tp.size());
  d.tm = (const int*)up(h.tm.data(), sizeof(int) * h.tm.size());
  d.fp = (const int*)up(h.fp.data(), sizeof(int) * h.fp.size());
  d.fm = (const int*)up(h.fm.data(), sizeof(int) * h.fm.size());
  d.len = (const double*)up(h.len.data(), sizeof(double) * h.len.size());
  // include/includes.hpp:53-63 with freq as a runtime value
  const double PI = 4.0 * std::atan(1.0);
  const double eps = 8.854187817e-12, c0 = 3.0e8, mu = 4 * PI * 1e-7;
  d.eps = eps;
  d.mu = mu;
  d.omega = 2.0 * PI * h.freq;
  d.kappa = 2 * PI / (c0 / h.freq);
  d.C0 = 1.0 / (4 * PI * eps);
}
void DeviceGeometry::release() {
  for (void* p : bufs) (void)hipFree(p);
  bufs.clear();
}

namespace {

template <typename T>
struct Box {
  bool ILActive = true;
  // L2P has orthonormal columns / P2M orthonormal rows (set when they come from an RRQR's Q)
  bool orthL2P = false, orthP2M = false;
  double maxPivot_L2P = 0.0, maxPivot_P2M = 0.0;
  double cx = 0, cy = 0, cz = 0;
  std::vector<int> IL, NB;
  bool Eliminated = false;
  int NumMultipoles = 0, NumLocals = 0;
  std::map<int, DMat<T>> M2L, P2P, M2P, P2L, L2P_f, P2M_f, M2M_f;
  DMat<T> L2M_f, L2P, P2M, Ac, Ar;
  std::vector<int> charge, in_charge, in_check, out_charge, out_check;
  DMat<T> luP2P, luL2M;
  int* ipP2P = nullptr;
  int* ipL2M = nullptr;
  DMat<T> invP2P, invL2M;  // inverses of the factored blocks, formed for repeated solves
  // solve vectors (n x nrhs)
  DMat<T> particles, multipoles, locals, particle_rhs, multipole_rhs, local_rhs;
};

template <typename T>
class Solver : public SolverBase {
 public:
  using R = decltype(T().re);
  Solver(std::shared_ptr<DeviceGeometry> g, const SolverParams& p, Ctx& c) : G(g), P(p), cx(c) {}
  // Box blocks are plain views; the solver owns every pool block it allocated and hands
  // all of them back to the pool when it goes away, so a process can build solvers one
  // after another without holding the earlier ones' HBM (the next solver reuses the cached
  // blocks; the pool returns them to the driver when it nears its limit).
  ~Solver() override {
    cx.sync();
    for (auto& kv : owned_) cx.pool.free(kv.first);
    owned_.clear();
  }

  // ------------------------------------------------------------ block helpers
  void* palloc(size_t bytes) {
    void* p = cx.pool.alloc(bytes);
    owned_[p] = bytes;
    own_in_use_ += bytes;
    own_peak_ = std::max(own_peak_, own_in_use_);
    return p;
  }
  void pfree(void* p) {
    if (!p) return;
    auto it = owned_.find(p);
    if (it == owned_.end()) throw std::runtime_error("solver: free of a block it does not own");
    own_in_use_ -= it->second;
    owned_.erase(it);
    cx.pool.free(p);
  }
  DMat<T> newmat(int r, int c) {
    DMat<T> m;
    m.r = r;
    m.c = c;
    m.ld = std::max(r, 1);
    if ((size_t)r * c > 0) m.p = (T*)palloc(sizeof(T) * (size_t)r * c);
    return m;
  }
  void drop(DMat<T>& m) {
    if (m.p) pfree(m.p);
    m = DMat<T>();
  }
  void put(DMat<T>& dst, DMat<T> src) {
    if (dst.p && dst.p != src.p) pfree(dst.p);
    dst = src;
  }
  DMat<T> zeros(int r, int c) {
    DMat<T> m = newmat(r, c);
    zero2d(m, cx.st);
    return m;
  }
  DMat<T> clone(const DMat<T>& a) {
    DMat<T> m = newmat(a.r, a.c);
    copy2d(m, a, cx.st);
    return m;
  }
  struct Group {
    std::vector<GemmOp<T>> ops;
  };
  // C = alpha*op(A)*op(B) + beta*C. chain = 1 sums into the same C as the previous op
  // of the group inside one output tile (several products into one block).
  void add(Group& g, DMat<T> C, DMat<T> A, int opA, DMat<T> B, int opB, T alpha, T beta, int chain = 0) {
    int m = opA == OP_N ? A.r : A.c, k = opA == OP_N ? A.c : A.r;
    int kb = opB == OP_N ? B.r : B.c, n = opB == OP_N ? B.c : B.r;
    if (k != kb || C.r != m || C.c != n) throw std::runtime_error("gemm: shape mismatch");
    if ((size_t)m * n == 0) return;
    stats_.gemm_flops += 8.0 * m * (double)n * k;
    if (k == 0) {  // empty inner dimension: C = beta*C
      if (beta.re == R(0) && beta.im == R(0)) zero2d(C, cx.st);
      return;
    }
    GemmOp<T> o{};
    o.A = A.p; o.lda = A.ld; o.opA = opA;
    o.B = B.p; o.ldb = B.ld; o.opB = opB;
    o.C = C.p; o.ldc = C.ld;
    o.m = m; o.n = n; o.k = k;
    o.alpha = alpha; o.beta = beta;
    o.chain = chain;
    g.ops.push_back(o);
  }
  void run(Group& g, int tag = -1) {
    gemm_grouped(g.ops, cx, tag == PROF_SCHUR ? 1 : 0, tag);
    g.ops.clear();
  }
  DMat<T> mul(DMat<T> A, int opA, DMat<T> B, int opB, T alpha = T(1, 0)) {
    int m = opA == OP_N ? A.r : A.c, n = opB == OP_N ? B.c : B.r;
    DMat<T> C = newmat(m, n);
    Group g;
    add(g, C, A, opA, B, opB, alpha, T(0, 0));
    run(g);
    return C;
  }
  // X = X - prod into an existing block, or a new block -prod when X is empty (T:3936-3947)
  void assign_or_sub(Group& g, DMat<T>& X, DMat<T> A, int opA, DMat<T> B, int opB) {
    int m = opA == OP_N ? A.r : A.c, n = opB == OP_N ? B.c : B.r;
    if (X.size() == 0) {
      put(X, newmat(m, n));
      add(g, X, A, opA, B, opB, T(-1, 0), T(0, 0));
    } else {
      add(g, X, A, opA, B, opB, T(-1, 0), T(1, 0));
    }
  }
  void factor(DMat<T> A, DMat<T>& lu, int*& ip) {
    put(lu, clone(A));
    if (ip) pfree(ip);
    ip = (int*)palloc(sizeof(int) * 2 * std::max(A.r, 1));  // pivots + permutation
    double t0 = wall();
    int h = cx.prof.begin(cx.st);
    getrf(lu, ip, cx);
    cx.prof.end(cx.st, h, PROF_GETRF, 8.0 / 3.0 * A.r * (double)A.r * A.r);
    stats_.t_getrf += wall() - t0;
    stats_.getrf_flops += 8.0 / 3.0 * A.r * (double)A.r * A.r;
  }
  void solve_inplace(DMat<T> lu, const int* ip, DMat<T> B) {
    stats_.getrs_flops += 8.0 * lu.r * (double)lu.r * B.c;
    int h = cx.prof.begin(cx.st);
    getrs(lu, ip, B, cx);
    cx.prof.end(cx.st, h, PROF_GETRS, 8.0 * lu.r * (double)lu.r * B.c);
  }
  std::vector<double> norms(const std::vector<DMat<T>>& v) {
    std::vector<double> out(v.size(), 0.0);
    if (v.empty()) return out;
    sqnorms(v, cx.d_scalars, cx);
    const double* h = cx.fetch_scalars(v.size());
    for (size_t i = 0; i < v.size(); ++i) out[i] = std::sqrt(h[i]);
    return out;
  }
  // Index lists travel through the pinned staging ring (stream ordered, no allocation).
  int* upload_ints(const std::vector<int>& v) {
    if (v.empty()) return nullptr;
    return (int*)cx.push(v.data(), sizeof(int) * v.size());
  }
  // A^{-1} rhs as a new block: a product with the stored inverse when repeated solves were
  // prepared, else triangular solves with the LU factors.
  DMat<T> solved(DMat<T> lu, const int* ip, DMat<T> inv, DMat<T> rhs) {
    if (inv.p) {
      stats_.getrs_flops += 8.0 * inv.r * (double)inv.c * rhs.c;
      return mul(inv, OP_N, rhs, OP_N);
    }
    DMat<T> s = clone(rhs);
    solve_inplace(lu, ip, s);
    return s;
  }
  DMat<T> inverse_of(DMat<T> lu, const int* ip) {
    DMat<T> inv = newmat(lu.r, lu.r);
    fill_identity(inv, T(1, 0), cx.st);
    solve_inplace(lu, ip, inv);
    return inv;
  }
  bool inverses_ready_ = false;
  void prepare_repeated_solves() override {
    if (!factorized_) throw std::runtime_error("prepare_repeated_solves before factorize");
    if (inverses_ready_) return;
    for (int j = startLevel; j <= nLevels; ++j)
      for (int k = 0; k < nBoxes[j]; ++k) {
        Box<T>& b = tree[j][k];
        if (b.luP2P.p && b.luP2P.r > 0) put(b.invP2P, inverse_of(b.luP2P, b.ipP2P));
        if (b.luL2M.p && b.luL2M.r > 0) put(b.invL2M, inverse_of(b.luL2M, b.ipL2M));
      }
    if (luBase.p && luBase.r > 0) put(invBase, inverse_of(luBase, ipBase));
    inverses_ready_ = true;
  }
  void free_tmp_ints() {}
  // Blocks Z(rows_i, cols_i) for many pairs in one launch.
  void entries(const std::vector<const std::vector<int>*>& rows, const std::vector<const std::vector<int>*>& cols,
               std::vector<DMat<T>*>& outs) {
    std::vector<EntryTask<T>> tasks;
    int tiles = 0;
    size_t base = cx.pushed;
    auto flush = [&]() {
      if (tasks.empty()) return;
      double ne = 0;
      for (auto& t : tasks) ne += (double)t.nr * t.nc;
      const EntryTask<T>* d = (const EntryTask<T>*)cx.push(tasks.data(), sizeof(EntryTask<T>) * tasks.size());
      int h = cx.prof.begin(cx.st);
      launch_entries<T>(G->d, G->h.kind, d, (int)tasks.size(), tiles, cx.st);
      cx.prof.end(cx.st, h, PROF_ENTRIES, ne);
      tasks.clear();
      tiles = 0;
      base = cx.pushed;
    };
    for (size_t i = 0; i < rows.size(); ++i) {
      int nr = (int)rows[i]->size(), nc = (int)cols[i]->size();
      put(*outs[i], newmat(nr, nc));
      if (nr == 0 || nc == 0) continue;
      // the index lists of one launch must all stay live in the staging ring
      if (cx.pushed - base > cx.cap / 4 || tiles > (1 << 26)) flush();
      EntryTask<T> t{};
      t.rows = upload_ints(*rows[i]);
      t.cols = upload_ints(*cols[i]);
      t.nr = nr;
      t.nc = nc;
      t.out = outs[i]->p;
      t.ld = outs[i]->ld;
      t.tile0 = tiles;
      tiles += entry_tiles(nr, nc);
      tasks.push_back(t);
    }
    flush();
  }

  // ------------------------------------------------------------ tree set-up (T:52-331, 425-668)
  void setup() {
    const HostGeometry& h = G->h;
    N = h.n_unknowns();
    std::vector<double> loc;
    h.locations(loc);
    px.resize(N); py.resize(N); pz.resize(N);
    for (int i = 0; i < N; ++i) { px[i] = loc[3 * i]; py[i] = loc[3 * i + 1]; pz[i] = loc[3 * i + 2]; }
    gcx = gcy = gcz = 0;
    for (int i = 0; i < N; ++i) { gcx += px[i]; gcy += py[i]; gcz += pz[i]; }
    if (N > 0) { gcx /= N; gcy /= N; gcz /= N; }
    TOL_POW = P.tol_pow;
    fill_in_TOL_POW = P.fillin_tol;
    fillInTolerance = std::pow(10.0, -1.0 * TOL_POW);
    const double* bb = P.bbox;
    L = std::max({bb[1] - bb[0], bb[3] - bb[2], bb[5] - bb[4]});
    int q = N / P.leafSize;
    int lv = q > 0 ? (int)std::floor(std::log((double)q) / std::log(8.0)) : 0;
    if (P.variant == 0) { nLevels = std::max(2, lv); eta = std::sqrt(3.0) * 2; }
    else { nLevels = std::max(3, lv); eta = std::sqrt(3.0); }
    nBoxes.assign(1, 1);
    side.assign(1, L);
    for (int j = 1; j <= nLevels; ++j) { nBoxes.push_back(8 * nBoxes[j - 1]); side.push_back(0.5 * side[j - 1]); }
    for (int j = 0; j < 32; ++j) rankVec[j] = 0;
    int temp = 5, i = 0;
    for (int j = nLevels; j >= 2; --j) {
      rankVec[j] = temp;
      ++i;
      if (i % 3 == 0) temp += 5;
    }
    if (P.variant == 1)
      for (int j = 0; j < 32; ++j) rankVec[j] = 1 << 28;
    tree.assign(nLevels + 1, std::vector<Box<T>>());
    for (int j = 0; j <= nLevels; ++j) tree[j].resize(nBoxes[j]);
    tree[0][0].cx = (bb[0] + bb[1]) / 2.0;
    tree[0][0].cy = (bb[2] + bb[3]) / 2.0;
    tree[0][0].cz = (bb[4] + bb[5]) / 2.0;
    static const int sx[8] = {-1, 1, 1, -1, -1, 1, 1, -1}, sy[8] = {-1, -1, 1, 1, -1, -1, 1, 1},
                     sz[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
    for (int j = 0; j < nLevels; ++j) {
      double sh = 0.5 * side[j + 1];
      for (int k = 0; k < nBoxes[j]; ++k)
        for (int l = 0; l < 8; ++l) {
          Box<T>& c = tree[j + 1][8 * k + l];
          c.cx = tree[j][k].cx + sx[l] * sh;
          c.cy = tree[j][k].cy + sy[l] * sh;
          c.cz = tree[j][k].cz + sz[l] * sh;
        }
    }
    tree[0][0].NB.push_back(0);
    for (int j = 0; j < nLevels; ++j)
      for (int k = 0; k < nBoxes[j]; ++k)
        for (int ci = 0; ci < 8; ++ci) {
          int child = 8 * k + ci;
          Box<T>& cb = tree[j + 1][child];
          for (int nn : tree[j][k].NB)
            for (int c = 0; c < 8; ++c) {
              int other = nn * 8 + c;
              const Box<T>& ob = tree[j + 1][other];
              double dx = cb.cx - ob.cx, dy = cb.cy - ob.cy, dz = cb.cz - ob.cz;
              double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
              double etaL = (j + 1 == nLevels) ? eta / 2.0 : eta;  // T:281-284
              if (dist > etaL * side[j + 1] + side[nLevels] / 2.0) cb.IL.push_back(other);
              else cb.NB.push_back(other);
            }
        }
    startLevel = -1;
    for (int j = 0; j <= nLevels && startLevel < 0; ++j)
      for (int k = 0; k < nBoxes[j]; ++k)
        if (!tree[j][k].IL.empty()) { startLevel = j; break; }
    if (startLevel < 0) startLevel = nLevels;
    // T:576-625: "<=" split against the box centre
    for (int p = 0; p < N; ++p) tree[0][0].charge.push_back(p);
    for (int j = 0; j < nLevels; ++j)
      for (int k = 0; k < nBoxes[j]; ++k) {
        const Box<T>& b = tree[j][k];
        for (int idx : b.charge) {
          int ch;
          if (pz[idx] <= b.cz) {
            if (px[idx] <= b.cx) ch = (py[idx] <= b.cy) ? 0 : 3;
            else ch = (py[idx] <= b.cy) ? 1 : 2;
          } else {
            if (px[idx] <= b.cx) ch = (py[idx] <= b.cy) ? 4 : 7;
            else ch = (py[idx] <= b.cy) ? 5 : 6;
          }
          tree[j + 1][8 * k + ch].charge.push_back(idx);
        }
      }
    for (int j = nLevels - 1; j >= 0; --j)
      for (int k = 0; k < nBoxes[j]; ++k) {
        auto& cl = tree[j][k].charge;
        cl.clear();
        for (int c = 0; c < 8; ++c) {
          auto& s = tree[j + 1][8 * k + c].charge;
          cl.insert(cl.end(), s.begin(), s.end());
        }
      }
    stats_.nLevels = nLevels;
    stats_.startLevel = startLevel;
  }

  // ------------------------------------------------------------ skeletons (T:681-760, 917-1256)
  void childPoints(int j, int k, bool inCheck, std::vector<int>& out) {
    if (j == nLevels) {
      out.insert(out.end(), tree[j][k].charge.begin(), tree[j][k].charge.end());
      return;
    }
    for (int c = 0; c < 8; ++c) {
      Box<T>& ch = tree[j + 1][8 * k + c];
      auto& v = inCheck ? ch.in_check : ch.out_charge;
      out.insert(out.end(), v.begin(), v.end());
    }
  }
  // The rank predictor's inputs for box k at level j (an ACA problem with `own` points on
  // the box side): mean interaction-list centre distance / box side, kappa * box side,
  // log10 tolerance, level, direction cosines of the box centre about the point cloud's
  // centroid, log2(1 + own).
  void rank_features(int j, int k, int own, float* f) {
    const Box<T>& b = tree[j][k];
    double d = 0;
    for (int l : b.IL) {
      const Box<T>& o = tree[j][l];
      d += std::sqrt((b.cx - o.cx) * (b.cx - o.cx) + (b.cy - o.cy) * (b.cy - o.cy) + (b.cz - o.cz) * (b.cz - o.cz));
    }
    f[0] = b.IL.empty() ? 0.0f : (float)(d / b.IL.size() / side[j]);
    f[1] = G->h.kind == KIND_EFIE ? (float)(G->d.kappa * side[j]) : 0.0f;
    f[2] = (float)(-TOL_POW);
    f[3] = (float)j;
    double ux = b.cx - gcx, uy = b.cy - gcy, uz = b.cz - gcz, un = std::sqrt(ux * ux + uy * uy + uz * uz);
    if (un > 0) { ux /= un; uy /= un; uz /= un; }
    f[4] = (float)ux; f[5] = (float)uy; f[6] = (float)uz;
    f[7] = (float)std::log2(1.0 + own);
  }
  // Adaptive ACA of one level's problems (incoming: inout 0, outgoing: 1), with predicted
  // ranks when a model is set, and the rank log.
  void run_aca(int j, int inout, std::vector<AcaProblem>& probs, const std::vector<int>& which,
               std::vector<AcaResult>& res) {
    const double tol = std::pow(10.0, -TOL_POW);
    std::vector<float> feat(which.size() * MLP_IN);
    for (size_t i = 0; i < which.size(); ++i) {
      const AcaProblem& p = probs[which[i]];
      rank_features(j, which[i], (int)(inout == 0 ? p.rows.size() : p.cols.size()), &feat[i * MLP_IN]);
    }
    std::vector<int> pred(which.size(), 0);
    if (P.rank_model && !which.empty()) {
      P.rank_model->predict_ranks(feat.data(), (int)which.size(), pred.data());
      for (size_t i = 0; i < which.size(); ++i) {  // a rank never exceeds the short side
        const AcaProblem& p = probs[which[i]];
        pred[i] = std::max(1, std::min(pred[i], (int)std::min(p.rows.size(), p.cols.size())));
        probs[which[i]].pred = pred[i];
      }
    }
    AcaStats ast;
    ast.burst = P.aca_burst;
    double t0 = wall();
    int h = cx.prof.begin(cx.st);
    aca_batched<T>(G->d, G->h.kind, tol, probs, res, cx, ast);
    cx.prof.end(cx.st, h, PROF_ACA, 0.0);
    stats_.t_aca += wall() - t0;
    stats_.aca_steps += ast.steps;
    stats_.aca_syncs += ast.syncs;
    if (verbose)
      std::fprintf(stderr, "[rcs] aca level %d dir %d: %lld steps, %.3f s, waiting %.3f s, growing %.3f s\n", j, inout,
                   ast.steps, wall() - t0, ast.wait_s, ast.grow_s);
    for (size_t i = 0; i < which.size(); ++i) {
      rank_log_.insert(rank_log_.end(), feat.begin() + i * MLP_IN, feat.begin() + (i + 1) * MLP_IN);
      rank_log_.push_back((float)res[which[i]].rank);
      rank_log_.push_back((float)pred[i]);
      rank_log_.push_back((float)j);
      rank_log_.push_back((float)inout);
      const AcaProblem& p = probs[which[i]];
      rank_log_.push_back((float)std::min(p.rows.size(), p.cols.size()));
      rank_log_.push_back((float)std::max(p.rows.size(), p.cols.size()));
    }
  }
  const std::vector<float>& rank_log() const override { return rank_log_; }

  void getNodes() {
    for (int j = nLevels; j >= startLevel; --j) {
      const int nb = nBoxes[j];
      // incoming (T:1155-1256)
      {
        std::vector<AcaProblem> probs(nb);
        std::vector<int> which;
        for (int k = 0; k < nb; ++k) {
          Box<T>& b = tree[j][k];
          childPoints(j, k, true, probs[k].rows);
          for (int kIL : b.IL) childPoints(j, kIL, false, probs[k].cols);
          if (!probs[k].rows.empty() && !probs[k].cols.empty()) which.push_back(k);
          else if (probs[k].cols.empty()) { b.ILActive = false; b.in_check = probs[k].rows; }
          if (probs[k].rows.empty() || probs[k].cols.empty()) probs[k].rows.clear(), probs[k].cols.clear();
        }
        std::vector<AcaResult> res;
        run_aca(j, 0, probs, which, res);
        trace("level %d incoming skeletons (%d boxes)", j, (int)which.size());
        std::vector<const std::vector<int>*> rr, cc;
        std::vector<DMat<T>*> outs;
        std::vector<std::vector<int>> pivcols(nb);
        for (int k : which) {
          Box<T>& b = tree[j][k];
          for (int r : res[k].row_bases) b.in_check.push_back(probs[k].rows.at(r));
          for (int c : res[k].col_bases) { b.in_charge.push_back(probs[k].cols.at(c)); pivcols[k].push_back(probs[k].cols.at(c)); }
          rr.push_back(&probs[k].rows);
          cc.push_back(&pivcols[k]);
          outs.push_back(&b.Ac);  // Ac = K(rows, pivot columns)
        }
        entries(rr, cc, outs);
      }
      // outgoing (T:994-1077)
      {
        std::vector<AcaProblem> probs(nb);
        std::vector<int> which;
        for (int k = 0; k < nb; ++k) {
          Box<T>& b = tree[j][k];
          childPoints(j, k, false, probs[k].cols);
          for (int kIL : b.IL) childPoints(j, kIL, true, probs[k].rows);
          if (!probs[k].rows.empty() && !probs[k].cols.empty()) which.push_back(k);
          else if (probs[k].rows.empty()) { b.ILActive = false; b.out_charge = probs[k].cols; }
          if (probs[k].rows.empty() || probs[k].cols.empty()) probs[k].rows.clear(), probs[k].cols.clear();
        }
        std::vector<AcaResult> res;
        run_aca(j, 1, probs, which, res);
        trace("level %d outgoing skeletons (%d boxes)", j, (int)which.size());
        std::vector<const std::vector<int>*> rr, cc;
        std::vector<DMat<T>*> outs;
        std::vector<std::vector<int>> pivrows(nb);
        for (int k : which) {
          Box<T>& b = tree[j][k];
          for (int r : res[k].row_bases) { b.out_check.push_back(probs[k].rows.at(r)); pivrows[k].push_back(probs[k].rows.at(r)); }
          for (int c : res[k].col_bases) b.out_charge.push_back(probs[k].cols.at(c));
          rr.push_back(&pivrows[k]);
          cc.push_back(&probs[k].cols);
          outs.push_back(&b.Ar);  // Ar = K(pivot rows, cols)
        }
        entries(rr, cc, outs);
      }
      get_L2P_P2M_level(j);
    }
  }
  void get_L2P_P2M_level(int j) {  // T:681-732, 755-760
    const int nb = nBoxes[j];
    std::vector<const std::vector<int>*> rr, cc;
    std::vector<DMat<T>*> outs;
    std::vector<DMat<T>> D1(nb), D2(nb);
    for (int k = 0; k < nb; ++k) {
      Box<T>& b = tree[j][k];
      size_t nin = b.in_check.size(), nout = b.out_charge.size();
      if (nin < nout) {
        b.out_check.resize(std::min(b.out_check.size(), nin));
        b.out_charge.resize(nin);
      } else {
        b.in_check.resize(nout);
        if (b.in_charge.size() > nout) b.in_charge.resize(nout);
      }
      if (b.ILActive) {
        if (!b.in_check.empty()) { rr.push_back(&b.in_check); cc.push_back(&b.in_charge); outs.push_back(&D1[k]); }
        if (!b.out_check.empty()) { rr.push_back(&b.out_check); cc.push_back(&b.out_charge); outs.push_back(&D2[k]); }
      }
    }
    entries(rr, cc, outs);
    std::vector<DMat<T>> acT, arT;
    std::vector<int> acK, arK;
    for (int k = 0; k < nb; ++k) {
      Box<T>& b = tree[j][k];
      const int r = (int)b.in_check.size();
      if (b.ILActive) {
        DMat<T> temp1 = b.Ac.cols(0, std::min(r, b.Ac.c));
        if (r > 0) {
          // L2P = temp1 * D1^{-1} through the LU of D1^H (T:694-696)
          DMat<T> D1h = newmat(r, r);
          adjoint2d(D1h, D1[k], cx.st);
          DMat<T> lu;
          int* ip = nullptr;
          factor(D1h, lu, ip);
          DMat<T> X = newmat(temp1.c, temp1.r);
          adjoint2d(X, temp1, cx.st);
          solve_inplace(lu, ip, X);
          DMat<T> L2P = newmat(X.c, X.r);
          adjoint2d(L2P, X, cx.st);
          put(b.L2P, L2P);
          drop(X); drop(lu); drop(D1h);
          pfree(ip);
        } else {
          put(b.L2P, zeros(b.Ac.r, r));
        }
        const int ro = (int)b.out_check.size();
        DMat<T> temp2 = b.Ar.rows(0, std::min(ro, b.Ar.r));
        if (ro > 0) {
          DMat<T> lu;
          int* ip = nullptr;
          factor(D2[k], lu, ip);
          DMat<T> X = newmat(temp2.r, temp2.c);
          copy2d(X, temp2, cx.st);
          solve_inplace(lu, ip, X);
          put(b.P2M, X);
          drop(lu);
          pfree(ip);
        } else {
          put(b.P2M, zeros(ro, b.Ar.c));
        }
      } else {
        DMat<T> I1 = newmat((int)b.in_check.size(), (int)b.in_check.size());
        fill_identity(I1, T(1, 0), cx.st);
        put(b.L2P, I1);
        DMat<T> I2 = newmat((int)b.out_charge.size(), (int)b.out_charge.size());
        fill_identity(I2, T(1, 0), cx.st);
        put(b.P2M, I2);
      }
      if (b.Ac.size() > 0) { acT.push_back(b.Ac.cols(0, std::min(r, b.Ac.c))); acK.push_back(k); }
      if (b.Ar.size() > 0) {
        const int ro = (int)b.out_check.size();
        DMat<T> t2 = b.Ar.rows(0, std::min(ro, b.Ar.r));
        DMat<T> adj = newmat(t2.c, t2.r);
        adjoint2d(adj, t2, cx.st);
        arT.push_back(adj);
        arK.push_back(k);
      }
      drop(D1[k]);
      drop(D2[k]);
    }
    std::vector<double> mp;
    max_col_norms(acT, mp, cx);
    for (size_t i = 0; i < acK.size(); ++i) tree[j][acK[i]].maxPivot_L2P = mp[i];
    max_col_norms(arT, mp, cx);
    for (size_t i = 0; i < arK.size(); ++i) tree[j][arK[i]].maxPivot_P2M = mp[i];
    for (auto& a : arT) drop(a);
  }
  void assemble_M2L() {  // T:1258-1273
    std::vector<const std::vector<int>*> rr, cc;
    std::vector<DMat<T>*> outs;
    for (int j = startLevel; j <= nLevels; ++j)
      for (int k = 0; k < nBoxes[j]; ++k)
        for (int kIL : tree[j][k].IL) {
          rr.push_back(&tree[j][k].in_check);
          cc.push_back(&tree[j][kIL].out_charge);
          outs.push_back(&tree[j][k].M2L[kIL]);
        }
    entries(rr, cc, outs);
  }
  void free_points() {  // T:1275-1284
    for (int j = startLevel; j <= nLevels; ++j)
      for (int k = 0; k < nBoxes[j]; ++k) {
        Box<T>& b = tree[j][k];
        b.out_charge.clear();
        b.in_check.clear();
        drop(b.Ar);
        drop(b.Ac);
      }
  }
  void initialise_phase() {  // T:3354-3370
    for (int j = startLevel; j <= nLevels; ++j)
      for (int k = 0; k < nBoxes[j]; ++k) {
        Box<T>& b = tree[j][k];
        b.NumMultipoles = b.P2M.r;
        b.NumLocals = b.L2P.c;
      }
  }
  void initialise_P2P_Leaf_Level() {  // T:3372-3386
    std::vector<const std::vector<int>*> rr, cc;
    std::vector<DMat<T>*> outs;
    const int j = nLevels;
    for (int k = 0; k < nBoxes[j]; ++k) {
      for (int nn : tree[j][k].NB) {
        if (nn == k) continue;
        rr.push_back(&tree[j][k].charge);
        cc.push_back(&tree[j][nn].charge);
        outs.push_back(&tree[j][k].P2P[nn]);
      }
      rr.push_back(&tree[j][k].charge);
      cc.push_back(&tree[j][k].charge);
      outs.push_back(&tree[j][k].P2P[k]);
    }
    entries(rr, cc, outs);
  }

  void assemble() override {
    double t0 = wall();
    trace_t0 = t0;
    cx.pool.reset_peak();
    setup();
    trace("setup: N=%d nLevels=%d", N, nLevels);
    getNodes();
    trace("skeletons done, startLevel=%d", startLevel);
    assemble_M2L();
    if (verbose) cx.sync();
    trace("M2L blocks done");
    free_points();
    initialise_phase();
    if (verbose) cx.sync();
    trace("L2P/P2M done");
    initialise_P2P_Leaf_Level();
    cx.sync();
    trace("assembly done");
    stats_.t_assemble = wall() - t0;
  }

  // ------------------------------------------------------------ elimination
  bool well_separated(int j, int k, int l) {  // T:4044-4055
    if (k == l) return false;
    for (int nn : tree[j][k].NB)
      if (nn == l) return false;
    return true;
  }
  void initialise_P2P_NonLeafLevel(int j) {  // T:3413-3476
    std::vector<CopyOp<T>> cops;
    auto blk = [&](DMat<T>& dst, int r0, int c0, const DMat<T>& src) {
      if (src.size() == 0) return;
      if (r0 + src.r > dst.r || c0 + src.c > dst.c) throw std::runtime_error("P2P non-leaf block out of range");
      CopyOp<T> o{};
      o.src = src.p; o.lds = src.ld; o.dst = dst.p + r0 + (size_t)c0 * dst.ld; o.ldd = dst.ld;
      o.r = src.r; o.c = src.c; o.adj = 0;
      cops.push_back(o);
    };
    for (int k = 0; k < nBoxes[j]; ++k) {
      for (int nn : tree[j][k].NB) {
        if (nn == k) continue;
        int n_rows = 0, n_cols = 0;
        for (int a = 0; a < 8; ++a) {
          int t = 0;
          for (int b = 0; b < 8; ++b) t += tree[j + 1][8 * k + a].M2L[8 * nn + b].c;
          n_cols = std::max(n_cols, t);
        }
        for (int b = 0; b < 8; ++b) {
          int t = 0;
          for (int a = 0; a < 8; ++a) t += tree[j + 1][8 * k + a].M2L[8 * nn + b].r;
          n_rows = std::max(n_rows, t);
        }
        DMat<T> Pm = zeros(n_rows, n_cols);
        int row = 0, col = 0;
        for (int a = 0; a < 8; ++a) {
          for (int b = 0; b < 8; ++b) {
            DMat<T>& M = tree[j + 1][8 * k + a].M2L[8 * nn + b];
            if (M.size() > 0) {
              blk(Pm, row, col, M);
              col += M.c;
            }
          }
          row += tree[j + 1][8 * k + a].M2L[8 * nn + 0].r;
          col = 0;
        }
        put(tree[j][k].P2P[nn], Pm);
      }
      int n_rows = 0, n_cols = 0;
      for (int a = 0; a < 8; ++a) n_rows += tree[j + 1][8 * k + a].NumLocals;
      for (int b = 0; b < 8; ++b) n_cols += tree[j + 1][8 * k + b].NumMultipoles;
      DMat<T> Pm = zeros(n_rows, n_cols);
      int row = 0, col = 0;
      for (int a = 0; a < 8; ++a) {
        Box<T>& ca = tree[j + 1][8 * k + a];
        for (int b = 0; b < 8; ++b) {
          Box<T>& cb = tree[j + 1][8 * k + b];
          DMat<T>& M = ca.M2L[8 * k + b];
          if (M.r == ca.NumLocals && M.c == cb.NumMultipoles) blk(Pm, row, col, M);
          else if (!(M.size() == 0 && (size_t)ca.NumLocals * cb.NumMultipoles == 0))
            throw std::runtime_error("P2P self block size mismatch");
          col += cb.NumMultipoles;
        }
        row += ca.NumLocals;
        col = 0;
      }
      put(tree[j][k].P2P[k], Pm);
    }
    copy_batched(cops, cx);
  }
  void deleteM2L(int j) {  // T:3478-3504
    for (int k = 0; k < nBoxes[j]; ++k)
      for (int nn : tree[j][k].NB)
        for (int a = 0; a < 8; ++a)
          for (int b = 0; b < 8; ++b) drop(tree[j + 1][8 * k + a].M2L[8 * nn + b]);
  }

  // Concatenate blocks column-wise into a new matrix with `rows` rows.
  DMat<T> hcat(int rows, const std::vector<DMat<T>>& parts, std::vector<int>& offs) {
    int total = 0;
    offs.clear();
    for (auto& p : parts) {
      if (p.r != rows && p.size() > 0) throw std::runtime_error("hcat: row mismatch");
      offs.push_back(total);
      total += p.c;
    }
    DMat<T> X = newmat(rows, total);
    std::vector<CopyOp<T>> cops;
    for (size_t i = 0; i < parts.size(); ++i) {
      if (parts[i].size() == 0) continue;
      CopyOp<T> o{};
      o.src = parts[i].p; o.lds = parts[i].ld; o.dst = X.p + (size_t)offs[i] * X.ld; o.ldd = X.ld;
      o.r = rows; o.c = parts[i].c; o.adj = 0;
      cops.push_back(o);
    }
    copy_batched(cops, cx);
    return X;
  }

  void eliminate_x(int j, int k) {  // T:3561-3611, 3653-3791, 3813-4042
    Box<T>& bk = tree[j][k];
    DMat<T>& Pkk = bk.P2P[k];
    if (Pkk.size() != 0) {
      factor(Pkk, bk.luP2P, bk.ipP2P);
      std::vector<DMat<T>> parts{bk.L2P};
      std::vector<int> nbs;
      for (int nn : bk.NB) {
        if (nn == k) continue;
        nbs.push_back(nn);
        parts.push_back(tree[j][nn].Eliminated ? bk.M2P[nn] : bk.P2P[nn]);
      }
      std::vector<int> offs;
      DMat<T> X = hcat(Pkk.r, parts, offs);
      solve_inplace(bk.luP2P, bk.ipP2P, X);
      DMat<T> XL = X.cols(offs[0], parts[0].c);
      std::map<int, DMat<T>> Xn;
      for (size_t i = 0; i < nbs.size(); ++i) Xn[nbs[i]] = X.cols(offs[i + 1], parts[i + 1].c);
      double t0 = wall();
      Group g;
      put(bk.L2M_f, newmat(bk.P2M.r, XL.c));
      add(g, bk.L2M_f, bk.P2M, OP_N, XL, OP_N, T(-1, 0), T(0, 0));
      for (int nn : nbs) {
        DMat<T>& tgt = tree[j][nn].Eliminated ? bk.M2M_f[nn] : bk.P2M_f[nn];
        put(tgt, newmat(bk.P2M.r, Xn[nn].c));
        add(g, tgt, bk.P2M, OP_N, Xn[nn], OP_N, T(-1, 0), T(0, 0));
      }
      for (int nn : nbs) {
        Box<T>& bn = tree[j][nn];
        DMat<T> Ank = bn.Eliminated ? bn.P2L[k] : bn.P2P[k];
        put(bn.L2P_f[k], newmat(Ank.r, XL.c));
        add(g, bn.L2P_f[k], Ank, OP_N, XL, OP_N, T(-1, 0), T(0, 0));
        for (int pp : nbs) {
          const bool pe = tree[j][pp].Eliminated;
          DMat<T>& C = !bn.Eliminated ? (!pe ? bn.P2P[pp] : bn.M2P[pp]) : (!pe ? bn.P2L[pp] : bn.M2L[pp]);
          assign_or_sub(g, C, Ank, OP_N, Xn[pp], OP_N);
        }
      }
      run(g, PROF_SCHUR);
      stats_.t_schur += wall() - t0;
      drop(X);
    } else {  // T:3656-3791: zero blocks of the right shapes
      put(bk.L2M_f, zeros(bk.P2M.r, bk.L2P.c));
      for (int nn : bk.NB) {
        if (nn == k) continue;
        if (!tree[j][nn].Eliminated) put(bk.P2M_f[nn], zeros(bk.P2M.r, bk.P2P[nn].c));
        else put(bk.M2M_f[nn], zeros(bk.P2M.r, bk.M2P[nn].c));
      }
      for (int nn : bk.NB) {
        if (nn == k) continue;
        Box<T>& bn = tree[j][nn];
        if (!bn.Eliminated) put(bn.L2P_f[k], zeros(bn.P2P[k].r, bk.L2P.c));
        else put(bn.L2P_f[k], zeros(bn.P2L[k].r, bk.L2P.c));
        for (int pp : bk.NB) {
          if (pp == k) continue;
          const bool pe = tree[j][pp].Eliminated;
          if (!bn.Eliminated) {
            if (!pe) { if (bn.P2P[pp].size() == 0) put(bn.P2P[pp], zeros(bn.P2P[k].r, bk.P2P[pp].c)); }
            else { if (bn.M2P[pp].size() == 0) put(bn.M2P[pp], zeros(bn.P2P[k].r, bk.M2P[pp].c)); }
          } else {
            if (!pe) { if (bn.P2L[pp].size() == 0) put(bn.P2L[pp], zeros(bn.P2L[k].r, bk.P2P[pp].c)); }
            else (void)bn.M2L[pp];
          }
        }
      }
    }
  }

  void eliminate_z(int j, int k) {  // T:4081-4113, 4150-4509
    Box<T>& bk = tree[j][k];
    if (bk.L2M_f.size() > 0) {
      factor(bk.L2M_f, bk.luL2M, bk.ipL2M);
      const int NM = bk.NumMultipoles;
      DMat<T> mI = newmat(NM, NM);
      fill_identity(mI, T(-1, 0), cx.st);
      std::vector<DMat<T>> parts{mI};
      std::vector<int> nbs;
      for (int nn : bk.NB) {
        if (nn == k) continue;
        nbs.push_back(nn);
        parts.push_back(tree[j][nn].Eliminated ? bk.M2M_f[nn] : bk.P2M_f[nn]);
      }
      std::vector<int> offs;
      DMat<T> Y = hcat(bk.L2M_f.r, parts, offs);
      drop(mI);
      solve_inplace(bk.luL2M, bk.ipL2M, Y);
      DMat<T> Yself = Y.cols(offs[0], NM);
      put(bk.M2L[k], clone(Yself));
      std::map<int, DMat<T>> Yn;
      for (size_t i = 0; i < nbs.size(); ++i) {
        int nn = nbs[i];
        Yn[nn] = Y.cols(offs[i + 1], parts[i + 1].c);
        DMat<T>& tgt = tree[j][nn].Eliminated ? bk.M2L[nn] : bk.P2L[nn];
        put(tgt, clone(Yn[nn]));
      }
      double t0 = wall();
      Group g;
      for (int nn : nbs) {
        Box<T>& bn = tree[j][nn];
        DMat<T> Lf = bn.L2P_f[k];
        DMat<T>& ty = !bn.Eliminated ? bn.M2P[k] : bn.M2L[k];
        assign_or_sub(g, ty, Lf, OP_N, Yself, OP_N);
        for (int pp : nbs) {
          const bool pe = tree[j][pp].Eliminated;
          DMat<T>& C = !bn.Eliminated ? (!pe ? bn.P2P[pp] : bn.M2P[pp]) : (!pe ? bn.P2L[pp] : bn.M2L[pp]);
          if (C.size() == 0 && (size_t)Lf.r * Yn[pp].c > 0) throw std::runtime_error("eliminate_z: missing fill-in block");
          add(g, C, Lf, OP_N, Yn[pp], OP_N, T(-1, 0), T(1, 0));
        }
      }
      run(g, PROF_SCHUR);
      stats_.t_schur += wall() - t0;
      drop(Y);
      // fill-in pairs to recompress (T:4413-4418, 4445-4453), pushed in loop order
      std::vector<DMat<T>> nm;
      std::vector<std::array<int, 3>> cand;  // (type, nn, pp)
      for (int nn : nbs) {
        Box<T>& bn = tree[j][nn];
        for (int pp : nbs) {
          const bool pe = tree[j][pp].Eliminated;
          if (!bn.Eliminated && !pe && well_separated(j, nn, pp) && nn < pp) {
            cand.push_back({0, nn, pp});
            nm.push_back(bn.P2P[pp]);
            nm.push_back(bn.M2L[pp]);
          } else if (bn.Eliminated && !pe && well_separated(j, nn, pp)) {
            cand.push_back({1, nn, pp});
            nm.push_back(bn.P2L[pp]);
            nm.push_back(bn.M2L[pp]);
          }
        }
      }
      std::vector<double> v = norms(nm);
      for (size_t i = 0; i < cand.size(); ++i) {
        double f = v[2 * i], m = v[2 * i + 1];
        if (m > 0 && f / m > fillInTolerance) {
          if (cand[i][0] == 0) P2Plist.push_back({cand[i][1], cand[i][2]});
          else P2L_M2P.push_back({cand[i][1], cand[i][2]});
        }
      }
    } else {  // T:4150-4296
      put(bk.M2L[k], newmat(0, bk.NumMultipoles));
      for (int nn : bk.NB) {
        if (nn == k) continue;
        if (!tree[j][nn].Eliminated) put(bk.P2L[nn], newmat(0, bk.P2M_f[nn].c));
        else put(bk.M2L[nn], newmat(0, bk.M2M_f[nn].c));
      }
      for (int nn : bk.NB) {
        if (nn == k) continue;
        Box<T>& bn = tree[j][nn];
        if (!bn.Eliminated) {
          if (bn.M2P[k].size() == 0) put(bn.M2P[k], zeros(bn.L2P_f[k].r, bk.NumMultipoles));
        } else {
          if (bn.M2L[k].size() == 0) put(bn.M2L[k], zeros(bn.L2P_f[k].r, bk.NumMultipoles));
        }
      }
    }
  }

  // ------------------------------------------------------------ fill-in recompression
  struct QR {
    DMat<T> Q;
    double maxPivot;
  };
  // Independent RRQRs (same level tolerances and fill-in cap) through shared launches.
  //
  // orth[i]: the first parts[i] columns of As[i] are known to be orthonormal (the Q of an
  // earlier RRQR). The first partition's pivoted QR then only re-derives their span: every
  // pivot is 1 to working precision, so it neither stops early nor truncates. For such a
  // matrix the first partition is taken as it is, and the second partition's RRQR runs on
  // the remaining columns projected twice onto the complement of that span. Its pivots, the
  // stopping test on them and the fill-in cap are those of the second partition of the
  // full RRQR; the norms it pivots on are computed from the projected columns instead of
  // downdated (the same values up to rounding). Q = [leading columns | new columns].
  std::vector<QR> do_rrqr_batch(const std::vector<DMat<T>>& As, const std::vector<int>& parts,
                                const std::vector<double>& pivots, int j, const std::vector<char>& orth) {
    const size_t nb = As.size();
    std::vector<RRQRTask<T>> tk(nb);
    std::vector<char> proj(nb, 0);
    std::vector<DMat<T>> Pp(nb), C(nb);
    double t0 = wall();
    for (size_t i = 0; i < nb; ++i) {
      const int r = parts[i], m = As[i].r, n = As[i].c;
      proj[i] = orth[i] && r > 0 && r < m && r < n && !full_rrqr_;
    }
    for (int pass = 0; pass < 2; ++pass) {  // P <- P - Q1 (Q1^H P), twice
      Group g1, g2;
      for (size_t i = 0; i < nb; ++i) {
        if (!proj[i]) continue;
        const int r = parts[i];
        DMat<T> Q1 = As[i].cols(0, r);
        if (pass == 0) Pp[i] = clone(As[i].cols(r, As[i].c - r));
        if (pass == 0) C[i] = newmat(r, Pp[i].c);
        add(g1, C[i], Q1, OP_C, Pp[i], OP_N, T(1, 0), T(0, 0));
        add(g2, Pp[i], Q1, OP_N, C[i], OP_N, T(-1, 0), T(1, 0));
      }
      run(g1);
      run(g2);
    }
    for (size_t i = 0; i < nb; ++i) {
      drop(C[i]);
      const int m = As[i].r;
      if (proj[i]) {
        tk[i].A = Pp[i];
        tk[i].partition = 0;
        tk[i].fillInRank = std::min(rankVec[j], m - parts[i]);
      } else {
        tk[i].A = As[i];
        tk[i].partition = parts[i];
        tk[i].fillInRank = rankVec[j];
      }
      tk[i].inputPivot = pivots[i];
      tk[i].Q = newmat(m, std::max(1, std::min(m, tk[i].A.c)));
    }
    int h = cx.prof.begin(cx.st);
    rrqr_batched(tk, std::pow(10.0, -TOL_POW), std::pow(10.0, -fill_in_TOL_POW), cx);
    double fl = 0;
    std::vector<QR> out(nb);
    for (size_t i = 0; i < nb; ++i) {
      const RRQRInfo& inf = tk[i].info;
      const int lead = proj[i] ? parts[i] : 0;
      double r = inf.rank, m = tk[i].A.r, n = tk[i].A.c;
      fl += 8.0 * (2.0 * m * n * r - r * r * (m + n) + 2.0 * r * r * r / 3.0) + 8.0 * (2.0 * m * r * r - 2.0 * r * r * r / 3.0);
      out[i].Q = newmat(As[i].r, lead + inf.rank);
      if (lead > 0) copy2d(out[i].Q.cols(0, lead), As[i].cols(0, lead), cx.st);
      if (inf.rank > 0) copy2d(out[i].Q.cols(lead, inf.rank), tk[i].Q.cols(0, inf.rank), cx.st);
      out[i].maxPivot = inf.maxPivot;
      drop(tk[i].Q);
      if (proj[i]) drop(Pp[i]);
    }
    cx.prof.end(cx.st, h, PROF_RRQR, fl);
    stats_.t_compress += wall() - t0;
    stats_.n_rrqr += (long long)nb;
    stats_.qr_flops += fl;
    return out;
  }
  DMat<T> hcat2(DMat<T> A, int opA, DMat<T> B, int opB) {  // [op(A) | op(B)]
    int ra = opA == OP_N ? A.r : A.c, ca = opA == OP_N ? A.c : A.r;
    int rb = opB == OP_N ? B.r : B.c, cb = opB == OP_N ? B.c : B.r;
    if (ra != rb) throw std::runtime_error("hcat2: row mismatch");
    DMat<T> X = newmat(ra, ca + cb);
    std::vector<CopyOp<T>> cops(2);
    cops[0] = CopyOp<T>{A.p, X.p, A.ld, X.ld, ra, ca, opA, 0};
    cops[1] = CopyOp<T>{B.p, X.p + (size_t)ca * X.ld, B.ld, X.ld, rb, cb, opB, 0};
    copy_batched(cops, cx);
    return X;
  }
  void update_other_M2Ls_at_k_old(int j, int k, int b, DMat<T> newL2P) {  // T:3159-3188
    Box<T>& bk = tree[j][k];
    DMat<T> R = mul(newL2P, OP_C, bk.L2P, OP_N);
    std::vector<DMat<T>*> tg;
    for (auto& kv : bk.P2L) tg.push_back(&kv.second);
    for (int kIL : bk.IL)
      if (kIL != b) tg.push_back(&bk.M2L[kIL]);
    for (int kN : bk.NB)
      if (kN != k && tree[j][kN].Eliminated && bk.Eliminated) tg.push_back(&bk.M2L[kN]);
    if (bk.Eliminated) tg.push_back(&bk.M2L[k]);
    Group g;
    std::vector<DMat<T>> fresh;
    for (DMat<T>* X : tg) {
      if (X->r != R.c) throw std::runtime_error("update_other_M2Ls_at_k: shape mismatch");
      DMat<T> Y = newmat(R.r, X->c);
      add(g, Y, R, OP_N, *X, OP_N, T(1, 0), T(0, 0));
      fresh.push_back(Y);
    }
    run(g, PROF_UPDATE);
    for (size_t i = 0; i < tg.size(); ++i) put(*tg[i], fresh[i]);
    drop(R);
  }
  void update_other_M2Ls_at_b_old(int j, int k, int b, DMat<T> newP2Mt) {  // T:3252-3288
    Box<T>& bb = tree[j][b];
    DMat<T> R = mul(bb.P2M, OP_N, newP2Mt, OP_N);
    std::vector<DMat<T>*> tg;
    for (auto& kv : bb.P2L) tg.push_back(&tree[j][kv.first].M2P[b]);
    for (int kIL : bb.IL)
      if (kIL != k) tg.push_back(&tree[j][kIL].M2L[b]);
    for (int kN : bb.NB)
      if (kN != b && tree[j][kN].Eliminated && bb.Eliminated) tg.push_back(&tree[j][kN].M2L[b]);
    if (bb.Eliminated) tg.push_back(&bb.M2L[b]);
    Group g;
    std::vector<DMat<T>> fresh;
    for (DMat<T>* X : tg) {
      if (X->c != R.r) throw std::runtime_error("update_other_M2Ls_at_b: shape mismatch");
      DMat<T> Y = newmat(X->r, R.c);
      add(g, Y, *X, OP_N, R, OP_N, T(1, 0), T(0, 0));
      fresh.push_back(Y);
    }
    run(g, PROF_UPDATE);
    for (size_t i = 0; i < tg.size(); ++i) put(*tg[i], fresh[i]);
    drop(R);
  }
  void update_parent_L2P_old(int j, int k, DMat<T> newL2P) {  // T:3290-3314
    int kP = k / 8, ck = k % 8;
    Box<T>& par = tree[j - 1][kP];
    std::vector<int> r0(8), nr(8);
    int row = 0;
    for (int c = 0; c < 8; ++c) { r0[c] = row; nr[c] = tree[j][kP * 8 + c].NumLocals; row += nr[c]; }
    DMat<T> Rm = mul(newL2P, OP_C, tree[j][k].L2P, OP_N);
    DMat<T> Uk = mul(Rm, OP_N, par.L2P.rows(r0[ck], nr[ck]), OP_N);
    int total = row - nr[ck] + Uk.r;
    DMat<T> Pn = newmat(total, par.L2P.c);
    std::vector<CopyOp<T>> cops;
    int out = 0;
    for (int c = 0; c < 8; ++c) {
      DMat<T> src = (c == ck) ? Uk : par.L2P.rows(r0[c], nr[c]);
      if (src.size()) cops.push_back(CopyOp<T>{src.p, Pn.p + out, src.ld, Pn.ld, src.r, src.c, 0, 0});
      out += src.r;
    }
    copy_batched(cops, cx);
    put(par.L2P, Pn);
    par.orthL2P = false;
    drop(Rm);
    drop(Uk);
  }
  void update_parent_P2M_old(int j, int b, DMat<T> newP2Mt) {  // T:3316-3340
    int bP = b / 8, cb = b % 8;
    Box<T>& par = tree[j - 1][bP];
    std::vector<int> c0(8), nc(8);
    int col = 0;
    for (int c = 0; c < 8; ++c) { c0[c] = col; nc[c] = tree[j][bP * 8 + c].NumMultipoles; col += nc[c]; }
    DMat<T> Rm = mul(tree[j][b].P2M, OP_N, newP2Mt, OP_N);
    DMat<T> Vb = mul(par.P2M.cols(c0[cb], nc[cb]), OP_N, Rm, OP_N);
    int total = col - nc[cb] + Vb.c;
    DMat<T> Pn = newmat(par.P2M.r, total);
    std::vector<CopyOp<T>> cops;
    int out = 0;
    for (int c = 0; c < 8; ++c) {
      DMat<T> src = (c == cb) ? Vb : par.P2M.cols(c0[c], nc[c]);
      if (src.size()) cops.push_back(CopyOp<T>{src.p, Pn.p + (size_t)out * Pn.ld, src.ld, Pn.ld, src.r, src.c, 0, 0});
      out += src.c;
    }
    copy_batched(cops, cx);
    put(par.P2M, Pn);
    par.orthP2M = false;
    drop(Rm);
    drop(Vb);
  }
  // compress_P2P_myRRQR (T:1573-1656) for both directions of compress_P2P_Two_Ways:
  // d = 0 is (k, b), d = 1 is (b, k). The two directions read disjoint state (bases,
  // fill-ins, pivots and M2L blocks of opposite roles), so their RRQRs share launches.
  // Outputs per direction: newL2P of the row box, newP2M^H of the column box, new M2L.
  void compress_P2P_both(int j, int k, int b, DMat<T> newL2P[2], DMat<T> newP2Mt[2], DMat<T> M2L[2]) {
    const int ra[2] = {k, b}, ca[2] = {b, k};
    std::vector<DMat<T>> A(2);
    std::vector<int> part(2);
    std::vector<double> piv(2);
    for (int d = 0; d < 2; ++d) {
      Box<T>& br = tree[j][ra[d]];
      A[d] = hcat2(br.L2P, OP_N, br.P2P[ca[d]], OP_N);
      part[d] = br.L2P.c;
      piv[d] = br.maxPivot_L2P;
    }
    std::vector<char> orth1{tree[j][ra[0]].orthL2P, tree[j][ra[1]].orthL2P};
    std::vector<QR> q1 = do_rrqr_batch(A, part, piv, j, orth1);
    DMat<T> Rk[2], Rkp[2];
    for (int d = 0; d < 2; ++d) {
      Box<T>& br = tree[j][ra[d]];
      drop(A[d]);
      br.maxPivot_L2P = std::max(br.maxPivot_L2P, q1[d].maxPivot);
      newL2P[d] = q1[d].Q;
      Rk[d] = newmat(newL2P[d].c, br.L2P.c);
      Rkp[d] = newmat(newL2P[d].c, br.P2P[ca[d]].c);
    }
    {
      Group g;
      for (int d = 0; d < 2; ++d) {
        Box<T>& br = tree[j][ra[d]];
        add(g, Rk[d], newL2P[d], OP_C, br.L2P, OP_N, T(1, 0), T(0, 0));
        add(g, Rkp[d], newL2P[d], OP_C, br.P2P[ca[d]], OP_N, T(1, 0), T(0, 0));
      }
      run(g);
    }
    for (int d = 0; d < 2; ++d) {
      Box<T>& bc = tree[j][ca[d]];
      A[d] = hcat2(bc.P2M, OP_C, Rkp[d], OP_C);
      part[d] = bc.P2M.r;
      piv[d] = bc.maxPivot_P2M;
    }
    std::vector<char> orth2{tree[j][ca[0]].orthP2M, tree[j][ca[1]].orthP2M};
    std::vector<QR> q2 = do_rrqr_batch(A, part, piv, j, orth2);
    DMat<T> Rn[2], Rnp[2], t[2];
    for (int d = 0; d < 2; ++d) {
      Box<T>& bc = tree[j][ca[d]];
      drop(A[d]);
      bc.maxPivot_P2M = std::max(bc.maxPivot_P2M, q2[d].maxPivot);
      newP2Mt[d] = q2[d].Q;
      Rn[d] = newmat(newP2Mt[d].c, bc.P2M.r);
      Rnp[d] = newmat(newP2Mt[d].c, Rkp[d].r);
    }
    {
      Group g;
      for (int d = 0; d < 2; ++d) {
        Box<T>& bc = tree[j][ca[d]];
        add(g, Rn[d], newP2Mt[d], OP_C, bc.P2M, OP_C, T(1, 0), T(0, 0));
        add(g, Rnp[d], newP2Mt[d], OP_C, Rkp[d], OP_C, T(1, 0), T(0, 0));
        DMat<T>& M = tree[j][ra[d]].M2L[ca[d]];
        if (M.size() > 0) {
          t[d] = newmat(Rk[d].r, M.c);
          add(g, t[d], Rk[d], OP_N, M, OP_N, T(1, 0), T(0, 0));
        }
      }
      run(g);
    }
    {
      Group g;
      for (int d = 0; d < 2; ++d) {  // M2L = Rnp^H + (Rk M) Rn^H
        M2L[d] = newmat(Rnp[d].c, Rnp[d].r);
        adjoint2d(M2L[d], Rnp[d], cx.st);
        if (t[d].size() > 0) add(g, M2L[d], t[d], OP_N, Rn[d], OP_C, T(1, 0), T(1, 0));
      }
      run(g);
    }
    for (int d = 0; d < 2; ++d) { drop(Rk[d]); drop(Rkp[d]); drop(Rn[d]); drop(Rnp[d]); drop(t[d]); }
  }
  void square_up(DMat<T> L2P, DMat<T> P2Mt, DMat<T>& L2Pm, DMat<T>& P2Mtm) {  // T:2168-2183
    if (L2P.c < P2Mt.c) { P2Mtm = P2Mt.cols(0, L2P.c); L2Pm = L2P; }
    else { L2Pm = L2P.cols(0, P2Mt.c); P2Mtm = P2Mt; }
  }
  void compress_P2P_Two_Ways(int j, int k, int b) {  // T:2116-2306
    if (tree[j][b].M2L[k].size() == 0 || tree[j][k].M2L[b].size() == 0) return;
    stats_.n_compress_p2p++;
    DMat<T> nL2P[2], nP2Mt[2], nM2L[2];
    compress_P2P_both(j, k, b, nL2P, nP2Mt, nM2L);
    DMat<T> nL2Pk = nL2P[0], nP2Mtb = nP2Mt[0], M2Lkb = nM2L[0];
    DMat<T> nL2Pb = nL2P[1], nP2Mtk = nP2Mt[1], M2Lbk = nM2L[1];
    DMat<T> L2Pk_m, P2Mtk_m, L2Pb_m, P2Mtb_m;
    square_up(nL2Pk, nP2Mtk, L2Pk_m, P2Mtk_m);
    square_up(nL2Pb, nP2Mtb, L2Pb_m, P2Mtb_m);
    put(tree[j][k].M2L[b], clone(M2Lkb.rows(0, L2Pk_m.c).cols(0, P2Mtb_m.c)));
    update_other_M2Ls_at_k_old(j, k, b, L2Pk_m);
    update_other_M2Ls_at_b_old(j, k, b, P2Mtb_m);
    if (j > startLevel) {
      update_parent_L2P_old(j, k, L2Pk_m);
      update_parent_P2M_old(j, b, P2Mtb_m);
    }
    put(tree[j][k].L2P, clone(L2Pk_m));
    tree[j][k].orthL2P = true;
    {
      DMat<T> t = newmat(P2Mtb_m.c, P2Mtb_m.r);
      adjoint2d(t, P2Mtb_m, cx.st);
      put(tree[j][b].P2M, t);
      tree[j][b].orthP2M = true;
    }
    tree[j][k].NumLocals = tree[j][k].L2P.c;
    tree[j][b].NumMultipoles = tree[j][b].P2M.r;
    put(tree[j][b].M2L[k], clone(M2Lbk.rows(0, L2Pb_m.c).cols(0, P2Mtk_m.c)));
    update_other_M2Ls_at_k_old(j, b, k, L2Pb_m);
    update_other_M2Ls_at_b_old(j, b, k, P2Mtk_m);
    if (j > startLevel) {
      update_parent_L2P_old(j, b, L2Pb_m);
      update_parent_P2M_old(j, k, P2Mtk_m);
    }
    put(tree[j][b].L2P, clone(L2Pb_m));
    tree[j][b].orthL2P = true;
    {
      DMat<T> t = newmat(P2Mtk_m.c, P2Mtk_m.r);
      adjoint2d(t, P2Mtk_m, cx.st);
      put(tree[j][k].P2M, t);
      tree[j][k].orthP2M = true;
    }
    tree[j][b].NumLocals = tree[j][b].L2P.c;
    tree[j][k].NumMultipoles = tree[j][k].P2M.r;
    drop(nL2Pk); drop(nP2Mtb); drop(M2Lkb); drop(nL2Pb); drop(nP2Mtk); drop(M2Lbk);
  }
  void compress_M2P_P2L(int j, int k, int b) {  // T:2454-2596, 2878-2959 (k not eliminated, b eliminated)
    if (tree[j][b].M2L[k].size() == 0 || tree[j][k].M2L[b].size() == 0) return;
    stats_.n_compress_p2l++;
    Box<T>& bk = tree[j][k];
    Box<T>& bb = tree[j][b];
    // the column and row RRQRs read disjoint inputs: one batched launch sequence
    DMat<T> Acol = hcat2(bk.L2P, OP_N, bk.M2P[b], OP_N);
    DMat<T> Arow = hcat2(bk.P2M, OP_C, bb.P2L[k], OP_C);
    std::vector<QR> qq = do_rrqr_batch({Acol, Arow}, {bk.L2P.c, bk.P2M.r}, {bk.maxPivot_L2P, bk.maxPivot_P2M}, j,
                                       {(char)bk.orthL2P, (char)bk.orthP2M});
    drop(Acol);
    drop(Arow);
    const QR& q1 = qq[0];
    const QR& q2 = qq[1];
    DMat<T> nL2P = q1.Q;
    bk.maxPivot_L2P = std::max(bk.maxPivot_L2P, q1.maxPivot);
    DMat<T> Rk = mul(nL2P, OP_C, bk.L2P, OP_N);
    DMat<T> M2Lkb = mul(nL2P, OP_C, bk.M2P[b], OP_N);  // Rk' then += Rk*M2L
    {
      Group g;
      add(g, M2Lkb, Rk, OP_N, bk.M2L[b], OP_N, T(1, 0), T(1, 0));
      run(g);
    }
    DMat<T> nP2Mt = q2.Q;
    bk.maxPivot_P2M = std::max(bk.maxPivot_P2M, q2.maxPivot);
    DMat<T> Rn = mul(nP2Mt, OP_C, bk.P2M, OP_C);
    DMat<T> Rnp = mul(nP2Mt, OP_C, bb.P2L[k], OP_C);
    DMat<T> M2Lbk = newmat(Rnp.c, Rnp.r);
    adjoint2d(M2Lbk, Rnp, cx.st);
    {
      Group g;
      add(g, M2Lbk, bb.M2L[k], OP_N, Rn, OP_C, T(1, 0), T(1, 0));
      run(g);
    }
    DMat<T> L2Pm, P2Mtm;
    square_up(nL2P, nP2Mt, L2Pm, P2Mtm);
    put(bk.M2L[b], clone(M2Lkb.rows(0, L2Pm.c)));
    update_other_M2Ls_at_k_old(j, k, b, L2Pm);
    if (j > startLevel) update_parent_L2P_old(j, k, L2Pm);
    put(bk.L2P, clone(L2Pm));
    bk.orthL2P = true;
    bk.NumLocals = bk.L2P.c;
    put(bb.M2L[k], clone(M2Lbk.cols(0, P2Mtm.c)));
    update_other_M2Ls_at_b_old(j, b, k, P2Mtm);
    if (j > startLevel) update_parent_P2M_old(j, k, P2Mtm);
    {
      DMat<T> t = newmat(P2Mtm.c, P2Mtm.r);
      adjoint2d(t, P2Mtm, cx.st);
      put(bk.P2M, t);
      bk.orthP2M = true;
    }
    bk.NumMultipoles = bk.P2M.r;
    drop(nL2P); drop(Rk); drop(M2Lkb); drop(nP2Mt); drop(Rn); drop(Rnp); drop(M2Lbk);
  }

  void factorize() override {  // T:4599-4747
    double t0 = wall();
    for (int j = nLevels; j >= startLevel; --j) {
      if (j != nLevels) {
        initialise_P2P_NonLeafLevel(j);
        deleteM2L(j);
      }
      for (int k = 0; k < nBoxes[j]; ++k) {
        double tc = wall();
        for (size_t i = 0; i < P2Plist.size(); ++i) {
          int nn = P2Plist[i].first, pp = P2Plist[i].second;
          if (k == nn || k == pp) {
            P2Plist.erase(P2Plist.begin() + i);
            --i;
            if (tree[j][nn].P2P[pp].size() != 0) {
              compress_P2P_Two_Ways(j, nn, pp);
              drop(tree[j][nn].P2P[pp]);
              drop(tree[j][pp].P2P[nn]);
            }
          }
        }
        for (size_t i = 0; i < P2L_M2P.size(); ++i) {
          int nn = P2L_M2P[i].first, pp = P2L_M2P[i].second;
          if (k == nn || k == pp) {
            P2L_M2P.erase(P2L_M2P.begin() + i);
            --i;
            if (tree[j][nn].P2L[pp].size() != 0) {
              compress_M2P_P2L(j, pp, nn);
              drop(tree[j][nn].P2L[pp]);
              drop(tree[j][pp].M2P[nn]);
            }
          }
        }
        (void)tc;
        tree[j][k].Eliminated = true;
        if (k % 16 == 0) trace("eliminate level %d box %d", j, k);
        eliminate_x(j, k);
        eliminate_z(j, k);
      }
    }
    trace("base level: %d boxes", nBoxes[startLevel]);
    factor_base_level();
    cx.sync();
    trace("factorization done");
    stats_.t_factor = wall() - t0;
    stats_.peak_bytes = cx.pool.peak();
    factorized_ = true;
  }

  // ------------------------------------------------------------ operator apply
  // y = Z x with the assembled (not yet factorized) compressed operator: the leaf near
  // field P2P plus the nested far field L2P M2L P2M (upward pass through P2M, M2L at every
  // level, downward pass through L2P), every level one grouped launch. x and y are device
  // complex128, N x nrhs (ld N), in input order.
  void apply_device(const void* x_dev, int nrhs, void* y_dev) override {
    if (factorized_) throw std::runtime_error("operator apply after factorize (the blocks were eliminated)");
    const int jl = nLevels;
    // box order: the concatenated leaf charge lists
    std::vector<int> order, poff(nBoxes[jl] + 1, 0);
    for (int k = 0; k < nBoxes[jl]; ++k) {
      const auto& ch = tree[jl][k].charge;
      order.insert(order.end(), ch.begin(), ch.end());
      poff[k + 1] = poff[k] + (int)ch.size();
    }
    DMat<T> X = newmat(N, nrhs), Xb = newmat(N, nrhs), Yb = newmat(N, nrhs);
    to_T(X, DMat<cpx<double>>{(cpx<double>*)x_dev, N, nrhs, N});
    int* d_order = (int*)palloc(sizeof(int) * std::max(N, 1));  // lives across many launches
    cx.upload(d_order, order.data(), sizeof(int) * order.size());
    gather_rows(Xb, X, d_order, cx.st);
    // per level contiguous multipole / local buffers, box after box
    std::vector<std::vector<int>> moff(nLevels + 1), loff(nLevels + 1);
    std::vector<DMat<T>> mult(nLevels + 1), loc(nLevels + 1);
    for (int j = startLevel; j <= nLevels; ++j) {
      moff[j].assign(nBoxes[j] + 1, 0);
      loff[j].assign(nBoxes[j] + 1, 0);
      for (int k = 0; k < nBoxes[j]; ++k) {
        moff[j][k + 1] = moff[j][k] + tree[j][k].P2M.r;
        loff[j][k + 1] = loff[j][k] + tree[j][k].L2P.c;
      }
      mult[j] = newmat(moff[j][nBoxes[j]], nrhs);
      loc[j] = zeros(loff[j][nBoxes[j]], nrhs);
    }
    auto slice = [&](DMat<T>& m, int r0, int nr) { return m.rows(r0, nr); };
    // upward: leaf P2M on the particles, then P2M on the children's multipoles
    for (int j = nLevels; j >= startLevel; --j) {
      Group g;
      for (int k = 0; k < nBoxes[j]; ++k) {
        const Box<T>& b = tree[j][k];
        if (b.P2M.size() == 0) continue;
        DMat<T> src = (j == nLevels) ? slice(Xb, poff[k], poff[k + 1] - poff[k])
                                     : slice(mult[j + 1], moff[j + 1][8 * k], moff[j + 1][8 * k + 8] - moff[j + 1][8 * k]);
        if (src.r != b.P2M.c) throw std::runtime_error("apply: P2M does not match its sources");
        add(g, slice(mult[j], moff[j][k], b.P2M.r), b.P2M, OP_N, src, OP_N, T(1, 0), T(0, 0));
      }
      run(g);
    }
    // M2L at every level: locals of k from the multipoles of its interaction list
    for (int j = startLevel; j <= nLevels; ++j) {
      Group g;
      for (int k = 0; k < nBoxes[j]; ++k) {
        Box<T>& b = tree[j][k];
        const int nl = b.L2P.c;
        if (nl == 0) continue;
        for (int l : b.IL) {
          auto it = b.M2L.find(l);
          if (it == b.M2L.end() || it->second.size() == 0) continue;
          add(g, slice(loc[j], loff[j][k], nl), it->second, OP_N, slice(mult[j], moff[j][l], tree[j][l].P2M.r), OP_N,
              T(1, 0), T(1, 0), 1);
        }
      }
      run(g);
    }
    // downward: L2P of a parent adds to its children's locals
    for (int j = startLevel; j < nLevels; ++j) {
      Group g;
      for (int k = 0; k < nBoxes[j]; ++k) {
        const Box<T>& b = tree[j][k];
        if (b.L2P.size() == 0) continue;
        DMat<T> dst = slice(loc[j + 1], loff[j + 1][8 * k], loff[j + 1][8 * k + 8] - loff[j + 1][8 * k]);
        if (dst.r != b.L2P.r) throw std::runtime_error("apply: L2P does not match the children's locals");
        add(g, dst, b.L2P, OP_N, slice(loc[j], loff[j][k], b.L2P.c), OP_N, T(1, 0), T(1, 0));
      }
      run(g);
    }
    // leaf: particles from locals plus the near field
    {
      Group g;
      zero2d(Yb, cx.st);
      for (int k = 0; k < nBoxes[jl]; ++k) {
        Box<T>& b = tree[jl][k];
        const int n = poff[k + 1] - poff[k];
        if (n == 0) continue;
        DMat<T> yk = slice(Yb, poff[k], n);
        if (b.L2P.size() > 0)
          add(g, yk, b.L2P, OP_N, slice(loc[jl], loff[jl][k], b.L2P.c), OP_N, T(1, 0), T(1, 0), 1);
        for (int nn : b.NB) {
          auto it = b.P2P.find(nn);
          if (it == b.P2P.end() || it->second.size() == 0) continue;
          add(g, yk, it->second, OP_N, slice(Xb, poff[nn], poff[nn + 1] - poff[nn]), OP_N, T(1, 0), T(1, 0), 1);
        }
      }
      run(g);
    }
    DMat<T> Y = newmat(N, nrhs);
    scatter_rows(Y, Yb, d_order, cx.st);
    from_T(DMat<cpx<double>>{(cpx<double>*)y_dev, N, nrhs, N}, Y);
    for (int j = startLevel; j <= nLevels; ++j) { drop(mult[j]); drop(loc[j]); }
    drop(X); drop(Xb); drop(Yb); drop(Y);
    pfree(d_order);
  }
  int size() const override { return N; }
  Ctx& context() override { return cx; }
  // T:4772-4810: the base-level system is factored once here; the reference builds
  // and factors it inside every back substitution.
  void factor_base_level() {
    const int j = startLevel;
    NMb = 0;
    NLb = 0;
    for (int k = 0; k < nBoxes[j]; ++k) { NMb += tree[j][k].NumMultipoles; NLb += tree[j][k].NumLocals; }
    if (NLb == 0 || NMb == 0) return;
    if (NLb != NMb) throw std::runtime_error("base-level system is not square");
    DMat<T> A = zeros(NLb, NMb);
    std::vector<CopyOp<T>> cops;
    int ri = 0;
    for (int r = 0; r < nBoxes[j]; ++r) {
      int si = 0;
      for (int s = 0; s < nBoxes[j]; ++s) {
        auto it = tree[j][r].M2L.find(s);
        if (it != tree[j][r].M2L.end() && it->second.r == tree[j][r].NumLocals &&
            it->second.c == tree[j][s].NumMultipoles && it->second.size() > 0)
          cops.push_back(CopyOp<T>{it->second.p, A.p + ri + (size_t)si * A.ld, it->second.ld, A.ld, it->second.r,
                                   it->second.c, 0, 0});
        si += tree[j][s].NumMultipoles;
      }
      ri += tree[j][r].NumLocals;
    }
    copy_batched(cops, cx);
    factor(A, luBase, ipBase);
    drop(A);
  }

  // ------------------------------------------------------------ solve
  DMat<T> vcat_children_mrhs(int j, int k, int nrhs) {
    int len = 0;
    for (int c = 0; c < 8; ++c) len += tree[j + 1][8 * k + c].multipole_rhs.r;
    DMat<T> v = newmat(len, nrhs);
    std::vector<CopyOp<T>> cops;
    int off = 0;
    for (int c = 0; c < 8; ++c) {
      DMat<T>& m = tree[j + 1][8 * k + c].multipole_rhs;
      if (m.size()) cops.push_back(CopyOp<T>{m.p, v.p + off, m.ld, v.ld, m.r, nrhs, 0, 0});
      off += m.r;
    }
    copy_batched(cops, cx);
    return v;
  }
  void solve(const double* rhs, int nrhs, double* x, double* x_box) override {
    DMat<cpx<double>> B64{(cpx<double>*)palloc(sizeof(cpx<double>) * (size_t)N * nrhs), N, nrhs, N};
    DMat<cpx<double>> X64{(cpx<double>*)palloc(sizeof(cpx<double>) * (size_t)N * nrhs), N, nrhs, N};
    DMat<cpx<double>> XB64{x_box ? (cpx<double>*)palloc(sizeof(cpx<double>) * (size_t)N * nrhs) : nullptr,
                           N, nrhs, N};
    cx.upload(B64.p, rhs, sizeof(cpx<double>) * (size_t)N * nrhs);
    solve_device(B64.p, nrhs, X64.p, XB64.p);
    HIPCHK(hipMemcpyAsync(x, X64.p, sizeof(cpx<double>) * X64.size(), hipMemcpyDeviceToHost, cx.st));
    if (x_box)
      HIPCHK(hipMemcpyAsync(x_box, XB64.p, sizeof(cpx<double>) * XB64.size(), hipMemcpyDeviceToHost, cx.st));
    cx.sync();
    pfree(B64.p);
    pfree(X64.p);
    if (XB64.p) pfree(XB64.p);
  }
  // rhs / x / x_box: device complex128, N x nrhs (ld N); x in input order, x_box (may be
  // null) in box order. Stream ordered; no host synchronisation.
  void solve_device(const void* rhs_dev, int nrhs, void* x_dev, void* xbox_dev) override {
    if (!factorized_) throw std::runtime_error("solve before factorize");
    double t0 = wall();
    // assign_Leaf_rhs (T:670-678)
    DMat<T> B = newmat(N, nrhs);
    to_T(B, DMat<cpx<double>>{(cpx<double>*)rhs_dev, N, nrhs, N});
    for (int k = 0; k < nBoxes[nLevels]; ++k) {
      Box<T>& b = tree[nLevels][k];
      put(b.particle_rhs, newmat((int)b.charge.size(), nrhs));
      if (!b.charge.empty()) gather_rows(b.particle_rhs, B, upload_ints(b.charge), cx.st);
    }
    double ts = wall();
    // rhs_eliminate_phase_efficient (T:4749-4770)
    for (int j = nLevels; j >= startLevel; --j)
      for (int k = 0; k < nBoxes[j]; ++k) tree[j][k].Eliminated = false;
    for (int j = nLevels; j >= startLevel; --j) {
      if (j != nLevels)
        for (int k = 0; k < nBoxes[j]; ++k) put(tree[j][k].particle_rhs, vcat_children_mrhs(j, k, nrhs));
      for (int k = 0; k < nBoxes[j]; ++k) {
        Box<T>& bk = tree[j][k];
        bk.Eliminated = true;
        if (bk.P2P[k].size() != 0) {  // T:3613-3653
          DMat<T> s = solved(bk.luP2P, bk.ipP2P, bk.invP2P, bk.particle_rhs);
          Group g;
          put(bk.local_rhs, newmat(bk.P2M.r, nrhs));
          add(g, bk.local_rhs, bk.P2M, OP_N, s, OP_N, T(-1, 0), T(0, 0));
          for (int nn : bk.NB) {
            if (nn == k) continue;
            Box<T>& bn = tree[j][nn];
            if (!bn.Eliminated) add(g, bn.particle_rhs, bn.P2P[k], OP_N, s, OP_N, T(-1, 0), T(1, 0));
            else add(g, bn.multipole_rhs, bn.P2L[k], OP_N, s, OP_N, T(-1, 0), T(1, 0));
          }
          run(g);
          drop(s);
        } else {
          put(bk.local_rhs, zeros(bk.P2M.r, nrhs));
        }
        if (bk.L2M_f.size() > 0) {  // T:4115-4148
          DMat<T> s = solved(bk.luL2M, bk.ipL2M, bk.invL2M, bk.local_rhs);
          Group g;
          for (int nn : bk.NB) {
            if (nn == k) continue;
            Box<T>& bn = tree[j][nn];
            if (!bn.Eliminated) add(g, bn.particle_rhs, bn.L2P_f[k], OP_N, s, OP_N, T(-1, 0), T(1, 0));
            else add(g, bn.multipole_rhs, bn.L2P_f[k], OP_N, s, OP_N, T(-1, 0), T(1, 0));
          }
          run(g);
          put(bk.multipole_rhs, s);
        } else {
          put(bk.multipole_rhs, newmat(0, nrhs));
        }
      }
    }
    // back_substitution_phase (T:4812-4865)
    {
      const int j = startLevel;
      if (NLb > 0 && NMb > 0) {
        DMat<T> b = newmat(NLb, nrhs);
        std::vector<CopyOp<T>> cops;
        int ri = 0;
        for (int r = 0; r < nBoxes[j]; ++r) {
          DMat<T>& m = tree[j][r].multipole_rhs;
          if (m.size()) cops.push_back(CopyOp<T>{m.p, b.p + ri, m.ld, b.ld, m.r, nrhs, 0, 0});
          ri += tree[j][r].NumLocals;
        }
        copy_batched(cops, cx);
        put(b, solved(luBase, ipBase, invBase, b));
        int idx = 0;
        for (int k = 0; k < nBoxes[j]; ++k) {
          int nm = tree[j][k].NumMultipoles;
          DMat<T> mm = newmat(nm, nrhs);
          if (nm) copy2d(mm, b.rows(idx, nm), cx.st);
          put(tree[j][k].multipoles, mm);
          idx += nm;
        }
        drop(b);
      }
    }
    for (int j = startLevel; j <= nLevels; ++j) {
      for (int k = nBoxes[j] - 1; k >= 0; --k) {
        Box<T>& bk = tree[j][k];
        if (bk.P2P[k].size() != 0) {
          DMat<T> rhs1 = clone(bk.particle_rhs);
          Group g;
          for (int nn : bk.NB) {
            if (nn == k) continue;
            Box<T>& bn = tree[j][nn];
            if (!bn.Eliminated) add(g, rhs1, bk.P2P[nn], OP_N, bn.particles, OP_N, T(-1, 0), T(1, 0), 1);
            else add(g, rhs1, bk.M2P[nn], OP_N, bn.multipoles, OP_N, T(-1, 0), T(1, 0), 1);
          }
          run(g);
          DMat<T> t = solved(bk.luP2P, bk.ipP2P, bk.invP2P, rhs1);
          DMat<T> rhs2 = clone(bk.multipoles);
          {
            Group g2;
            add(g2, rhs2, bk.P2M, OP_N, t, OP_N, T(-1, 0), T(1, 0));
            run(g2);
          }
          if (bk.L2M_f.size() != 0) {  // Atemp == L2M_f
            put(bk.locals, solved(bk.luL2M, bk.ipL2M, bk.invL2M, rhs2));
            drop(rhs2);
          } else {
            drop(rhs2);
            put(bk.locals, newmat(0, nrhs));
          }
          {
            Group g3;
            add(g3, rhs1, bk.L2P, OP_N, bk.locals, OP_N, T(-1, 0), T(1, 0));
            run(g3);
          }
          put(bk.particles, solved(bk.luP2P, bk.ipP2P, bk.invP2P, rhs1));
          drop(rhs1);
          drop(t);
        } else {
          put(bk.particles, newmat(0, nrhs));
        }
        bk.Eliminated = false;
        if (j != nLevels) {
          int idx = 0;
          for (int c = 0; c < 8; ++c) {
            Box<T>& ch = tree[j + 1][8 * k + c];
            DMat<T> mm = newmat(ch.NumMultipoles, nrhs);
            if (ch.NumMultipoles) copy2d(mm, bk.particles.rows(idx, ch.NumMultipoles), cx.st);
            put(ch.multipoles, mm);
            idx += ch.NumMultipoles;
          }
        }
      }
    }
    // getx (T:4874-4907) and reorder (T:466-476)
    DMat<T> Xb = zeros(N, nrhs);
    std::vector<CopyOp<T>> cops;
    int off = 0;
    std::vector<int> order;
    for (int k = 0; k < nBoxes[nLevels]; ++k) {
      Box<T>& b = tree[nLevels][k];
      if (b.particles.size()) cops.push_back(CopyOp<T>{b.particles.p, Xb.p + off, b.particles.ld, Xb.ld, b.particles.r, nrhs, 0, 0});
      off += (int)b.charge.size();
      order.insert(order.end(), b.charge.begin(), b.charge.end());
    }
    copy_batched(cops, cx);
    DMat<T> Xo = zeros(N, nrhs);
    scatter_rows(Xo, Xb, upload_ints(order), cx.st);
    from_T(DMat<cpx<double>>{(cpx<double>*)x_dev, N, nrhs, N}, Xo);
    if (xbox_dev) from_T(DMat<cpx<double>>{(cpx<double>*)xbox_dev, N, nrhs, N}, Xb);
    cx.sync();
    stats_.t_solve = wall() - ts;
    free_tmp_ints();
    drop(Xb); drop(Xo); drop(B);
    (void)t0;
  }
  void to_T(DMat<T> dst, DMat<cpx<double>> src);
  void from_T(DMat<cpx<double>> dst, DMat<T> src);
  void download(DMat<T> src, double* host);

  SolverStats stats() const override {
    SolverStats s = stats_;
    s.factor_bytes = 0;
    for (int j = startLevel; j <= nLevels; ++j)
      for (int k = 0; k < (int)tree[j].size(); ++k) {
        const Box<T>& b = tree[j][k];
        auto acc = [&](const DMat<T>& m) { s.factor_bytes += (double)m.size() * sizeof(T); };
        acc(b.luP2P); acc(b.luL2M); acc(b.P2M); acc(b.L2P);
        for (auto& kv : b.P2P) if (std::find(b.NB.begin(), b.NB.end(), kv.first) != b.NB.end()) acc(kv.second);
        for (auto& kv : b.M2P) acc(kv.second);
        for (auto& kv : b.P2L) acc(kv.second);
        for (auto& kv : b.L2P_f) acc(kv.second);
      }
    s.factor_bytes += (double)luBase.size() * sizeof(T);
    return s;
  }
  bool block(int level, int box, int which, int key, int& r, int& c, std::vector<double>& out) override {
    if (level < 0 || level > nLevels || box < 0 || box >= nBoxes[level]) return false;
    Box<T>& b = tree[level][box];
    DMat<T>* m = nullptr;
    auto find = [&](std::map<int, DMat<T>>& mp) -> DMat<T>* {
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
      default: return false;
    }
    if (!m) return false;
    r = m->r;
    c = m->c;
    out.assign(2 * m->size(), 0.0);
    if (m->size()) {
      DMat<T> t = clone(*m);
      std::vector<double> tmp(2 * m->size());
      download(t, tmp.data());
      out = tmp;
      drop(t);
    }
    return true;
  }
  // getAvgLeafSize (T:490-501), getAvgMultipoles (T:531-551; integer averages as there) and
  // findMemory / findMemoryAfterFactorize (T:4511-4571: stored matrix entries x 8 bytes in
  // GB) from the shapes of the blocks held now.
  void tree_summary(double* out) const override {
    long long sum = 0, nb = 0;
    for (const Box<T>& b : tree[nLevels])
      if (!b.charge.empty()) { ++nb; sum += (long long)b.charge.size(); }
    out[0] = nb ? (double)(sum / nb) : 0.0;
    sum = 0;
    nb = 0;
    for (int j = nLevels; j >= startLevel; --j)
      for (const Box<T>& b : tree[j])
        if (b.P2M.r > 0 && !b.IL.empty()) { ++nb; sum += b.P2M.r; }
    out[1] = nb ? (double)(sum / nb) : 0.0;
    auto entries = [](const std::map<int, DMat<T>>& m, int key) {
      auto it = m.find(key);
      return it == m.end() ? 0.0 : (double)it->second.size();
    };
    double bases = 0, far = 0, near_leaf = 0, near_all = 0;
    for (int j = nLevels; j >= startLevel; --j)
      for (const Box<T>& b : tree[j]) {
        bases += (double)b.L2P.size() + (double)b.P2M.size();
        for (int n : b.IL) far += entries(b.M2L, n);
        double nearb = 0;
        for (int n : b.NB) nearb += entries(b.P2P, n);
        near_all += nearb;
        if (j == nLevels) near_leaf += nearb;
      }
    out[2] = (bases + far + near_leaf) * 8 * 1e-9;
    out[3] = (bases + near_all) * 8 * 1e-9;
  }
  void ranks(int level, std::vector<int>& out) override {
    out.clear();
    if (level < 0 || level > nLevels) return;
    for (auto& b : tree[level]) {
      out.push_back(b.NumMultipoles);
      out.push_back(b.NumLocals);
      out.push_back((int)b.charge.size());
    }
  }

 private:
  std::shared_ptr<DeviceGeometry> G;
  SolverParams P;
  Ctx& cx;
  SolverStats stats_;
  std::unordered_map<void*, size_t> owned_;  // pool blocks this solver holds, with their sizes
  size_t own_in_use_ = 0, own_peak_ = 0;
  bool factorized_ = false;
  int N = 0, nLevels = 0, startLevel = 0, TOL_POW = 6, fill_in_TOL_POW = 6;
  double L = 0, eta = 0, fillInTolerance = 0;
  int rankVec[32];
  std::vector<int> nBoxes;
  std::vector<double> side, px, py, pz;
  double gcx = 0, gcy = 0, gcz = 0;  // centroid of the unknown locations
  std::vector<float> rank_log_;
  std::vector<std::vector<Box<T>>> tree;
  std::vector<std::pair<int, int>> P2Plist, P2L_M2P;
  DMat<T> luBase, invBase;
  int* ipBase = nullptr;
  int NMb = 0, NLb = 0;
  // RCS_VERBOSE=1 prints one line per phase to stderr (flushed)
  bool verbose = std::getenv("RCS_VERBOSE") != nullptr;
  // RCS_FULL_RRQR=1 runs the full partitioned RRQR even when the leading block is orthonormal
  bool full_rrqr_ = std::getenv("RCS_FULL_RRQR") != nullptr;
  double trace_t0 = wall();
  void trace(const char* fmt, int a = 0, int b = 0) {
    if (!verbose) return;
    std::fprintf(stderr, "[rcs %8.3f s] ", wall() - trace_t0);
    std::fprintf(stderr, fmt, a, b);
    std::fprintf(stderr, " (pool %.2f GiB)\n", cx.pool.in_use() / 1073741824.0);
    std::fflush(stderr);
  }
};

template <>
void Solver<cpx<double>>::to_T(DMat<cpx<double>> dst, DMat<cpx<double>> src) { copy2d(dst, src, cx.st); }
template <>
void Solver<cpx<float>>::to_T(DMat<cpx<float>> dst, DMat<cpx<double>> src) { cast_c128_to_c64(dst, src, cx.st); }
template <>
void Solver<cpx<double>>::from_T(DMat<cpx<double>> dst, DMat<cpx<double>> src) { copy2d(dst, src, cx.st); }
template <>
void Solver<cpx<float>>::from_T(DMat<cpx<double>> dst, DMat<cpx<float>> src) { cast_c64_to_c128(dst, src, cx.st); }
template <>
void Solver<cpx<double>>::download(DMat<cpx<double>> src, double* host) {
  HIPCHK(hipMemcpyAsync(host, src.p, sizeof(cpx<double>) * src.size(), hipMemcpyDeviceToHost, cx.st));
  cx.sync();
}
template <>
void Solver<cpx<float>>::download(DMat<cpx<float>> src, double* host) {
  DMat<cpx<double>> t;
  t.r = src.r; t.c = src.c; t.ld = src.r;
  t.p = (cpx<double>*)palloc(sizeof(cpx<double>) * src.size());
  cast_c64_to_c128(t, src, cx.st);
  HIPCHK(hipMemcpyAsync(host, t.p, sizeof(cpx<double>) * src.size(), hipMemcpyDeviceToHost, cx.st));
  cx.sync();
  pfree(t.p);
}

}  // namespace

std::unique_ptr<SolverBase> make_solver(std::shared_ptr<DeviceGeometry> g, const SolverParams& p, int precision,
                                        Ctx& cx) {
  if (precision == PREC_C128) return std::unique_ptr<SolverBase>(new Solver<cpx<double>>(g, p, cx));
  if (precision == PREC_C64) return std::unique_ptr<SolverBase>(new Solver<cpx<float>>(g, p, cx));
  throw std::runtime_error("unknown precision");
}

}  // namespace rcs
