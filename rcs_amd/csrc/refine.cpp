// FP64 iterative refinement of a (possibly complex64) AIFMM factorization: restarted
// flexible GMRES on Z x = b, right-preconditioned by the factorization's solve, with the
// residual formed by a separately assembled FP64 compressed operator. The preconditioned
// vectors z_k = M^{-1} v_k are kept and the update is x += [z_0 .. z_k] y: a complex64
// solve is linear only up to its rounding, which plain right-preconditioned GMRES (update
// M^{-1} (V y)) assumes exact. Classical Gram-Schmidt run twice (CGS2) orthogonalises each
// new Krylov vector with two products against the basis. All vectors live in HBM; the host
// only keeps the small Hessenberg system and its Givens rotations (one synchronisation per
// iteration).
//
// No counterpart in the reference (its driver checks a dense-LU residual, test/sphere.cpp:
// 331-338); this is the refinement step of the mixed-precision path.
#include "refine.hpp"

#include <cmath>
#include <complex>

namespace rcs {

namespace {
using Z = cpx<double>;
using cd = std::complex<double>;

DMat<Z> dmat(Ctx& cx, int r, int c) {
  DMat<Z> m;
  m.r = r;
  m.c = c;
  m.ld = std::max(r, 1);
  m.p = (Z*)cx.pool.alloc(sizeof(Z) * std::max<size_t>((size_t)r * c, 1));
  return m;
}
void gemm1(Ctx& cx, DMat<Z> C, DMat<Z> A, int opA, DMat<Z> B, Z alpha, Z beta) {
  std::vector<GemmOp<Z>> g(1);
  g[0] = GemmOp<Z>{};
  g[0].A = A.p; g[0].lda = A.ld; g[0].opA = opA;
  g[0].B = B.p; g[0].ldb = B.ld; g[0].opB = OP_N;
  g[0].C = C.p; g[0].ldc = C.ld;
  g[0].m = C.r; g[0].n = C.c; g[0].k = opA == OP_N ? A.c : A.r;
  g[0].alpha = alpha; g[0].beta = beta;
  gemm_grouped(g, cx);
}
double norm2(Ctx& cx, DMat<Z> v) {
  sqnorms(std::vector<DMat<Z>>{v}, cx.d_scalars, cx);
  return std::sqrt(cx.fetch_scalars(1)[0]);
}
}  // namespace

RefineResult gmres_refine(SolverBase& op, SolverBase& pre, const cpx<double>* b_dev, int nrhs, cpx<double>* x_dev,
                          double tol, int maxit, int restart, Ctx& cx) {
  const int N = op.size();
  if (pre.size() != N) throw std::runtime_error("refine: operator and factorization differ in size");
  if (restart < 1) restart = 30;
  RefineResult res;
  pre.prepare_repeated_solves();
  DMat<Z> V = dmat(cx, N, restart + 1), Zk = dmat(cx, N, restart), r = dmat(cx, N, 1), w = dmat(cx, N, 1);
  DMat<Z> h = dmat(cx, restart + 1, 2);  // two CGS passes
  for (int c = 0; c < nrhs; ++c) {
    DMat<Z> b{const_cast<Z*>(b_dev) + (size_t)c * N, N, 1, N};
    DMat<Z> x{x_dev + (size_t)c * N, N, 1, N};
    const double bn = norm2(cx, b);
    pre.solve_device(b.p, 1, x.p, nullptr);  // x0 = M^{-1} b
    int it = 0;
    double rel = 0;
    for (;;) {
      // r = b - Z x
      op.apply_device(x.p, 1, w.p);
      copy2d(r, b, cx.st);
      axpy2d(r, w, Z(-1, 0), cx.st);
      const double beta = norm2(cx, r);
      rel = bn > 0 ? beta / bn : 0;
      if (rel <= tol || it >= maxit || beta == 0) break;
      DMat<Z> v0 = V.cols(0, 1);
      zero2d(v0, cx.st);
      axpy2d(v0, r, Z(1.0 / beta, 0), cx.st);
      std::vector<std::vector<cd>> H(restart + 1, std::vector<cd>(restart, cd(0)));
      std::vector<cd> g(restart + 1, cd(0)), cs(restart), sn(restart);
      g[0] = beta;
      int k = 0;
      for (; k < restart && it < maxit; ++k, ++it) {
        // z_k = M^{-1} v_k, w = Z z_k
        pre.solve_device(V.cols(k, 1).p, 1, Zk.cols(k, 1).p, nullptr);
        op.apply_device(Zk.cols(k, 1).p, 1, w.p);
        // CGS2 against v_0..v_k
        DMat<Z> Vk = V.cols(0, k + 1);
        DMat<Z> h1 = h.cols(0, 1).rows(0, k + 1), h2 = h.cols(1, 1).rows(0, k + 1);
        gemm1(cx, h1, Vk, OP_C, w, Z(1, 0), Z(0, 0));
        gemm1(cx, w, Vk, OP_N, h1, Z(-1, 0), Z(1, 0));
        gemm1(cx, h2, Vk, OP_C, w, Z(1, 0), Z(0, 0));
        gemm1(cx, w, Vk, OP_N, h2, Z(-1, 0), Z(1, 0));
        sqnorms(std::vector<DMat<Z>>{w}, cx.d_scalars, cx);
        std::vector<Z> hh(2 * (k + 1));
        HIPCHK(hipMemcpyAsync(hh.data(), h1.p, sizeof(Z) * (k + 1), hipMemcpyDeviceToHost, cx.st));
        HIPCHK(hipMemcpyAsync(hh.data() + k + 1, h2.p, sizeof(Z) * (k + 1), hipMemcpyDeviceToHost, cx.st));
        const double wn = std::sqrt(cx.fetch_scalars(1)[0]);  // synchronises
        for (int i = 0; i <= k; ++i) H[i][k] = cd(hh[i].re + hh[k + 1 + i].re, hh[i].im + hh[k + 1 + i].im);
        H[k + 1][k] = wn;
        DMat<Z> vn = V.cols(k + 1, 1);
        zero2d(vn, cx.st);
        if (wn > 0) axpy2d(vn, w, Z(1.0 / wn, 0), cx.st);
        // Givens rotations on the host
        for (int i = 0; i < k; ++i) {
          cd t = std::conj(cs[i]) * H[i][k] + std::conj(sn[i]) * H[i + 1][k];
          H[i + 1][k] = -sn[i] * H[i][k] + cs[i] * H[i + 1][k];
          H[i][k] = t;
        }
        const double a = std::abs(H[k][k]), bb = std::abs(H[k + 1][k]);
        const double den = std::sqrt(a * a + bb * bb);
        if (den == 0) { cs[k] = 1; sn[k] = 0; }
        else {
          cs[k] = H[k][k] / den;
          sn[k] = H[k + 1][k] / den;
        }
        H[k][k] = std::conj(cs[k]) * H[k][k] + std::conj(sn[k]) * H[k + 1][k];
        H[k + 1][k] = 0;
        g[k + 1] = -sn[k] * g[k];
        g[k] = std::conj(cs[k]) * g[k];
        if (std::abs(g[k + 1]) <= tol * bn || wn == 0) { ++k; ++it; break; }
      }
      // y = H^{-1} g (upper triangular k x k), x += [z_0 .. z_{k-1}] y
      std::vector<cd> y(k);
      for (int i = k - 1; i >= 0; --i) {
        cd s = g[i];
        for (int j = i + 1; j < k; ++j) s -= H[i][j] * y[j];
        y[i] = s / H[i][i];
      }
      std::vector<Z> yz(k);
      for (int i = 0; i < k; ++i) yz[i] = Z(y[i].real(), y[i].imag());
      DMat<Z> yd = h.cols(0, 1).rows(0, k);
      cx.upload(yd.p, yz.data(), sizeof(Z) * k);
      gemm1(cx, x, Zk.cols(0, k), OP_N, yd, Z(1, 0), Z(1, 0));
      ++res.restarts;
    }
    res.iters = std::max(res.iters, it);
    res.relres = std::max(res.relres, rel);
  }
  cx.sync();
  for (DMat<Z>* m : {&V, &Zk, &r, &w, &h}) cx.pool.free(m->p);
  return res;
}

}  // namespace rcs
