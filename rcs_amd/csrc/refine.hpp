// Mixed-precision refinement: restarted GMRES with an FP64 operator and a factorization
// as right preconditioner (see refine.cpp).
#pragma once
#include "tree.hpp"

namespace rcs {

struct RefineResult {
  int iters = 0;       // largest GMRES iteration count over the right-hand sides
  int restarts = 0;    // total restart cycles
  double relres = 0;   // largest final ||b - Z x|| / ||b|| (Z = the FP64 operator)
};

// b, x: device complex128, N x nrhs (ld N), input order. x is overwritten (starts from
// the preconditioner's solution M^{-1} b). op must be assembled and not factorized; pre
// must be factorized. Stops when ||b - Z x|| <= tol ||b|| or after maxit iterations.
RefineResult gmres_refine(SolverBase& op, SolverBase& pre, const cpx<double>* b, int nrhs, cpx<double>* x,
                          double tol, int maxit, int restart, Ctx& cx);

}  // namespace rcs
