// Batched adaptive cross approximation (LowRank<K>::ACA_only_nodes2, src/ACA.cpp:639-859).
//
// All boxes of one tree level (one direction: incoming or outgoing) run the pivoted
// cross approximation together: every step is one launch that generates the residual
// "long" vector of every active box (row or column of the box x interaction-list
// block, entries computed in the kernel), a pivot launch (wavefront argmax), a
// normalisation launch and one launch for the short vectors, stopping test and next
// pivot. Per-box state stays in HBM; the host only reads the active flags and ranks.
#pragma once
#include "ctx.hpp"
#include "dense.hpp"
#include "entries.hpp"

namespace rcs {

struct AcaProblem {
  std::vector<int> rows, cols;  // global unknown indices
};
struct AcaResult {
  int rank = 0;
  std::vector<int> row_bases, col_bases;  // positions into rows / cols
};

// Runs ACA on every problem (tolerance tol) and returns the skeletons.
template <typename T>
void aca_batched(const DevGeometry& g, int kind, double tol, const std::vector<AcaProblem>& probs,
                 std::vector<AcaResult>& out, Ctx& cx);

// out[b] = max column 2-norm of mats[b] (|R(0,0)| of a column-pivoted QR).
template <typename T>
void max_col_norms(const std::vector<DMat<T>>& mats, std::vector<double>& out, Ctx& cx);

}  // namespace rcs
