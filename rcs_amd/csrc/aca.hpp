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
  int pred = 0;                 // predicted rank (0: none); sizes U/V and the first step burst
};
struct AcaResult {
  int rank = 0;
  std::vector<int> row_bases, col_bases;  // positions into rows / cols
};
struct AcaStats {
  int burst = 1;            // in: ACA steps launched between host checks (after the first pass)
  long long steps = 0;      // out: step launches
  long long syncs = 0;      // out: host synchronisations
};

// Runs ACA on every problem (tolerance tol) and returns the skeletons. The skeletons do not
// depend on the burst length or on the predictions: a box stops stepping on its own
// criterion inside the kernels.
template <typename T>
void aca_batched(const DevGeometry& g, int kind, double tol, const std::vector<AcaProblem>& probs,
                 std::vector<AcaResult>& out, Ctx& cx, AcaStats& stats);

// out[b] = max column 2-norm of mats[b] (|R(0,0)| of a column-pivoted QR).
template <typename T>
void max_col_norms(const std::vector<DMat<T>>& mats, std::vector<double>& out, Ctx& cx);

}  // namespace rcs
