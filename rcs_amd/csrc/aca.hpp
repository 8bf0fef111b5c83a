// Batched adaptive cross approximation (LowRank<K>::ACA_only_nodes2, src/ACA.cpp:639-859).
//
// All boxes of one tree level (one direction: incoming or outgoing) run the pivoted
// cross approximation together, five launches per step:
//  - long entries: one workgroup per 1024-entry chunk of every active box's long side. It
//    normalises the previous step's long vector into V and generates this step's long
//    entries (the matrix entries are computed in the kernel).
//  - long: same grid. It takes the previous step's dots with every earlier V column (for
//    the Frobenius estimate) and forms this step's residual long vector from the same
//    pass over V.
//  - pivot: one workgroup per box. It finishes the previous step's Frobenius estimate,
//    applies the stopping test, then picks the long pivot (first maximum).
//  - short entries: one workgroup per 256 short indices.
//  - short: one workgroup per box. It forms the residual short vector and its dots with
//    the earlier U columns, and picks the next short pivot.
// Per-box state stays in HBM; the host only reads the active flags and ranks.
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
  double grow_s = 0;        // out: seconds spent growing U/V (allocation and copies)
  double wait_s = 0;        // out: seconds the host waited for step bursts
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
