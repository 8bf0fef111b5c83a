// Batched MoM entry generation (kernel::getMatrix over index lists, src/kernel.cpp:49-61).
#pragma once
#include "geometry.hpp"

namespace rcs {

// One block Z(rows, cols) written column-major to out with leading dimension ld.
// rows == nullptr means the contiguous range row0 .. row0+nr-1 (same for cols).
template <typename T>
struct EntryTask {
  const int* rows;
  const int* cols;
  int row0, col0;
  int nr, nc;
  T* out;
  int ld;
  int tile0;  // first 16x16 tile index of this task in the launch
};

template <typename T>
void launch_entries(const DevGeometry& g, int kind, const EntryTask<T>* d_tasks, int ntask, int ntiles,
                    hipStream_t st);

inline int entry_tiles(int nr, int nc) { return ((nr + 15) / 16) * ((nc + 15) / 16); }

}  // namespace rcs
