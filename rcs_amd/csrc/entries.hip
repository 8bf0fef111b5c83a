// MoM entry generation on the GPU: RWG EFIE (src/EFIEMatrix.cpp:247-783) and the
// piecewise-constant Laplace single layer (src/ElectrostaticsMatrix.cpp:140-340).
//
// One thread computes one entry Z(rows[i], cols[j]) of a block. Many blocks (index-list
// pairs) go through one launch: the grid enumerates 16x16 tiles of every block, found
// from a prefix sum of tile counts. Threads of a wavefront cover consecutive rows of a
// column, so the column-major stores coalesce. R is the arithmetic type (double for the
// complex128 path, float for the complex64 path); geometry is read in FP64 and rounded.
#include "entries_dev.hpp"

namespace rcs {

template <typename R, typename T>
__global__ void __launch_bounds__(256) entries_kernel(DevGeometry g, int kind, const EntryTask<T>* tasks, int ntask) {
  // locate the task owning this tile
  const int tile = blockIdx.x;
  int lo = 0, hi = ntask - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (tasks[mid].tile0 <= tile) lo = mid;
    else hi = mid - 1;
  }
  const EntryTask<T> tk = tasks[lo];
  const int t = tile - tk.tile0;
  const int tiles_r = (tk.nr + 15) >> 4;
  const int tr = t % tiles_r, tc = t / tiles_r;
  const int r = tr * 16 + (threadIdx.x & 15);
  const int c = tc * 16 + (threadIdx.x >> 4);
  if (r >= tk.nr || c >= tk.nc) return;
  const int i = tk.rows ? tk.rows[r] : tk.row0 + r;
  const int j = tk.cols ? tk.cols[c] : tk.col0 + c;
  cpx<R> v = kernel_entry<R>(g, kind, i, j);
  T out;
  out.re = v.re;
  out.im = v.im;
  tk.out[(size_t)r + (size_t)c * tk.ld] = out;
}

template <typename T>
void launch_entries(const DevGeometry& g, int kind, const EntryTask<T>* d_tasks, int ntask, int ntiles,
                    hipStream_t st) {
  if (ntiles <= 0) return;
  using R = decltype(T().re);
  hipLaunchKernelGGL((entries_kernel<R, T>), dim3(ntiles), dim3(256), 0, st, g, kind, d_tasks, ntask);
  HIPCHK(hipGetLastError());
}

template void launch_entries<cpx<double>>(const DevGeometry&, int, const EntryTask<cpx<double>>*, int, int, hipStream_t);
template void launch_entries<cpx<float>>(const DevGeometry&, int, const EntryTask<cpx<float>>*, int, int, hipStream_t);

}  // namespace rcs
