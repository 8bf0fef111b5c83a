// AIFMM hierarchical solver on one GPU (FMM2DTree<K>, include/AIFMMTree2D_3D_memory_optimized.hpp,
// and AIFMM<K>, src/AIFMM.cpp).
//
// The host keeps the tree bookkeeping (interaction lists, per-box block maps, pending
// fill-in pairs, elimination order) exactly as the reference orders it; every block
// lives in HBM and every arithmetic operation is a HIP kernel. Work that the reference
// performs as separate small products inside its per-box loops is issued as grouped
// launches: one multi-RHS LU solve per eliminated box, one grouped GEMM for all Schur
// fill-ins of that box, one grouped GEMM per basis update.
#pragma once
#include <map>
#include <memory>

#include "aca.hpp"
#include "ctx.hpp"
#include "dense.hpp"
#include "entries.hpp"
#include "geometry.hpp"

namespace rcs {

class RankMLP;

struct SolverParams {
  double bbox[6];
  int leafSize = 64, tol_pow = 6, fillin_tol = 6, fillin_rank = 0;
  int variant = 0;  // 0: memory-optimized tree (reference default); 1: AIFMMTree2D_3D parameters
  RankMLP* rank_model = nullptr;  // predicted ACA ranks (sizes and step bursts); null: none
  int aca_burst = 1;              // ACA steps between host checks after the predicted pass
};

struct SolverStats {
  int nLevels = 0, startLevel = 0;
  long long n_rrqr = 0, n_compress_p2p = 0, n_compress_p2l = 0;
  double t_assemble = 0, t_factor = 0, t_solve = 0;
  double gemm_flops = 0, getrf_flops = 0, getrs_flops = 0, qr_flops = 0;
  double t_getrf = 0, t_schur = 0, t_compress = 0;
  size_t peak_bytes = 0;
  double factor_bytes = 0;  // bytes of the stored factor read once by a solve
  long long aca_steps = 0, aca_syncs = 0;  // ACA step launches and host checks of the assembly
  double t_aca = 0;
};

// One ACA problem of the skeleton construction: the predictor's 8 features, the rank the
// adaptive ACA found, the rank predicted (0 without a model), the level, 0 incoming / 1
// outgoing.
constexpr int RANK_LOG_COLS = 14;

class DeviceGeometry {
 public:
  HostGeometry h;
  DevGeometry d{};
  Ctx* cx = nullptr;
  std::vector<void*> bufs;
  void upload(Ctx& c);
  void release();
  ~DeviceGeometry() { release(); }
};

class SolverBase {
 public:
  virtual ~SolverBase() {}
  virtual void assemble() = 0;
  virtual void factorize() = 0;
  // rhs and x are interleaved complex128 arrays of N x nrhs (column-major), x in input order
  virtual void solve(const double* rhs, int nrhs, double* x, double* x_box) = 0;
  // The same on device complex128 arrays (ld N), stream ordered; x_box may be null.
  virtual void solve_device(const void* rhs, int nrhs, void* x, void* x_box) = 0;
  // y = Z x with the assembled compressed operator (before factorize), device arrays.
  virtual void apply_device(const void* x, int nrhs, void* y) = 0;
  // Before many solves with few right-hand sides (iterative refinement): form the inverses
  // of the factored diagonal blocks once, so that every later solve multiplies by them
  // instead of running triangular solves.
  virtual void prepare_repeated_solves() {}
  virtual int size() const = 0;
  virtual Ctx& context() = 0;
  // rows of RANK_LOG_COLS floats, one per ACA problem of the assembly
  virtual const std::vector<float>& rank_log() const = 0;
  virtual SolverStats stats() const = 0;
  virtual void ranks(int level, std::vector<int>& out) = 0;
  // The tree statistics the reference driver prints: getAvgLeafSize, getAvgMultipoles,
  // findMemory and findMemoryAfterFactorize of the current state, in the reference's units.
  virtual void tree_summary(double* out4) const = 0;
  // Copy one stored block to the host as complex128 (which: 0 L2P, 1 P2M, 2 M2L[key],
  // 3 P2P[key], 4 M2P[key], 5 P2L[key], 6 L2M_f). Returns false if absent.
  virtual bool block(int level, int box, int which, int key, int& r, int& c, std::vector<double>& out) = 0;
};

std::unique_ptr<SolverBase> make_solver(std::shared_ptr<DeviceGeometry> g, const SolverParams& p, int precision,
                                        Ctx& cx);

}  // namespace rcs
