/*
 * C ABI of the MI355X AIFMM library (librcs_amd.so).
 *
 * Plain pointers and sizes only. Complex arrays are interleaved (re, im) doubles and
 * matrices are column-major. Every call returns 0 on success or a negative status;
 * rcs_last_error() gives the message of the last failure on the calling thread.
 * All GPU work runs on one device (the current HIP device at rcs_init) and one stream.
 *
 * Each entry point names the reference interface (vaishna77/RCS) it replaces.
 */
#ifndef RCS_AMD_H
#define RCS_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rcs_mesh_s* rcs_mesh_t;
typedef struct rcs_solver_s* rcs_solver_t;

enum { RCS_KERNEL_EFIE = 0, RCS_KERNEL_LAPLACE = 1 };
enum { RCS_PREC_C128 = 0, RCS_PREC_C64 = 1 };

/* Library set-up: HIP device, stream, HBM block pool. No counterpart in the reference
 * (it is a host-only program); called implicitly by the other entry points. */
int rcs_init(void);
const char* rcs_last_error(void);

/* EFIEMatrix(inputsToKernelClass) (src/EFIEMatrix.cpp:5-20, mesh ingest and RWG
 * numbering E:31-209) or ElectrostaticsMatrix(inputsToKernelClass)
 * (src/ElectrostaticsMatrix.cpp:5-9). nodes: n_nodes x 3 (the node file lines);
 * tris: n_tris x 3, 1-based node indices (the patch file lines); conductor (the optional
 * fourth column, E:78-79) may be NULL, which gives every patch conductor 0 as there.
 * freq replaces the compile-time constant of include/includes.hpp:58. */
int rcs_mesh_create(const double* nodes, int n_nodes, const int* tris, int n_tris, const int* conductor,
                    int kernel_kind, double freq, rcs_mesh_t* out);
/* The same mesh with host tables only (no device, no HIP call): enough for the numbering,
 * patches, right-hand side and RCS entry points below; rcs_entries, rcs_aca,
 * rcs_solver_create and rcs_operator_create reject it. */
int rcs_mesh_create_host(const double* nodes, int n_nodes, const int* tris, int n_tris, const int* conductor,
                         int kernel_kind, double freq, rcs_mesh_t* out);
void rcs_mesh_destroy(rcs_mesh_t m);
/* EFIEMatrix::N (E:97-99) / ElectrostaticsMatrix::N. */
int rcs_mesh_size(rcs_mesh_t m);
/* Unknown locations handed to inputsToSolverClass::locations: RWG edge midpoints
 * (test/sphere.cpp:143-152) or patch centroids. xyz: N x 3. */
int rcs_mesh_locations(rcs_mesh_t m, double* xyz);
/* RWG table in basis order: n1 n2 T+ T- free+ free- per basis (edge2PatchMap,
 * E:129-209 and getIndexOfThirdNode E:353-373), plus the edge lengths. */
int rcs_mesh_rwg(rcs_mesh_t m, int* out6, double* length);
/* Patch count and patchList (include/Patch.hpp; centroid, area and largestEdge from
 * E:72-95, the reference's largestEdge E:247-254 looking at two of the three edges): out5 holds
 * centroid x y z, area, largestEdge per patch; conductor (may be NULL) the conductor index. */
int rcs_mesh_npatch(rcs_mesh_t m);
int rcs_mesh_patches(rcs_mesh_t m, double* out5, int* conductor);

/* kernel::getMatrix(rows, cols) (src/kernel.cpp:49-61) over getMatrixEntry
 * (E:698-783 / src/ElectrostaticsMatrix.cpp:293-340), computed on the GPU in the
 * requested precision. out: nr x nc complex. */
int rcs_entries(rcs_mesh_t m, int precision, const int* rows, int nr, const int* cols, int nc, double* out);

/* EFIEMatrix::getRhs (E:785-829): plane wave with field vector E and phase vector
 * kvec (the reference hard-codes kvec = (0, 0, 1)). out: N complex. */
int rcs_rhs(rcs_mesh_t m, const double* E, const double* kvec, double* out);

/* getJAtCentroids (E:1100-1130) + getRCS_OnePointQuadrature (E:916-930) at n angles. */
int rcs_rcs(rcs_mesh_t m, const double* coef, const double* theta, const double* phi, int n, double* out);
/* EFIEMatrix::getJAtCentroids (E:1100-1130): J is n_patch x 3 complex, column-major (the
 * reference's MatrixXcd JAtCentroids). */
int rcs_j_at_centroids(rcs_mesh_t m, const double* coef, double* J);
/* EFIEMatrix::getRCS_OnePointQuadrature(theta, phi, JAtCentroids) (E:916-930) at n angles, m^2. */
int rcs_rcs_from_j(rcs_mesh_t m, const double* J, const double* theta, const double* phi, int n, double* out);

/* LowRank<K>(K, tol, rows, cols).ACA_only_nodes2 (include/ACA.hpp:16,26; src/ACA.cpp:639-859)
 * on the GPU. row_bases/col_bases must hold min(nr, nc) entries. */
int rcs_aca(rcs_mesh_t m, int precision, double tol, const int* rows, int nr, const int* cols, int nc, int* rank,
            int* row_bases, int* col_bases);

/* RRQR(A, TOL_POW1, TOL_POW2, partition, inputPivot, flag, fillInRank)
 * .colPivHouseholderQRWithPartition() (include/rrqr.hpp:20-22, src/rrqr.cpp:69-182)
 * on the GPU. Q must hold m x min(m, n) complex; the first *rank columns are Q_trunc. */
int rcs_rrqr(int precision, const double* A, int m, int n, int tol_pow1, int tol_pow2, int partition,
             double input_pivot, int fillin_rank, double* Q, int* rank, double* max_pivot);

/* Eigen::PartialPivLU<Mat>(A).solve(B) as used at E:831-837 and throughout the tree:
 * blocked GPU GETRF (first-max partial pivoting) + GETRS. */
int rcs_lu_solve(int precision, const double* A, int n, const double* B, int nrhs, double* X);

/* Dense complex product C = op(A) op(B) (op: 0 none, 1 conjugate transpose) on the GPU,
 * the operation behind every Mat product of the tree. */
int rcs_gemm(int precision, int opA, int opB, int m, int n, int k, const double* A, int lda, const double* B,
             int ldb, double* C);

/* AIFMM<K>(inputsToSolverClass, K*) (src/AIFMM.cpp:6-93): tree set-up, nested ACA
 * skeletons, L2P/P2M, M2L and leaf P2P assembly, all on the GPU. bbox = xmin xmax ymin
 * ymax zmin zmax of the locations (test/sphere.cpp:159-174). variant 0 = the compiled
 * memory-optimised tree, 1 = AIFMMTree2D_3D parameters. */
int rcs_solver_create(rcs_mesh_t m, const double* bbox, int leaf_size, int tol_pow, int fillin_tol, int fillin_rank,
                      int precision, int variant, rcs_solver_t* out);
void rcs_solver_destroy(rcs_solver_t s);
/* AIFMM<K>::factorize (src/AIFMM.cpp:95-100 -> FMM2DTree::eliminate_phase_efficient T:4599-4747). */
int rcs_solver_factorize(rcs_solver_t s);
/* AIFMM<K>::backSubstitute1/2/3 + getPhi (src/AIFMM.cpp:119-147): nrhs right-hand sides
 * (N x nrhs complex). x gets getPhi's input order; x_box (may be NULL) gets solve()'s
 * box order (src/AIFMM.cpp:103-110). */
int rcs_solver_solve(rcs_solver_t s, const double* rhs, int nrhs, double* x, double* x_box);
/* FP64 compressed operator for residuals: the AIFMM<K> assembly (src/AIFMM.cpp:6-93, tree
 * set-up, nested ACA skeletons, M2L, leaf P2P) at its own tolerance, never factorized, and
 * applied as y = Z x through the near field plus the nested far field (upward P2M, M2L,
 * downward L2P). No counterpart in the reference, whose driver measures the error of a
 * solve against the dense matrix (test/sphere.cpp:331-338); it is the FP64 residual of the
 * mixed-precision path. x, y: N x nrhs complex, input order. */
typedef struct rcs_operator_s* rcs_operator_t;
int rcs_operator_create(rcs_mesh_t m, const double* bbox, int leaf_size, int tol_pow, rcs_operator_t* out);
void rcs_operator_destroy(rcs_operator_t o);
int rcs_operator_apply(rcs_operator_t o, const double* x, int nrhs, double* y);
/* Restarted GMRES(restart) on Z x = rhs with the FP64 operator o, right-preconditioned by
 * the factorized solver s (any precision), CGS2 orthogonalisation, until
 * ||rhs - Z x|| <= tol ||rhs|| or maxit iterations. x: N x nrhs complex (input order).
 * iters / relres (may be NULL): largest iteration count and final relative residual.
 * No counterpart in the reference (see rcs_operator_create). */
int rcs_solver_refine(rcs_solver_t s, rcs_operator_t o, const double* rhs, int nrhs, double* x, double tol, int maxit,
                      int restart, int* iters, double* relres);

/* Learned ACA rank predictor. No counterpart in the reference: its ACA (src/ACA.cpp:639-859)
 * only stops adaptively, one step at a time. An MLP (8 inputs, 2 x 64 GELU, 1 output =
 * log(1 + rank); BF16 weights, FP32 accumulate) with HIP forward, backward and Adam kernels.
 * params: 4801 floats (W1[64][8], b1[64], W2[64][64], b2[64], W3[64], b3); mean/scale: the
 * 8-feature standardisation. X: n x 8 floats, row-major; t: targets log(1 + rank). */
typedef struct rcs_mlp_s* rcs_mlp_t;
int rcs_mlp_create(unsigned seed, rcs_mlp_t* out);
void rcs_mlp_destroy(rcs_mlp_t m);
int rcs_mlp_get(rcs_mlp_t m, float* params, float* mean, float* scale);
int rcs_mlp_set(rcs_mlp_t m, const float* params, const float* mean, const float* scale);
int rcs_mlp_forward(rcs_mlp_t m, const float* X, int n, float* y);
/* loss = mean (y - t)^2 and d loss / d params (4801 floats) at the BF16 weights. */
int rcs_mlp_gradient(rcs_mlp_t m, const float* X, const float* t, int n, float* grad, double* loss);
/* Full-batch Adam (iters steps, learning rate lr) after standardising with X's statistics. */
int rcs_mlp_train(rcs_mlp_t m, const float* X, const float* t, int n, int iters, float lr, double* loss);
/* Solvers created after this call predict every ACA rank with m (NULL: none) to size the
 * ACA buffers and launch that many steps before the first host check; the adaptive stop
 * still decides where each ACA ends (a box keeps iterating past a low prediction). */
int rcs_set_rank_model(rcs_mlp_t m);
/* ACA steps launched between host checks once the predicted pass is over (default 1, the
 * reference's one step per check). Skeletons do not depend on it. */
int rcs_set_aca_burst(int steps);
/* Selects the library context ("lane": HIP stream, HBM pool, staging buffer, profiler) of the
 * calling host thread; lane 0 is the default. Currently only lane 0 is accepted (other lanes
 * return an error until concurrent lanes have been verified on the GPU). Meshes, solvers and operators keep the lane
 * they were created in, and calls on them run on its stream, so independent problems created
 * in different lanes from different host threads run concurrently on the GPU (independent
 * scatterers). A solver and the operator refining it must share a lane; the objects of one
 * lane are used from one host thread at a time (a rank model set with rcs_set_rank_model
 * runs its predictions in its own lane). No counterpart in the reference (its solver is
 * single-threaded per process). */
int rcs_set_lane(int lane);
/* Rank log of the assembly: 14 floats per ACA problem (8 predictor features, rank found,
 * rank predicted or 0, level, 0 incoming / 1 outgoing, short and long side lengths). Writes
 * up to cap_rows rows (out may be NULL) and returns the number of rows. */
int rcs_solver_rank_log(rcs_solver_t s, float* out, int cap_rows);
/* ACA step launches, host checks and seconds spent in the assembly's ACA. */
int rcs_solver_aca_stats(rcs_solver_t s, long long* steps, long long* syncs, double* seconds);

/* Statistics (the driver's prints, test/sphere.cpp:217-262):
 * info[0..4] = nLevels, startLevel, n_rrqr, n_compress_p2p, n_compress_p2l;
 * vals[0..10] = t_assemble, t_factor, t_solve, gemm_flops, getrf_flops, getrs_flops, qr_flops,
 *               t_getrf, t_schur, t_compress, factor_bytes; vals[11] = peak HBM bytes. */
int rcs_solver_stats(rcs_solver_t s, int* info, double* vals);
/* The statistics the reference driver prints through AIFMM::A (test/sphere.cpp:217-245):
 * out4 = getAvgLeafSize (T:490-501), getAvgMultipoles (T:531-551), findMemory and
 * findMemoryAfterFactorize (T:4511-4571, stored entries x 8 bytes in GB) of the current state. */
int rcs_solver_tree_summary(rcs_solver_t s, double* out4);
/* Per-box NumMultipoles, NumLocals, particle count at a level (3 ints per box). Returns
 * the number of boxes written, or a negative status. */
int rcs_solver_ranks(rcs_solver_t s, int level, int* out, int cap);

/* Read one stored block of the tree (FMM2DBox members, include/AIFMMBox.hpp:45-59):
 * which = 0 L2P, 1 P2M, 2 M2L[key], 3 P2P[key], 4 M2P[key], 5 P2L[key], 6 L2M_f.
 * shape receives rows, cols; out holds cap complex values. Returns 1 when the block
 * does not exist. */
int rcs_solver_block(rcs_solver_t s, int level, int box, int which, int key, int* shape, double* out,
                     long long cap);

/* Phase timing with HIP events on the library stream. No counterpart in the reference
 * (its driver prints omp_get_wtime phases, T:4610-4744); used by bench.py for the
 * roofline figure of the dominant kernel. Tags: 0 Schur fill-in GEMMs, 1 GETRF,
 * 2 GETRS, 3 entry generation, 4 ACA, 5 RRQR, 6 basis updates. work = flops
 * (entries for tag 3). */
int rcs_profile_enable(int on);
int rcs_profile_reset(void);
int rcs_profile_get(int tag, double* work, double* ms, long long* launches);

#ifdef __cplusplus
}
#endif
#endif
