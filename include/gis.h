/*
 * gis.h -- C ABI of the MI355X (gfx950) gis sparse iterative solver library.
 *
 * Replaces cuFoam's C++ template ABI gis/gis.h:1-95 (free functions taking
 * SingleGPU::SparseCSR / SingleGPU::Vector by reference) with a handle-based
 * extern "C" interface over plain pointers and sizes.  The library owns all
 * device memory behind the handle; the OpenFOAM lduMatrix::solver wrappers
 * (cuFoamCG.C:38-89 and siblings) hand over host arrays.  Every `_device`
 * variant takes device pointers instead (e.g. from torch tensors) and keeps
 * the data resident.
 *
 * Conventions
 *   - n_cells = N (matrix order), n_faces = M (one upper and one lower
 *     coefficient per face), int32 indices, FP64 values (gis/gis.h:7-8).
 *   - lower_addr[f] = owner, upper_addr[f] = neighbour, faces sorted by owner
 *     (OpenFOAM upper-triangular order).  A(owner, nbr) = upper[f],
 *     A(nbr, owner) = lower[f].
 *   - All calls are synchronous with respect to the host unless named
 *     `_async`; work runs on the handle's HIP stream.
 *   - Return value: GIS_OK, an error code (< 10), or a solve outcome (>= 10)
 *     which is not an error: the outputs are valid.
 *   - Handles are independent; one handle must not be used from two threads
 *     at once.
 */
#ifndef GIS_AMD_GIS_H
#define GIS_AMD_GIS_H

#ifdef __cplusplus
extern "C" {
#endif

#define GIS_API __attribute__((visibility("default")))

/* ---- status codes (replace exit(1) in check.h:21-29 and BiCG's bool) ---- */
#define GIS_OK 0
#define GIS_ERR_ARG 1          /* invalid argument or addressing                  */
#define GIS_ERR_HIP 2          /* HIP runtime error (message: gis_last_error)     */
#define GIS_ERR_SHAPE 3        /* size mismatch / limits (SolverSingleGPU.h:177)   */
#define GIS_ERR_STATE 4        /* pattern or values not set                      */
#define GIS_ERR_NOMEM 5
#define GIS_BREAKDOWN 10       /* BiCG |xi| < int_bound (SolverSingleGPU.h:629)   */
#define GIS_NOT_CONVERGED 11   /* iteration cap reached above the threshold       */
#define GIS_NAN 12             /* residual norm became NaN                         */
#define GIS_ERR_SYNC 13        /* a cross-workgroup reduction did not complete     */

/* ---- solver methods (gis/gis.h:13-87; wrappers cuFoam<Name>) ---- */
#define GIS_CG 0
#define GIS_BICG 1
#define GIS_JACOBI 2
#define GIS_MR 3
#define GIS_MRJACOBI 4
#define GIS_JMR 5              /* PreconditionedMRJacobi (gis/gis.cu:84-92)        */
#define GIS_MRS 6
#define GIS_RICHARDSON 7

typedef struct gis_context *gis_handle;

/* Solve controls.  Defaults (gis_default_params) reproduce gis/gis.h:13-87. */
/* What the history records (DetailType, support.h:7-9): the residual norm
 * ||r_k||_2 per iteration, or the elapsed time in seconds since the loop started
 * (entry 0 is 0.0), stamped on the device's constant-rate wall clock when
 * iteration k completes (SolverSingleGPU.h:397-444 and every other method). */
#define GIS_DETAIL_ERROR 0
#define GIS_DETAIL_TIME 1

typedef struct {
    double error;        /* absolute ||r||_2 threshold (CG :412 uses norm > error) */
    long long max_steps; /* <= 0: unbounded (SolverSingleGPU.h:394-395)            */
    double omega;        /* Richardson factor; |omega| < 1e-15 -> OptimizePower     */
    int power_steps;     /* Richardson power iterations (30)                        */
    double int_bound;    /* BiCG breakdown bound (1e-10)                            */
    int detail_type;     /* GIS_DETAIL_ERROR (default) or GIS_DETAIL_TIME: what
                            gis_solve / gis_solve_pcg write into hist            */
} gis_params;

/* solverPerformance fields the wrappers fill (cuFoamCG.C:80-83) */
typedef struct {
    double initial_residual; /* history.front(): unnormalised ||r_0||_2              */
    double final_residual;   /* history.back()                                      */
    int n_iterations;        /* history.size() - 1                                   */
    int status;              /* GIS_OK / GIS_BREAKDOWN / GIS_NOT_CONVERGED / GIS_NAN  */
    double norm_factor;      /* NormFactor(A, b, x) (gis/gis.cu:106-123)              */
} gis_performance;

GIS_API void gis_default_params(gis_params *p);

/* ---- lifetime ---- */
GIS_API int gis_create(gis_handle *h, int device);
GIS_API int gis_destroy(gis_handle h);
/* Run on a caller stream (hipStream_t; NULL is the default stream) instead of
 * the handle's own non-blocking stream; gis_reset_stream goes back to it. */
GIS_API int gis_set_stream(gis_handle h, void *stream);
GIS_API int gis_reset_stream(gis_handle h);
GIS_API int gis_synchronize(gis_handle h);
GIS_API const char *gis_status_string(int status);
GIS_API const char *gis_last_error(gis_handle h);

/* ---- matrix: LDU pattern (cached per mesh) and coefficients ----
 * gis_set_pattern replaces the addressing half of getSparseCSRMat
 * (gis/ldu2csr.h:7-42): the CSR pattern and the value permutation are built
 * on the device once and reused by every gis_set_values. */
GIS_API int gis_set_pattern(gis_handle h, int n_cells, int n_faces, const int *lower_addr,
                            const int *upper_addr);
GIS_API int gis_set_pattern_device(gis_handle h, int n_cells, int n_faces,
                                   const int *d_lower_addr, const int *d_upper_addr);
/* The value half of getSparseCSRMat (gis/ldu2csr.h:23-36): one device gather. */
GIS_API int gis_set_values(gis_handle h, const double *diag, const double *lower,
                           const double *upper);
GIS_API int gis_set_values_device(gis_handle h, const double *d_diag, const double *d_lower,
                                  const double *d_upper);
GIS_API int gis_get_size(gis_handle h, int *n_rows, long long *nnz);
/* The sliced layouts of the pattern (64-row slices): what one SpMV reads.
 * Schedule 0 reads sell_slots values and int32 columns; schedule 6 reads, per
 * slice with a column-offset table, 64 values per table entry, one mask per
 * row (1 byte in byte_mask_slices, else 4) and the table itself, and the
 * sliced-row slots (value + int32 column) of the slices without a table. */
typedef struct {
    long long n_rows;
    long long nnz;
    long long slices;            /* 64-row slices */
    long long sell_slots;        /* padded slots of the sliced-row layout */
    long long table_slices;      /* slices with a column-offset table */
    long long table_entries;     /* sum of their table lengths */
    long long byte_mask_slices;  /* tabled slices with one-byte row masks (<= 8 entries) */
    long long untabled_slots;    /* sliced-row slots of the slices without a table */
} gis_layout;
GIS_API int gis_get_layout(gis_handle h, gis_layout *out);
/* Symmetric storage (no counterpart in the reference; a storage form of the
 * SpMV that getSparseCSRMat, gis/ldu2csr.h:7-42, feeds): when every 64-row slice
 * of the pattern has a column-offset table and the values are bitwise symmetric
 * (checked on the device at each gis_set_values), the FP64 offset-table
 * schedule stores only the entries on and above the diagonal and reads each
 * entry below it from its mirror; results are bit-identical.  enable 0 / 1 (default 1;
 * the GAMG levels' own matrices follow the handle).
 * gis_get_symmetric_storage: active = 1 while in use; stored_slots = value
 * slots it keeps (0 where the pattern does not allow it). */
GIS_API int gis_set_symmetric_storage(gis_handle h, int enable);
GIS_API int gis_get_symmetric_storage(gis_handle h, int *active, long long *stored_slots);
/* Copy the device CSR (row_offsets[N+1], column_indices[nnz], values[nnz]) to host. */
GIS_API int gis_get_csr(gis_handle h, int *row_offsets, int *column_indices, double *values);
/* CSR of A^T (Transpose, gis/gis.cu:4-7): same pattern for LDU, lower <-> upper. */
GIS_API int gis_get_csr_transpose(gis_handle h, int *row_offsets, int *column_indices,
                                  double *values);

/* ---- BLAS-1 / SpMV on device vectors of length N (SolverSingleGPU.h:103-211) ---- */
GIS_API int gis_spmv_device(gis_handle h, const double *d_x, double *d_y);
GIS_API int gis_spmv_transpose_device(gis_handle h, const double *d_x, double *d_y);
GIS_API int gis_dot_device(gis_handle h, const double *d_a, const double *d_b, double *out);
/* Enqueue one SpMV on the handle stream without synchronising (timing loops). */
GIS_API int gis_spmv_device_async(gis_handle h, const double *d_x, double *d_y);
GIS_API int gis_spmv32_device_async(gis_handle h, const float *d_x, float *d_y);

/* ---- launch schedule (tuning only: every schedule gives bit-identical results).
 * Replaces the launch geometry that launch.h:116-175 derives from device
 * attributes and caches in statics. ---- */
typedef struct gis_schedule {
    int spmv64;          /* FP64 SpMV: 0 = sliced rows (64-row slices, one row per lane,
                            default); 1 = CSR with products staged in LDS, 2048 nonzeros
                            per chunk; 2 = the same with 4096; 3 = CSR unstaged;
                            4 = sliced rows, two slices per lane at a time;
                            5 = sliced rows with compressed columns (an int32 base
                            per 64 slots plus one signed byte per slot; slices
                            whose offsets do not fit read the int32 columns);
                            6 = sliced rows with a column-offset table per slice
                            and a bit mask per row (no column index per slot;
                            slices whose rows do not ascend or whose offsets
                            exceed 32 entries read the int32 columns), two
                            slices per lane at a time */
    int spmv32;          /* FP32 SpMV (mixed-precision inner loop), same ids */
    int block_iters;     /* iterations enqueued between host checks of the stop flag */
    int fp64_direction;  /* CG direction p = r + beta p: 0 formed inside the SpMV kernel's
                            gathers; 1 written by a separate streaming pass first */
    int fp32_direction;  /* the same for the FP32 inner loop */
    int coarsen64;       /* tiles (1024 rows each) per workgroup of the FP64 CG
                            iteration kernels: 1, 2 or 4 */
    int coarsen32;       /* the same for the FP32 inner loop */
} gis_schedule;
/* Negative fields leave the current value unchanged.  gis_set_pattern (and
 * gis_set_interfaces) start every new pattern on a schedule chosen without
 * timing: the autotuner's cached choice for (device, N, nnz) if there is one,
 * else schedule 6 (FP64 direction form 1, FP32 form 0) when the slices
 * with a column-offset table hold at least 95% of the sliced-row slots
 * (structured meshes, batches of them), else schedule 0. */
GIS_API int gis_set_schedule(gis_handle h, const gis_schedule *s);
GIS_API int gis_get_schedule(gis_handle h, gis_schedule *s);

/* ---- autotuner: times every (SpMV schedule x direction form x tiles per
 * workgroup) candidate on the
 * handle's matrix with `iterations` solver iterations per run (<= 0: 24) and
 * applies the fastest.  A candidate is admitted only if its SpMV equals the
 * sliced-row schedule bit for bit on a fixed vector and its run ends in the
 * bit-identical solver state.  The schedule in force is kept unless a
 * candidate is more than 3% faster.  The choice is cached per (device, N, nnz):
 * later calls (any handle) apply it without timing and set *from_cache = 1,
 * and so does gis_set_pattern for a pattern of that size.
 * targets: GIS_TUNE_FP64 | GIS_TUNE_FP32.  best_ms[2] (may be NULL): time of
 * the chosen candidate's run per target (0 when cached).  Requires values. ---- */
#define GIS_TUNE_FP64 1
#define GIS_TUNE_FP32 2
GIS_API int gis_autotune(gis_handle h, int targets, int iterations, gis_schedule *chosen,
                         int *from_cache, double *best_ms);
/* Drops every cached autotuner choice (all handles). */
GIS_API int gis_autotune_clear(void);

/* ---- kernel timing with HIP events on the handle stream (opt-in; the
 * DetailType::Time analogue of support.h:7-9 at kernel granularity).
 * gis_set_timing(h, mask) records the slots whose bit is set in mask (0 stops)
 * and zeroes the counters.  Slots: 0 = the SpMV-carrying kernel of each solver
 * iteration, 1 = the vector-update kernel of each iteration, 2 = the FP64
 * update kernel of gis_solve_mixed, 3 = gis_spmv*_device_async, 4 = the
 * separate CG direction pass (direction form 1).  Each timed
 * launch adds two event records to the stream. ---- */
GIS_API int gis_set_timing(gis_handle h, int slot_mask);
GIS_API int gis_get_timing(gis_handle h, int slot, double *total_ms, long long *count);
/* Time one launch in `period` (>= 1, default 1) of each recorded slot: each
 * timed launch adds a few microseconds of launch gap, so a timed solve samples;
 * count then reports the launches actually timed. */
GIS_API int gis_set_timing_sample(gis_handle h, int period);

/* ---- NormFactor (gis/gis.cu:106-123) ---- */
GIS_API int gis_norm_factor(gis_handle h, const double *b, const double *x, double *nf);
GIS_API int gis_norm_factor_device(gis_handle h, const double *d_b, const double *d_x,
                                   double *nf);
/* gis_norm_factor_device without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under stream
 * capture, serving loops with a fixed memory budget, harnesses that forbid
 * hipMalloc).  The same kernels and the same bits as gis_norm_factor_device, but
 * the reduction workspace and the device state live in a fixed device array
 * that the library holds on each GPU until the process ends (about 4.3 MB), and
 * d_b / d_x are read in place instead of staged.  The handle's pattern and
 * values must be set (those calls allocate).  At most 33,554,432 rows
 * (GIS_ERR_SHAPE above that).  All calls on one GPU share its array: the
 * library serializes them with a per-device lock held until the result has been
 * read back, so they never overlap on the GPU; calls that did overlap there
 * would race. */
GIS_API int gis_norm_factor_device_noalloc(gis_handle h, const double *d_b, const double *d_x,
                                           double *nf);

/* ---- solve A x = b with one method; x holds x0 on entry (host pointers).
 * hist (optional, hist_cap entries) receives the residual norms
 * ||r_0||, ||r_1||, ...; *hist_len gets the full count (iterations + 1). ---- */
GIS_API int gis_solve(gis_handle h, int method, double *x, const double *b,
                      const gis_params *params, double *hist, int hist_cap, int *hist_len);
GIS_API int gis_solve_device(gis_handle h, int method, double *d_x, const double *d_b,
                             const gis_params *params, double *hist, int hist_cap,
                             int *hist_len);

/* ---- the cuFoam<Name>::solve sequence (cuFoamCG.C:45-83 and siblings):
 * nf = NormFactor(A, source, psi); solve with error = tolerance * nf and
 * max_steps = max_iter (BiCG int_bound 1e-10, Richardson omega 0 / 30 power
 * steps, MRS invM = D^-1); psi is updated in place. ---- */
GIS_API int gis_foam_solve(gis_handle h, int method, double *psi, const double *source,
                           double tolerance, long long max_iter, gis_performance *perf);

/* ---- coupled interfaces on one device (no counterpart in the reference,
 * whose wrappers ignore interfaces and coupleBouCoeffs, cuFoamCG.C:30-31).
 * n_couples extra matrix entries A(row[j], col[j]), for couplings whose two
 * sides are cells of this mesh (cyclic patches): row != col, and the set is
 * structurally symmetric (the k-th (r, c) in entry order is mirrored by the
 * k-th (c, r)).  Each row keeps its LDU entries in the reference's order and
 * gets its coupling entries appended, in entry order.  Every solver, SpMV,
 * NormFactor and A^T then includes them.  Call after gis_set_pattern, which
 * clears them; n_couples = 0 clears them too.  Values are set with
 * gis_set_coupled_values; gis_set_values keeps the last coupling values
 * (zero after gis_set_interfaces).  OpenFOAM's Amul subtracts
 * coupleBouCoeffs * psi(neighbour), so coeffs[j] = -interfaceBouCoeffs of that face.
 * gis_adjoint returns no gradient for coupling entries. ---- */
GIS_API int gis_set_interfaces(gis_handle h, int n_couples, const int *row, const int *col);
GIS_API int gis_set_coupled_values(gis_handle h, const double *diag, const double *lower,
                                   const double *upper, const double *coeffs);
GIS_API int gis_set_coupled_values_device(gis_handle h, const double *d_diag,
                                          const double *d_lower, const double *d_upper,
                                          const double *d_coeffs);

/* ---- preconditioned CG (no counterpart in the reference, whose roadmap lists
 * preconditioners as future work, README.md:71-73).  The loop, stopping rule
 * (||r||_2 <= error) and history are those of GIS_CG; z = M r, rho = (r, z),
 * p = z + (rho / rho_prev) p.  precond GIS_PRECOND_NONE gives GIS_CG bit for
 * bit; GIS_PRECOND_DIAGONAL uses M = D^-1; GIS_PRECOND_GAMG uses one
 * symmetric V-cycle of an algebraic multigrid hierarchy: two pairwise
 * agglomeration passes per level (weights |upper| + |lower|), Galerkin
 * coarse LDU matrices, nu damped-Jacobi sweeps (factor omega) before and
 * after the coarse correction (scaled by cscale), dense inverse at <= 64
 * cells.  The hierarchy is built on first use after gis_set_pattern (from
 * the values then present) and its coarse values are refreshed after every
 * gis_set_values.  For symmetric positive definite A; GAMG does not take
 * interface entries. ---- */
#define GIS_PRECOND_NONE 0
#define GIS_PRECOND_DIAGONAL 1
#define GIS_PRECOND_GAMG 2
/* The same V-cycle in FP32 (level matrices and vectors in float, the rhs
 * rounded to float, the result widened); the Krylov loop, residuals and
 * reductions stay FP64, so the attainable residual is unchanged. */
#define GIS_PRECOND_GAMG32 3
/* OpenFOAM's DIC (diagonal incomplete Cholesky) from its published algorithm:
 * rD = 1 / (diag - sum over lower faces of upper^2 rD_owner) in face order,
 * z = rD r, a forward sweep over the faces and a backward sweep in reverse face
 * order.  The sweeps run level by level (cells whose faces are independent),
 * one launch per level: about 3n launches per application on an n^3 mesh.
 * Symmetric matrices (it reads upper); no interface entries. */
#define GIS_PRECOND_DIC 4
/* OpenFOAM's DILU (diagonal incomplete LU) for asymmetric matrices, from its
 * published algorithm: rD = 1 / (diag - sum over lower faces of
 * upper lower rD_owner) in face order; z = rD r, a forward sweep over the faces
 * with lower, a backward sweep in reverse face order with upper; M^-T exchanges
 * lower and upper.  The same level-scheduled sweeps as DIC.  PBiCG only
 * (gis_solve_pbicg); gis_solve_pcg rejects it.  No interface entries. */
#define GIS_PRECOND_DILU 5
GIS_API int gis_solve_pcg(gis_handle h, int precond, double *x, const double *b,
                          const gis_params *params, double *hist, int hist_cap, int *hist_len);
GIS_API int gis_solve_pcg_device(gis_handle h, int precond, double *d_x, const double *d_b,
                                 const gis_params *params, double *hist, int hist_cap,
                                 int *hist_len);
/* OpenFOAM's PCG entry with controls, as gis_foam_solve_controls. */
GIS_API int gis_foam_solve_pcg(gis_handle h, int precond, double *psi, const double *source,
                               double tolerance, double rel_tol, long long min_iter,
                               long long max_iter, gis_performance *perf);
/* ---- preconditioned BiCG (no counterpart in the reference; its BiCG,
 * SolverSingleGPU.h:568-675, and the cuFoamBiCG wrapper take no
 * preconditioner).  The loop, breakdown test (|xi| < int_bound), stopping rule
 * (||r||_2 < error) and history are those of GIS_BICG, with z = M^-1 r and
 * z* = M^-T r* in the directions (p = z + theta p, p* = z* + theta p*) and in
 * rho = (z, r*).  precond GIS_PRECOND_NONE gives GIS_BICG bit for bit;
 * GIS_PRECOND_DIAGONAL uses M = D (OpenFOAM's diagonal preconditioner: z_i =
 * r_i / a_ii); GIS_PRECOND_DILU uses OpenFOAM's DILU; others are GIS_ERR_ARG.
 * On a symmetric matrix the iterates with the diagonal (DILU) preconditioner
 * are those of gis_solve_pcg with the diagonal (DIC) one, bit for bit. ---- */
GIS_API int gis_solve_pbicg(gis_handle h, int precond, double *x, const double *b,
                            const gis_params *params, double *hist, int hist_cap,
                            int *hist_len);
GIS_API int gis_solve_pbicg_device(gis_handle h, int precond, double *d_x, const double *d_b,
                                   const gis_params *params, double *hist, int hist_cap,
                                   int *hist_len);
/* The wrapper sequence of cuFoamBiCG (int_bound 1e-10, error = tolerance *
 * NormFactor) with a preconditioner, and relTol / minIter as
 * gis_foam_solve_controls (0 and 0: the reference's rule). */
GIS_API int gis_foam_solve_pbicg(gis_handle h, int precond, double *psi, const double *source,
                                 double tolerance, double rel_tol, long long min_iter,
                                 long long max_iter, gis_performance *perf);

/* GAMG settings (defaults omega 0.8, nu 1, cscale 1.8). */
GIS_API int gis_set_gamg(gis_handle h, double omega, int nu, double cscale);

/* ---- failure-path tests only (no counterpart in the reference): with
 * polls > 0 every workgroup that collects reduction partials on the device
 * treats them as never written and gives up after `polls` polls, so every
 * solve (and dot / NormFactor) of the handle on more than one 1024-row tile
 * stops with GIS_ERR_SYNC; polls = 0 restores normal operation.  The kernels
 * still exit normally.  Applies to solves started after the call. ---- */
GIS_API int gis_set_slot_fault(gis_handle h, int polls);
/* The hierarchy (built if needed): level count, then one level's sizes,
 * aggregation map (cell -> next-level cell; levels below the last) and LDU
 * matrix (any output pointer may be NULL). */
GIS_API int gis_gamg_levels(gis_handle h, int *n_levels);
GIS_API int gis_gamg_level(gis_handle h, int level, int *n_cells, int *n_faces, int *agg,
                           int *lower_addr, int *upper_addr, double *diag, double *lower,
                           double *upper);

/* ---- gis_foam_solve with OpenFOAM's remaining solver controls
 * (lduMatrix::solver reads tolerance, relTol, minIter and maxIter from
 * fvSolution; the reference wrappers pass on tolerance and maxIter only,
 * cuFoamCG.C:68-74).  The residual is the method's own L2 norm and
 * threshold = max(tolerance * nf, rel_tol * initial residual); iteration k
 * ends the solve when (the method's convergence test against threshold holds
 * or k >= max_iter) and k >= min_iter.  rel_tol = 0 and min_iter = 0 give
 * gis_foam_solve exactly. ---- */
GIS_API int gis_foam_solve_controls(gis_handle h, int method, double *psi,
                                    const double *source, double tolerance, double rel_tol,
                                    long long min_iter, long long max_iter,
                                    gis_performance *perf);

/* ---- several right-hand sides sharing the matrix (no counterpart in the
 * reference, which solves one field component per call; for callers with the
 * three velocity components of one momentum matrix, or batched adjoint loads):
 * one SpMM per iteration streams the matrix once for all columns.  X and B are
 * nrhs x N, each column contiguous (X[k*N + i]); X holds x0 on entry.  Column
 * k follows exactly the single-RHS method (same scalars, same stopping rule, the
 * same bits).  method: GIS_CG or GIS_BICG; transpose != 0 solves A^T X = B.
 * iterations / final_residual / status: nrhs entries each (optional).  1 <= nrhs <= 8. ---- */
GIS_API int gis_solve_multi(gis_handle h, int method, int nrhs, double *X, const double *B,
                            const gis_params *params, int transpose, int *iterations,
                            double *final_residual, int *status);
GIS_API int gis_solve_multi_device(gis_handle h, int method, int nrhs, double *d_X,
                                   const double *d_B, const gis_params *params, int transpose,
                                   int *iterations, double *final_residual, int *status);

/* ---- mixed-precision refinement of several right-hand sides (no counterpart in
 * the reference).  Per column, restarted refinement: the FP64 residual
 * R = B - A X (an FP64 SpMM), an FP32 multi-RHS CG or BiCG on A D = R from D = 0
 * (FP32 matrix, vectors and updates; FP64 reductions) until its recursive
 * residual falls below max(inner_rel_tol ||R||_2, rel_tol ||b||_2 / 2) or
 * max_inner iterations, then
 * X += D in FP64; until ||b - A x||_2 <= rel_tol ||b||_2 (FP64, recomputed) or
 * max_outer corrections (<= 0: no cap).  BiCG breakdown bound 1e-10.
 * Per column: hist[k * hist_cap + j] = ||b - A x||_2 after j corrections (may
 * be NULL), outer_steps = number of entries of hist, iterations = inner
 * iterations summed, rel_residual = final ||b - A x|| / ||b||, status GIS_OK /
 * GIS_NOT_CONVERGED / GIS_NAN. ---- */
GIS_API int gis_solve_multi_mixed(gis_handle h, int method, int nrhs, double *X,
                                  const double *B, double rel_tol, double inner_rel_tol,
                                  long long max_outer, long long max_inner, double *hist,
                                  int hist_cap, int *outer_steps, int *iterations,
                                  double *rel_residual, int *status);
GIS_API int gis_solve_multi_mixed_device(gis_handle h, int method, int nrhs, double *d_X,
                                         const double *d_B, double rel_tol,
                                         double inner_rel_tol, long long max_outer,
                                         long long max_inner, double *hist, int hist_cap,
                                         int *outer_steps, int *iterations,
                                         double *rel_residual, int *status);

/* ---- adjoint of x = A^-1 b (no counterpart in the reference; for callers that
 * differentiate through the solve): for nrhs losses with dL_k/dx = W[k],
 * solves A^T lambda_k = W[k] with the same kernels on the transposed matrix
 * and writes dL_k/db = lambda_k (lambda, nrhs x N) and dL_k/dA_ij =
 * -lambda_k[i] x[j] on the pattern, folded into LDU form: d_diag (nrhs x N),
 * d_lower, d_upper (nrhs x n_faces).  x is the forward solution.  Gradient
 * outputs may be NULL. ---- */
GIS_API int gis_adjoint(gis_handle h, int method, int nrhs, const double *W, const double *x,
                        double *lambda, double *d_diag, double *d_lower, double *d_upper,
                        const gis_params *params, int *iterations, int *status);
GIS_API int gis_adjoint_device(gis_handle h, int method, int nrhs, const double *d_W,
                               const double *d_x, double *d_lambda, double *d_diag,
                               double *d_lower, double *d_upper, const gis_params *params,
                               int *iterations, int *status);

/* ---- batched independent systems (no counterpart in the reference, which
 * solves one matrix per call; for callers with many small or uneven meshes,
 * e.g. ensembles or per-region solves): CG on every system, one launch per CG
 * stage for the whole batch, stages replayed from a captured hipGraph when
 * use_graph != 0.  Each system follows exactly the single-system CG (same
 * bits, its own iteration count and its own absolute threshold error[s]).
 * Addressing is per system (local cell ids); faces, coefficients, x and b are
 * concatenated system by system.  The batch replaces the handle's matrix. ---- */
GIS_API int gis_batch_set_pattern(gis_handle h, int n_systems, const int *n_cells,
                                  const int *n_faces, const int *lower_addr,
                                  const int *upper_addr);
GIS_API int gis_batch_set_values(gis_handle h, const double *diag, const double *lower,
                                 const double *upper);
GIS_API int gis_batch_set_values_device(gis_handle h, const double *d_diag,
                                        const double *d_lower, const double *d_upper);
GIS_API int gis_batch_solve(gis_handle h, double *x, const double *b, const double *error,
                            long long max_steps, int use_graph, int *iterations,
                            double *final_residual, int *status);
GIS_API int gis_batch_solve_device(gis_handle h, double *d_x, const double *d_b,
                                   const double *error, long long max_steps, int use_graph,
                                   int *iterations, double *final_residual, int *status);

/* ---- mixed precision (no counterpart in the reference; for callers that want
 * an FP64-accurate CG solve at FP32 bandwidth): the Krylov inner loop runs in
 * FP32, FP64 is kept for the residual recompute r = b - A x, the correction
 * accumulation x += d and the norms, until ||b - A x||_2 / ||b||_2 <= rel_tol.
 * mode 0: reliable updates (search direction kept across FP64 updates);
 * mode 1: restarted iterative refinement (inner CG to inner_rel_tol, restart).
 * An FP64 update is triggered when the FP32 residual falls below
 * inner_rel_tol times the last true residual.  max_iter caps the inner
 * iterations (<= 0: unbounded).  hist gets the FP64 true residual norm at x0
 * and after every update.  Returns GIS_OK when converged, GIS_NOT_CONVERGED
 * or GIS_NAN otherwise (x is valid in every case). ---- */
GIS_API int gis_solve_mixed(gis_handle h, double *x, const double *b, double rel_tol,
                            double inner_rel_tol, long long max_iter, int mode, double *hist,
                            int hist_cap, int *hist_len, int *iterations, double *rel_residual);
GIS_API int gis_solve_mixed_device(gis_handle h, double *d_x, const double *d_b,
                                   double rel_tol, double inner_rel_tol, long long max_iter,
                                   int mode, double *hist, int hist_cap, int *hist_len,
                                   int *iterations, double *rel_residual);

#ifdef __cplusplus
}
#endif

#endif /* GIS_AMD_GIS_H */
