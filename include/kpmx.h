/*
 * kpmx -- C-ABI of the MI355X-native KPM Chebyshev-moment engine (libkpmx.so).
 *
 * This is the device boundary of the engine.  It replaces the reference's compute plugin
 *   cpb::kpm::Compute::Interface::moments(MomentsRef, Starter const&, AlgorithmConfig const&,
 *                                         OptimizedHamiltonian const&) const
 *   (cppcore/include/kpm/Core.hpp:21-38, implemented on the CPU by DefaultCompute,
 *    cppcore/include/kpm/default/Compute.hpp:9-27, cppcore/src/kpm/default/Compute.cpp:128-131)
 * with one entry point per `MomentsRef` alternative (cppcore/include/kpm/Moments.hpp:125-126).
 * The host C++ compute class (HipCompute) and the Python facade sit above it.
 *
 * Conventions
 *   - Every function returns KPMX_OK (0) or a nonzero KPMX_E* code; kpmx_last_error() gives the
 *     message of the last failure on the calling thread.  The host layer turns it into the
 *     reference's exception (std::runtime_error / RuntimeError).
 *   - Plain pointers and sizes only.  "host" buffers are caller-owned and read or written only
 *     during the call.  Device buffers are owned by the library (kpmx_mat, kpmx_block, kpmx_ctx).
 *   - Scalars: KPMX_F32 float, KPMX_F64 double, KPMX_C64 / KPMX_C128 complex stored as {re, im}
 *     pairs (the layout of std::complex<float/double>).
 *   - Matrices are given as CSR (int32 row pointers and column indices, row-major).  The matrix is
 *     the already scaled H2 = (H - b) * 2/a of the reference's OptimizedHamiltonian
 *     (cppcore/src/kpm/OptimizedHamiltonian.cpp:54-135), uploaded once.
 *   - Starter blocks are given vector-major: num_vectors x n, each vector contiguous, already in
 *     the matrix's (optimized) order.
 *   - Moments follow the reference collectors: mu_0 is stored halved
 *     (cppcore/src/kpm/default/collectors.cpp:5-106).
 *   - One HIP stream per context; a context is used from one thread at a time (the reference's
 *     KPM object has the same rule).
 */
#ifndef KPMX_H
#define KPMX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KPMX_OK 0
#define KPMX_EINVAL 1   /* invalid argument (size mismatch, bad scalar code, ...) */
#define KPMX_EHIP 2     /* HIP runtime error */
#define KPMX_ENOMEM 3   /* device allocation failed */

#define KPMX_F32 0
#define KPMX_F64 1
#define KPMX_C64 2
#define KPMX_C128 3

#define KPMX_FMT_CSR 0  /* CSR, one wavefront per row group                                  */
#define KPMX_FMT_SELL 1 /* sliced ELL with 64-row slices (single-vector recurrences)          */

/* Graph policy of a context for the diagonal recurrence (kpmx_ctx_set_graph_policy). */
#define KPMX_GRAPH_NEVER 0  /* launch every step from the host (default)                           */
#define KPMX_GRAPH_AUTO 1   /* starter blocks <= 64 MiB: capture on the second identical run          */
#define KPMX_GRAPH_ALWAYS 2 /* capture every recurrence                                             */
                        /* replaces kpm::MatrixFormat {CSR, ELL} (cppcore/include/kpm/Config.hpp:7) */

#define KPMX_KUBO_FP64 0        /* mu_nm contraction in FP64 (reference arithmetic)             */
#define KPMX_KUBO_FP32_SPLIT 1  /* FP32 MFMA with hi/lo split-precision compensation            */
#define KPMX_KUBO_BF16X3 2      /* bf16 MFMA on three-piece operand splits (fp32-accurate)      */

typedef struct kpmx_ctx kpmx_ctx;
typedef struct kpmx_mat kpmx_mat;
typedef struct kpmx_block kpmx_block;
typedef struct kpmx_batch kpmx_batch;
typedef struct kpmx_tape kpmx_tape;

/* Message of the last failed call on this thread ("" if none). */
const char* kpmx_last_error(void);
/* ABI version (major * 100 + minor). */
int kpmx_version(void);
/* Number of visible HIP devices. */
int kpmx_device_count(int* count);

/* A device context: one GPU, one stream, reusable workspaces.
 * Replaces the reference's DefaultCompute object (thread pool + SIMD batching),
 * cppcore/src/kpm/default/Compute.cpp:121-126.
 * Matrices, starter blocks, tapes and batches belong to the context they were made on: destroy
 * them before destroying the context. */
int kpmx_ctx_create(int device, kpmx_ctx** out);
int kpmx_ctx_destroy(kpmx_ctx* ctx);
int kpmx_ctx_synchronize(kpmx_ctx* ctx);
/* HIP-event timer on the context's stream (no counterpart in the reference, which times on the
 * host with Chrono, cppcore/include/utils/Chrono.hpp:32-71): start records an event, stop records
 * a second one, waits for it and returns the elapsed milliseconds. */
int kpmx_ctx_timer_start(kpmx_ctx* ctx);
int kpmx_ctx_timer_stop(kpmx_ctx* ctx, double* ms);

/* Upload a CSR matrix once (the scaled H2, or an operator such as the velocity).
 * format: KPMX_FMT_CSR or KPMX_FMT_SELL (SELL is built on the device side from the CSR).
 * Replaces OptimizedHamiltonian::matrix() as seen by the compute backend
 * (cppcore/include/kpm/OptimizedHamiltonian.hpp:117,157; ELL conversion
 *  cppcore/include/numeric/ellmatrix.hpp:64-82). */
int kpmx_mat_upload(kpmx_ctx* ctx, int scalar, int format, int64_t rows, int64_t nnz, const int32_t* indptr,
                    const int32_t* indices, const void* values, kpmx_mat** out);
/* Bytes of the matrix's value-table form (the form the fp64 column-pair step variants 45..59
 * read: slice pointers and widths, 4-byte indices and 1-byte codes of the off-diagonal entries,
 * the dense diagonal and the value table), or 0 when the matrix has none (no counterpart in the
 * reference). */
int kpmx_mat_coded_bytes(const kpmx_mat* m, int64_t* bytes);
int kpmx_mat_destroy(kpmx_mat* m);

/* Light cone of the next recurrences on this matrix (the reference's `optimal_size` algorithm,
 * cppcore/include/kpm/calc_moments.hpp:35-51,102-114 with SliceMap::optimal_size,
 * cppcore/include/kpm/OptimizedHamiltonian.hpp:54-96): step k >= 1 of a recurrence computes only
 * rows [0, rows[k]) of the BFS-ordered matrix, the rows that can be nonzero (or still reach the
 * collected rows); rows[0] is ignored (the first step always covers the whole matrix) and steps
 * past num_steps cover it too.  num_steps = 0 clears the cone.  The caller computes rows[] from
 * its slice map; for the diagonal recurrence step k is reference iteration k + 1. */
int kpmx_mat_set_step_rows(kpmx_mat* h2, int64_t num_steps, const int64_t* rows);

/* Upload starter vectors (host, vector-major num_vectors x n) and keep them resident.
 * Replaces make_r0 for vector and block starters (cppcore/include/kpm/Starter.hpp:36-50). */
int kpmx_block_upload(kpmx_ctx* ctx, int scalar, int64_t n, int64_t num_vectors, const void* r0, kpmx_block** out);
int kpmx_block_destroy(kpmx_block* b);

/* DiagonalMoments / BatchDiagonalMoments: mu_n = <r|T_n(H~)|r> for every vector of the block,
 * two moments per Chebyshev step (cppcore/include/kpm/calc_moments.hpp:35-89,
 * cppcore/src/kpm/default/collectors.cpp:5-34).  num_moments must be even.
 * mu: host, num_vectors x num_moments, scalar type of h2. */
int kpmx_moments_diagonal(const kpmx_mat* h2, const kpmx_block* r0, int64_t num_moments, void* mu);

/* Autotuner (no counterpart in the reference, whose CPU kernels have one fixed shape): time the
 * variants of the fused diagonal step kernel (three tile shapes, each on the full resident grid,
 * half of it, a quarter, twice or four times it, each with the tile's matrix entries staged
 * through workgroup LDS or handed out by lane shuffles within each wave) on this matrix and
 * starter block and keep the fastest for every later recurrence of that scalar type and vector
 * width on the context, timed on the first column tile the recurrence runs (see
 * kpmx_ctx_set_tile_width).  variant (optional): the chosen variant index, tile shape + 3 * grid
 * (0 resident, 1 half, 2 quarter, 3 twice, 4 four times) + 15 * staging (0 LDS, 1 shuffles,
 * 2 column pairs: an fp64 block of 8 or more vectors read as pairs of adjacent columns, 16 bytes
 * per lane, for other blocks run as 1; 3 column pairs on the matrix's value-table form: an SELL
 * of the off-diagonal entries with 8-bit codes into a table of at most 256 values and the
 * diagonal as a dense vector, built at upload when the off-diagonal entries take that few
 * distinct values, for other matrices run as 2).  On fp64 SELL matrices it then also times the
 * LDS two-step form (see kpmx_ctx_set_autotune_two_step) and keeps it, with its column-tile
 * width, when it is faster per vector. */
int kpmx_autotune_diagonal(const kpmx_mat* h2, const kpmx_block* r0, int* variant);
/* Pin the diagonal step variant (as numbered above) for one scalar type and vector width on the
 * context, e.g. a choice an earlier autotuning run reported (no counterpart in the reference). */
int kpmx_ctx_set_step_variant(kpmx_ctx* ctx, int scalar, int vector_width, int variant);

/* Pin the variant of the adjoint step kernel (numbered as the diagonal step variants; no
 * counterpart in the reference) on the context; -1 restores the default (wave-staged entries on
 * twice the resident grid, fp64 blocks of 8 or more vectors as column pairs). */
int kpmx_ctx_set_adjoint_variant(kpmx_ctx* ctx, int variant);

/* Widest column tile of the diagonal recurrence for one scalar type (1, 2, 4, ..., 64; no
 * counterpart in the reference).  A starter tile wider than this runs as several recurrences over
 * sub-tiles of this many columns (each copied out of the tile first); narrower tiles move fewer
 * bytes per gathered row, so one XCD's L2 holds the rows of a wider band, at the price of one
 * matrix read per sub-tile and step.  The moments are the same up to rounding.  Defaults: 16 for
 * fp64, 64 for the others.  get: the current width. */
int kpmx_ctx_set_tile_width(kpmx_ctx* ctx, int scalar, int width);
int kpmx_ctx_get_tile_width(kpmx_ctx* ctx, int scalar, int* width);

/* Two Chebyshev steps per launch for the diagonal recurrence of one scalar type (the GPU form of
 * the reference's interleaved sweep, cppcore/include/kpm/calc_moments.hpp:61-89): tile_rows > 0
 * (a multiple of 64) cuts the rows into tiles of that many; each workgroup computes the next
 * vector on its tile and on the tile's halo (the outside rows its rows reference), then the one
 * after on its tile, so a launch reads two vectors and writes two instead of three per step.
 * Runs on SELL matrices for column tiles whose rows are whole 128-byte lines (fp64 >= 16
 * vectors, complex128 >= 8, fp32 >= 32, complex64 >= 16) and without light-cone row limits;
 * anything else keeps single steps.  0 (the default) turns it off.  workgroups_per_cu > 0 runs
 * a persistent grid of that many workgroups per CU walking the tiles (0: one workgroup per tile).
 * The moments are the same up to rounding.  get: the current tile rows. */
int kpmx_ctx_set_two_step(kpmx_ctx* ctx, int scalar, int tile_rows, int workgroups_per_cu);
int kpmx_ctx_get_two_step(kpmx_ctx* ctx, int scalar, int* tile_rows);
/* The LDS form of the two-step recurrence (fp64 blocks of 8 or 16 vectors, complex128 blocks of 4
 * or 8; no counterpart in the reference): tiles of 256 rows, one 64-row slice per wave; each workgroup keeps r_k+1 of its tile
 * and of the tile's halo rows in LDS and the wave keeps r_k of its slice in registers, so a launch
 * reads two vectors and writes two; it runs on a persistent grid of workgroups_per_cu workgroups
 * per CU (0: as many as fit, at most 4), the kernel's registers bounded so that many fit.  Reads the value-table form where the single step would.  Other
 * scalar types and widths, and matrices whose tiles have too many halo rows for 64 KB of LDS
 * ((256 + halo rows) x 8 bytes x vectors), keep single steps.  kpmx_ctx_set_two_step turns it off
 * again. */
int kpmx_ctx_set_two_step_lds(kpmx_ctx* ctx, int scalar, int workgroups_per_cu);
/* enabled: whether the LDS two-step form is on for the scalar type; workgroups_per_cu: its grid
 * (0: as many as fit). */
int kpmx_ctx_get_two_step_lds(kpmx_ctx* ctx, int scalar, int* enabled, int* workgroups_per_cu);
/* Whether kpmx_autotune_diagonal may also choose the LDS two-step form (default 1): for fp64
 * and complex128 SELL matrices it then times column tiles of 64-byte rows (8 fp64 or 4 complex128
 * vectors) with 4 workgroups per CU and of 128-byte rows with 3 against the single steps per vector and keeps the two-step form, with its tile width,
 * when it is at least 2 % faster.  Settings chosen this way are reset by the next autotuning;
 * settings made with kpmx_ctx_set_tile_width / kpmx_ctx_set_two_step / kpmx_ctx_set_two_step_lds
 * are the caller's and are only changed when the autotuner picks the two-step form. */
int kpmx_ctx_set_autotune_two_step(kpmx_ctx* ctx, int enable);

/* Graph policy (no counterpart in the reference, whose CPU loop has no launch cost): with
 * KPMX_GRAPH_ALWAYS the diagonal recurrence of a starter block is captured into a HIP graph on its
 * first run (KPMX_GRAPH_AUTO: on its second run) and replayed by later runs with the same matrix,
 * number of moments, tuned variant and workspaces; the results are the same as with
 * KPMX_GRAPH_NEVER.  Capture and instantiation cost host time on the run that captures; on the
 * engine's kernels replay has not been faster than per-step launches, so the default is
 * KPMX_GRAPH_NEVER and the graph path is for callers whose host thread is busy. */
int kpmx_ctx_set_graph_policy(kpmx_ctx* ctx, int policy);

/* Same recurrence, results kept as raw per-step sums on the device; for timing loops.
 * Leaves the sums readable through kpmx_moments_diagonal_fetch. */
int kpmx_moments_diagonal_run(const kpmx_mat* h2, const kpmx_block* r0, int64_t num_moments);
int kpmx_moments_diagonal_fetch(const kpmx_mat* h2, const kpmx_block* r0, int64_t num_moments, void* mu);

/* kpmx_moments_diagonal for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  Same inputs, output and kernels as kpmx_moments_diagonal, but the recurrence's work
 * vectors, partial and raw sums and the collected moments live in a static device array of
 * 64 MiB per GPU (held by the process until it ends): 2 x n x (vector width) scalars (3 x when
 * the block is run as column sub-tiles, see kpmx_ctx_set_tile_width), plus (num_moments / 2) x
 * 4 x (vector width) doubles of sums, plus num_vectors x num_moments scalars, plus one step's
 * partial sums, must fit, otherwise KPMX_EINVAL with the sizes in the message.  No device memory
 * is allocated or freed (h2 and r0 are uploaded beforehand).  It always runs single steps (the
 * two-step forms need workspaces of their own), so where kpmx_moments_diagonal runs two steps per
 * launch the two results agree to rounding rather than bit for bit.  Calls that overlap on one
 * GPU share the array and race. */
int kpmx_moments_diagonal_noalloc(const kpmx_mat* h2, const kpmx_block* r0, int64_t num_moments, void* mu);

/* GenericMoments: mu_n = <beta| op T_n(H~) |r0> for the first vector of r0
 * (cppcore/src/kpm/default/collectors.cpp:36-58).  op may be null (op_nnz = 0): identity.
 * beta: host, n scalars.  op: host CSR (n x n).  mu: host, num_moments scalars. */
int kpmx_moments_generic(const kpmx_mat* h2, const kpmx_block* r0, int64_t num_moments, const void* beta,
                         int64_t op_nnz, const int32_t* op_indptr, const int32_t* op_indices, const void* op_values,
                         void* mu);

/* MultiUnitMoments: mu[i][n] = (T_n(H~) r0)[dest[i]] (cppcore/src/kpm/default/collectors.cpp:60-75).
 * mu: host, ndest x num_moments. */
int kpmx_moments_multi_unit(const kpmx_mat* h2, const kpmx_block* r0, int64_t num_moments, int64_t ndest,
                            const int32_t* dest, void* mu);

/* DenseMatrixMoments: out[c][n][:] = op T_n(H~) r0_c (op may be null) for every vector c of r0
 * (cppcore/src/kpm/default/collectors.cpp:77-106).  out: host, num_vectors x num_moments x n. */
int kpmx_moments_dense(const kpmx_mat* h2, const kpmx_block* r0, int64_t num_moments, const kpmx_mat* op, void* out);

/* Kubo-Bastin 2-D moments (cppcore/src/kpm/Core.cpp:119-150 with MomentMultiplication,
 * cppcore/src/kpm/Moments.cpp:92-101):
 *   mu_nm = (1/R) sum_r sum_i L_r[n,i] conj(R_r[m,i]),
 *   L_r rows = T_n(H~) (vl r),  R_r rows = vr T_m(H~) r,  r = the R vectors of `r0`.
 * mode: KPMX_KUBO_FP64, KPMX_KUBO_FP32_SPLIT or KPMX_KUBO_BF16X3.  mu_nm: host, M x M, scalar type of h2. */
int kpmx_kubo_moments(const kpmx_mat* h2, const kpmx_mat* vl, const kpmx_mat* vr, const kpmx_block* r0,
                      int64_t num_moments, int mode, void* mu_nm);

/* Adjoint gradient of the diagonal moments (no counterpart in the reference).
 * For L = sum_n w[n] mu_n with mu_n the moments of kpmx_moments_diagonal averaged over the vectors
 * of r0 (mu_0 halved), returns dL/dH2 for every stored entry of h2 (grad: host, nnz doubles, the
 * CSR order of the upload), by the reverse (adjoint) Chebyshev recurrence
 *   lambda_j = g_j + H2 lambda_{j+1} - lambda_{j+2},   dL/dH2 += lambda_{j+1} v_j^T,
 * fused with the gradient accumulation over the sparsity pattern.  checkpoint: steps between stored
 * vector pairs, the vectors in between recomputed by running the recurrence backwards
 * (v_{j-1} = H2 v_j - v_{j+1}) inside the adjoint steps, restarting from every stored pair; <= 0
 * keeps every vector (no recomputation).
 * Real H2 (KPMX_F32 / KPMX_F64) uploaded with KPMX_FMT_SELL; at most 64 vectors.
 * mu (optional): host, num_vectors x num_moments, the forward moments. */
int kpmx_adjoint_diagonal(const kpmx_mat* h2, const kpmx_block* r0, int64_t num_moments, const double* w,
                          int64_t checkpoint, double* grad, void* mu);
/* The same in two halves, for callers (autograd) that learn w after the forward pass: the forward
 * pass returns the moments and a tape holding the stored vectors on the device; the backward pass
 * takes w and returns the gradient; the tape can be run backward more than once. */
int kpmx_adjoint_forward(const kpmx_mat* h2, const kpmx_block* r0, int64_t num_moments, int64_t checkpoint, void* mu,
                         kpmx_tape** out);
int kpmx_adjoint_backward(kpmx_tape* tape, const double* w, double* grad);
/* Frees a tape; the vector storage of the largest finished tape stays with the context (until it
 * is destroyed) and is reused by the next tape that fits, so repeated gradients do not allocate. */
int kpmx_tape_destroy(kpmx_tape* tape);

/* Batched independent systems -- many disorder realizations of uneven size advanced together.
 * No single counterpart in the reference: it is what a caller of the reference does by running
 * one `KPM.calc_dos` (cppcore/src/kpm/Core.cpp:74-90) per realization, each with its own bounds
 * and its own default-seeded random starter; here every realization's diagonal moments come from
 * one block-diagonal recurrence.
 *   upload:   the block-diagonal matrix of the already scaled H2 of every system (CSR, global
 *             indices; system s owns rows [row_offsets[s], row_offsets[s+1]) and may only couple
 *             to them).
 *   starters: num_vectors (1..64) starters per system, host, vector-major num_vectors x N.
 *   run:      all num_moments / 2 steps.  use_graph (the run mode): 0 one launch per step (and
 *             per-system reduction), 1 the whole step sequence captured in one HIP graph on the
 *             first run for this num_moments and replayed afterwards, 2 one persistent kernel
 *             launch for all steps (grid barriers between steps; the partition's workgroups must
 *             fit the resident grid, else KPMX_EINVAL; returns after the run has finished).
 *   fetch:    mu: host, num_systems x num_vectors x num_moments (mu_0 halved). */
int kpmx_batch_upload(kpmx_ctx* ctx, int scalar, int64_t num_systems, const int64_t* row_offsets, int64_t nnz,
                      const int32_t* indptr, const int32_t* indices, const void* values, kpmx_batch** out);
int kpmx_batch_starters(kpmx_batch* b, int64_t num_vectors, const void* r0);
int kpmx_batch_run(kpmx_batch* b, int64_t num_moments, int use_graph);
/* Partition of the batched step's 64-row slices over the walkers of num_workgroups workgroups
 * (four per workgroup, one per wave; no counterpart in the reference).  walk_chunk == 0:
 * contiguous ranges.  walk_chunk > 0 (num_workgroups a multiple of 8): XCD-aware walk -- 8
 * contiguous regions, one per XCD, whose runs of walk_chunk slices are dealt round-robin to that
 * XCD's walkers, so the gathered rows of the band they sweep together stay in the XCD's L2.
 * Results stay within rounding of any other partition (only the grouping of each system's sums
 * changes). */
int kpmx_batch_set_partition(kpmx_batch* b, int num_workgroups, int walk_chunk);
/* Matrix form of the batched step (no counterpart in the reference).  The value-table form --
 * off-diagonal entries as a 4-byte column and a 1-byte code into a table of at most 256 values
 * per system, the diagonal as a dense vector over the rows -- is built at upload when every
 * system's off-diagonal values fit one table and no row holds two diagonal entries, and is then
 * the default.  enable = 0 selects the standard form (4-byte column + value per stored entry),
 * 1 the value-table form (KPMX_EINVAL when the batch has none), under every run mode.  Results
 * agree to rounding (the diagonal term is added last). */
int kpmx_batch_set_value_table(kpmx_batch* b, int enable);
/* Column tiles of the batch's vectors (no counterpart in the reference): the next
 * kpmx_batch_starters splits its num_vectors columns into tiles of at most `width` vectors
 * (0: the library's default for the scalar type), each run as its own recurrence over the whole
 * batch (the matrix is read once per tile, the walker's registers hold one tile's columns).
 * kpmx_batch_tiles reports the current split: num_tiles tiles of tile_vectors vectors (the last
 * may hold fewer). */
int kpmx_batch_set_tile_width(kpmx_batch* b, int width);
int kpmx_batch_tiles(const kpmx_batch* b, int64_t* num_tiles, int64_t* tile_vectors);
/* available / enabled (optional): whether the batch has / uses the value-table form;
 * matrix_bytes (optional): bytes of the selected form that one step reads -- every stored slot
 * (padding included) with its column and value or code, plus the diagonal and the tables. */
int kpmx_batch_format(const kpmx_batch* b, int* available, int* enabled, int64_t* matrix_bytes);
/* Autotuner of the batched step (no counterpart in the reference): times the step on this batch
 * and its starters over contiguous ranges for 1 to 256 workgroups per CU (four walkers, one per
 * wave, each) and over the XCD-aware walk (runs of 1, 2 or 4 slices on half, one, two and four
 * times the resident grid), and keeps the fastest partition.  num_workgroups, walk_chunk (optional): the chosen
 * partition. */
int kpmx_batch_autotune(kpmx_batch* b, int* num_workgroups, int* walk_chunk);
int kpmx_batch_fetch(kpmx_batch* b, int64_t num_moments, void* mu);
int kpmx_batch_destroy(kpmx_batch* b);

/* Extreme eigenvalues of the (unscaled) Hamiltonian `h` by the Lanczos procedure, converged to
 * `precision_percent` (cppcore/include/compute/lanczos.hpp:93-154, used by Bounds::compute_bounds,
 * cppcore/src/kpm/Bounds.cpp:21-31).  The matrix-vector products and reductions run on the GPU,
 * the small tridiagonal eigenproblem on the host. */
int kpmx_lanczos_bounds(const kpmx_mat* h, double precision_percent, double* min_energy, double* max_energy,
                        int* loops);

/* Reconstruction sums (cppcore/include/kpm/reconstruct.hpp:16-143).  The caller forms the
 * per-energy quantities in the moments' real type as the reference does (E~ = (E - b) / a,
 * theta = acos E~ and the prefactor); the device evaluates the M-term sums of every energy and
 * column.  real: KPMX_F32 or KPMX_F64, the type the sums run in.  Host buffers.
 *
 * SpectralDensity (reconstruct.hpp:16-54):
 *   out[e * cols + c] = pref[e] * sum_n mu[c * M + n] cos(n theta[e])
 * mu: cols x M reals; theta, pref: num_energy reals; out: num_energy x cols doubles. */
int kpmx_chebyshev_cos_series(kpmx_ctx* ctx, int real, int64_t num_moments, int64_t cols, const void* mu,
                              int64_t num_energy, const void* theta, const void* pref, double* out);
/* GreensFunction (reconstruct.hpp:59-85):
 *   out[e * cols + c] = pref[e] * sum_n mu[c * M + n] exp(-i n theta[e])
 * mu: cols x M complex ({re, im} reals); pref: num_energy complex; out: num_energy x cols complex
 * doubles. */
int kpmx_chebyshev_exp_series(kpmx_ctx* ctx, int real, int64_t num_moments, int64_t cols, const void* mu,
                              int64_t num_energy, const void* theta, const void* pref, double* out);
/* KuboBastin energy-sample sums (reconstruct.hpp:90-143) in double precision:
 *   out[q] = sum_nm mu[n * M + m] Gamma_nm(E_q) / (1 - E_q^2)^2,
 *   Gamma_nm = g(n, m) + conj(g(m, n)),  g(n, m) = (E - i m sqrt(1 - E^2)) e^{i m acos E} T_n(E),
 * evaluated as sum_n T_n(E) (mu a)_n + conj(a_n) (mu t)_n with a_m = (E - i m sqrt(1 - E^2))
 * e^{i m acos E} and t_m = T_m(E).  mu_nm: M x M complex doubles (damped moments); samples: the
 * scaled energies E_q; out: num_samples complex doubles.  The Fermi-Dirac integration over the
 * samples stays with the caller. */
int kpmx_kubo_bastin_sums(kpmx_ctx* ctx, int64_t num_moments, const double* mu_nm, int64_t num_samples,
                          const double* samples, double* out);

/* Streaming probe (no counterpart in the reference; the practical bandwidth reference for the
 * roofline reports): best GB/s over a few grid sizes of a non-temporal 16-byte copy of `bytes`
 * (bytes read + bytes written), of a 16-byte read-only sweep, and (triad_gbps, optional) of
 * b = a + b in place with b non-temporal -- the step kernel's two reads and one write per element
 * with the gathers made contiguous (3 x bytes moved) -- on buffers the call allocates and frees. */
int kpmx_stream_probe(kpmx_ctx* ctx, int64_t bytes, double* copy_gbps, double* read_gbps, double* triad_gbps);

#ifdef __cplusplus
}
#endif

#endif /* KPMX_H */
