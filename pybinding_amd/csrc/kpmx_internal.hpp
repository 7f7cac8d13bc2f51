// Internal launch interface between the C-ABI layer (kpmx_abi.hip) and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kpmx {

enum { MODE_DIAG = 0, MODE_DOT = 1, MODE_NONE = 2 };
enum { FORMAT_CSR = 0, FORMAT_SELL = 1 };

struct StepLaunch {
    int scalar = 1;          // 0 f32, 1 f64, 2 c64, 3 c128
    int format = FORMAT_CSR;
    int vw = 1;              // vectors per row (power of two <= 64)
    int mode = MODE_DIAG;
    int num_wg = 0;          // grid size; 0 = resident grid chosen by the launcher (SELL block kernel)
    int num_cu = 256;
    const int32_t* indptr = nullptr;   // CSR row pointers (CSR format)
    const int32_t* indices = nullptr;  // CSR or SELL column indices
    const void* vals = nullptr;
    const int64_t* slice_ptr = nullptr;  // SELL
    const int32_t* slice_width = nullptr;
    int max_width = 0;                   // widest SELL slice
    const void* x = nullptr;
    const void* yin = nullptr;
    void* yout = nullptr;
    const void* w = nullptr;
    void* copy_out = nullptr;
    int64_t copy_col_stride = 0;
    uint16_t* copy_split = nullptr;   // bf16 pieces of y_new (see store_split), plane stride copy_col_stride
    const int32_t* mark = nullptr;
    void* mark_out = nullptr;
    int64_t mark_stride = 0;
    double* partials = nullptr;
    int64_t n = 0;
    double hscale = 1.0;
    int sub_y = 1;
    bool query_grid = false;  // launch_step returns the resident grid instead of launching
    int tune = 0;             // tile-shape variant of the SELL block kernel (0 default, 1 smaller, 2 larger)
    // value-table form of the SELL matrix (off-diagonal entries as 8-bit codes into ctable, the
    // diagonal as a dense vector); null when the matrix has none
    const int64_t* cslice_ptr = nullptr;
    const int32_t* cslice_width = nullptr;
    const int32_t* cindices = nullptr;
    const uint8_t* ccodes = nullptr;
    const void* cdiag = nullptr;
    const void* ctable = nullptr;
    int cmax_width = 0;
    int ctable_size = 0;
};

// Slice sequence of walker b of nb of the batched step (the device kernel walks it and the host
// replays it to number the partial-sum slots).
//   chunk == 0: the contiguous slices [tile_begin[b], tile_begin[b+1]).
//   chunk > 0 (nb a multiple of 8): the slices are cut into 8 contiguous regions, region r
//     belonging to the walkers with b % 8 == r (on the XCD r of their workgroup), and inside a
//     region runs of `chunk` consecutive slices are dealt round-robin to its walkers, so the
//     resident walkers of one XCD sweep one narrow band of rows together.
// Slices only increase along a walk, so a walker meets every system in one stretch.
struct BatchWalk {
    int64_t s = 0, end = 0, jump = 0;
    int64_t left = 0, chunk = 0;
    __host__ __device__ BatchWalk(int64_t nslices, int nb, int b, int chunk_, const int64_t* tile_begin) {
        if (chunk_ <= 0) {
            s = tile_begin[b];
            end = tile_begin[b + 1];
            chunk = end - s + 1;  // never runs out
            jump = 0;
        } else {
            const int region = b % 8, local = b / 8, per_region_wg = nb / 8;
            const int64_t per_region = (nslices + 7) / 8;
            const int64_t start = region * per_region;
            end = start + per_region < nslices ? start + per_region : nslices;
            chunk = chunk_;
            s = start + (int64_t)local * chunk;
            jump = (int64_t)(per_region_wg - 1) * chunk;
        }
        left = chunk;
    }
    __host__ __device__ bool valid() const { return s < end; }
    __host__ __device__ void next() {
        ++s;
        if (--left == 0) {
            s += jump;
            left = chunk;
        }
    }
};

struct BatchedLaunch {
    int scalar = 1, vw = 4, num_wg = 0, max_width = 0;
    int chunk = 0;                          // walk of the slices (see BatchWalk)
    int64_t nslices = 0;
    const int32_t* indices = nullptr;
    const void* vals = nullptr;
    const int64_t* slice_ptr = nullptr;
    const int32_t* slice_width = nullptr;
    const int32_t* slice_sys = nullptr;
    const int64_t* tile_begin = nullptr;  // walkers + 1 (chunk == 0)
    const int64_t* slot_begin = nullptr;  // walkers: first stretch of every walker
    const int64_t* slot_index = nullptr;  // partial slot of every stretch
    const void* x = nullptr;
    const void* yin = nullptr;
    void* yout = nullptr;
    double* partials = nullptr;
    int64_t n = 0;
    double hscale = 1.0;
    int sub_y = 1;
    // the value-table form (codes != nullptr): indices / slice_ptr / slice_width / max_width
    // describe the off-diagonal entries only, each a 1-byte code into its system's table
    // (table + tab_off[system], at most 256 values of the value type), and the diagonal is the
    // dense vector diag over the padded rows
    const uint8_t* codes = nullptr;
    const void* diag = nullptr;
    const void* table = nullptr;
    const int32_t* tab_off = nullptr;
    // the persistent form (launch_batched_persistent): the whole recurrence in one launch
    const void* r0 = nullptr;
    void* w0 = nullptr;
    void* w1 = nullptr;
    double* sums = nullptr;             // steps x num_sys x width
    const int64_t* slot_first = nullptr;
    int64_t num_slots = 0, num_sys = 0, steps = 0;
    int width = 0;                      // 4 vw
    unsigned* barrier = nullptr;        // 4 words of device memory
};
// Two Chebyshev steps per launch over tiles of `tile_rows` rows (k_step2_sell): phase A writes
// r_{k+1} = s H2 x - xprev to out1 on every tile's rows and halo rows (h1_rows[h1_ptr[j] ..]),
// phase B writes r_{k+2} = H2 out1 - x to out2 on the tile's rows; the diagonal sums of the two
// steps go to partialsA / partialsB (step2_grid slots each, step2_partial_width doubles per slot,
// fp64 in the column-pair layout).
struct Step2Launch {
    int scalar = 1, vw = 16, max_width = 0;
    const int32_t* indices = nullptr;
    const void* vals = nullptr;
    const int64_t* slice_ptr = nullptr;
    const int32_t* slice_width = nullptr;
    const int64_t* h1_ptr = nullptr;
    const int32_t* h1_rows = nullptr;
    const void* xprev = nullptr;
    const void* x = nullptr;
    void* out1 = nullptr;
    void* out2 = nullptr;
    double* partialsA = nullptr;
    double* partialsB = nullptr;
    int64_t n = 0, ntiles = 0, tile_rows = 0;
    int num_wg = 0;  // persistent grid (a multiple of 8); 0: one workgroup per tile
    double hscaleA = 1.0;
    // The LDS form (k_step2_lds, lidx != nullptr; fp64 column pairs or complex128): tiles of STEP2_LDS_TILE
    // rows, one 64-row slice per wave; phase A keeps r_{k+1} of the tile's rows and of its at
    // most hmax halo rows in LDS and phase B gathers it there by lidx (every stored entry's column
    // as its position in the tile's rows followed by its halo rows); the wave keeps r_k of its
    // slice in registers across the barrier, so a launch reads r_{k-1}, r_k and writes r_{k+1},
    // r_{k+2} once each.  codes != nullptr: the value-table form (indices, slice_ptr and
    // slice_width then describe the off-diagonal entries, 1-byte codes into table, diag dense).
    const uint16_t* lidx = nullptr;
    int hmax = 0;
    int min_waves = 1;  // register budget: waves per SIMD to allow (1 ... 6; widths above 4 entries: 1)
    // packed halo slices (64 halo rows each; see k_step2_lds)
    const int64_t* hs_first = nullptr;
    const int64_t* hs_ptr = nullptr;
    const int32_t* hs_width = nullptr;
    const int32_t* hrow = nullptr;
    const int32_t* hidx = nullptr;
    const void* hval = nullptr;
    const uint8_t* hcode = nullptr;
    const void* hdiag = nullptr;
    const uint8_t* codes = nullptr;
    const void* diag = nullptr;
    const void* table = nullptr;
    int table_size = 0;
};
// rows of whole 128-byte lines (the duplicate halo stores must not share a line with other rows)
bool step2_supported(int scalar, int vw);
int step2_grid(int64_t ntiles, int num_wg);
int step2_partial_width(int scalar, int vw);
int launch_step2(const Step2Launch& S, hipStream_t stream);
constexpr int STEP2_LDS_TILE = 256;
constexpr size_t STEP2_LDS_MAX_BYTES = 65536;  // dynamic LDS of one workgroup of the LDS form
// fp64 blocks of 8 or 16 vectors (4 or 8 column pairs) and complex128 blocks of 4 or 8 vectors;
// LDS bytes of one workgroup
bool step2_lds_supported(int scalar, int vw);
size_t step2_lds_bytes(int scalar, int vw, int hmax);
// resident workgroups per CU of the LDS form (0 when it cannot run)
int step2_lds_blocks_per_cu(const Step2Launch& S);

// The batched step kernel runs one walker per wave.
constexpr int BATCH_WALKERS_PER_WG = 4;
int launch_batched_step(const BatchedLaunch& B, hipStream_t stream);
// Workgroups of the batched step kernel that B selects that fit on one CU at once.
int batched_blocks_per_cu(const BatchedLaunch& B);
// All steps of the batched recurrence in one persistent launch (partials: 2 x num_slots x width
// doubles); the grid (B.num_wg) must be resident -- at most num_cu x batched_persistent_blocks_per_cu;
// afterwards barrier[2] != 0 reports a grid that was not.
int launch_batched_persistent(const BatchedLaunch& B, hipStream_t stream);
int batched_persistent_blocks_per_cu(const BatchedLaunch& B);
void launch_reduce_segments(const double* partials, const int64_t* slot_first, int64_t num_sys, int width, double* sums,
                            hipStream_t stream);

struct AdjointLaunch {
    int scalar = 1, vw = 4, num_cu = 256;
    const int32_t* indices = nullptr;
    const void* vals = nullptr;
    const int64_t* slice_ptr = nullptr;
    const int32_t* slice_width = nullptr;
    const void* lam1 = nullptr;
    const void* lam2 = nullptr;
    void* lam_out = nullptr;
    const void* v_j = nullptr;
    // reverse recurrence (optional): v_prev = H2 v_j - v_next, from the product the step forms anyway
    const void* v_next = nullptr;
    void* v_prev = nullptr;
    double ca = 0, chv = 0, cg = 0, gscale = 1;
    int update = 1, max_width = 0;
    double* grad = nullptr;
    int64_t n = 0;
    int tune = 0;  // tile shape + NUM_TILE_SHAPES * grid fraction, as for the step kernel
};
int launch_adjoint_step(const AdjointLaunch& A, hipStream_t stream);

int launch_step(const StepLaunch& L, hipStream_t stream);
int step_grid(const StepLaunch& L);
// whether this launch runs the fp64 column-pair kernel, whose partial sums need the pair remap of
// launch_reduce_partials (pair_cols)
bool step_uses_pairs(const StepLaunch& L);
constexpr int NUM_TILE_SHAPES = 3;     // rows per wave per tile (k_step_sellb)
constexpr int NUM_GRID_VARIANTS = 5;   // resident grid, 1/2, 1/4, 2x, 4x
// tile shape x grid x staging (workgroup LDS, wave shuffles, fp64 column pairs, fp64 column
// pairs on the value-table form of the matrix)
constexpr int NUM_TUNE_VARIANTS = 60;
int64_t step_rows_per_tile(int format, int vw);
// sums[s][j] = sum over workgroups of partials[s][w][j] (width doubles per workgroup); with
// pair_cols > 0 the partials are column-pair sums of pair_cols columns and the output has the
// standard [4][pair_cols] layout
void launch_reduce_partials(const double* partials, int64_t num_wg, int width, int64_t nsteps, double* sums,
                            hipStream_t stream, int pair_cols = 0);
// diagonal moments of `cols` columns from the raw per-step sums (steps x 4 vw doubles) into mu
// (device; column c at mu + (c0 + c) * M scalars of the scalar type's real kind, complex as pairs)
void launch_diag_collect(int scalar, const double* sums, int vw, int64_t cols, int64_t M, int64_t c0, void* mu,
                         hipStream_t s);
void launch_gather_block(int scalar, const void* src, int64_t n, int64_t c0, int64_t cols, int vw, void* dst,
                         hipStream_t s);
void launch_scatter_block(int scalar, const void* src, int64_t n, int vw, int64_t c0, int64_t cols, void* dst,
                          hipStream_t s);
// columns [c0, c0 + w) of a row-major n x src_vw block into a row-major n x w block
void launch_copy_cols(int scalar, const void* src, int64_t n, int src_vw, int c0, int w, void* dst, hipStream_t s);
void launch_zero_tails(uint16_t* base, int64_t rows, int64_t ld, int64_t n, hipStream_t s);
// the bf16 pieces of scale * src (columns 0..cols-1) into split planes (stride `plane` elements)
void launch_scatter_split(int scalar, const void* src, int64_t n, int vw, int64_t cols, double scale, uint16_t* dst,
                          int64_t plane, hipStream_t s);
void launch_scatter_scaled(int scalar, const void* src, int64_t n, int vw, int64_t cols, double scale, void* dst,
                           int64_t col_stride, hipStream_t s);
// C (M x M, {re, im} doubles, or doubles for real scalars) += sum_k A[n][k] conj(B[m][k]) over
// k < kdim, A and B row-major with leading dimensions lda / ldb (elements).  Deterministic
// split-K: `work` must hold gemm_workspace_doubles(M, kdim, scalar, mode, num_cu) doubles.
int64_t gemm_workspace_doubles(int64_t M, int64_t kdim, int scalar, int mode, int num_cu);
// The bf16x3 contraction on pre-split operands (store_split planes: row stride ld, a multiple of 8
// with a zero tail beyond kdim; plane stride `plane`); workspace as for mode 2.
int launch_gemm_nc_planes(bool cplx, const uint16_t* A, const uint16_t* B, int64_t ld, int64_t plane, int64_t M,
                          int64_t kdim, double* C, double* work, int num_cu, hipStream_t s);
int launch_gemm_nc(int scalar, int mode, const void* A, int64_t lda, const void* B, int64_t ldb, int64_t M,
                   int64_t kdim, double* C, double* work, int num_cu, hipStream_t s);
// Reconstruction sums (kpmx_reconstruct.hip).  real: 0 float, 1 double.  Chebyshev series
// out[e][c] = pref[e] sum_n mu[c][n] cos(n theta[e]) (exp_series: mu, pref complex and
// exp(-i n theta[e]) in place of the cosine); returns -1 for an unsupported size or type.
int launch_cheb_series(int real, bool exp_series, const void* mu, int64_t M, int64_t cols, int64_t num_energy,
                       const void* theta, const void* pref, double* out, hipStream_t s);
// KuboBastin sample sums out[q] (complex) from the M x M complex moments; t (P M doubles),
// a (2 P M doubles) and partial (2 P groups doubles) are workspaces.
int kubo_sums_groups(int64_t M, int64_t P);
void launch_kubo_sums(const double* mu, int64_t M, const double* samples, int64_t P, double* t, double* a,
                      double* partial, int groups, double* out, hipStream_t s);
// streaming probes: kind 0 copies `bytes` from a to b, kind 1 reads a, kind 2 adds a into b in place
void launch_probe(int kind, const void* a, void* b, int64_t bytes, int grid, hipStream_t stream);
void launch_zero(int scalar, void* dst, int64_t count, hipStream_t s);
void launch_zero_f64(double* dst, int64_t count, hipStream_t s);
void launch_permute_f64(const double* src, const int64_t* pos, int64_t count, double* dst, hipStream_t s);
void launch_csr_matvec(int scalar, const int32_t* indptr, const int32_t* indices, const void* vals, const void* x,
                       int64_t n, void* y, hipStream_t s);

void launch_lanczos_spmv(int scalar, const int32_t* indptr, const int32_t* indices, const void* vals, const void* v1,
                         void* v0, int64_t n, double b_prev, double* partials, int nblocks, hipStream_t s);
void launch_lanczos_axpy(int scalar, const void* v1, void* v0, int64_t n, double a, double* partials, int nblocks,
                         hipStream_t s);
void launch_scale(int scalar, void* v, int64_t n, double f, hipStream_t s);

inline size_t scalar_bytes(int scalar) {
    switch (scalar) {
        case 0: return 4;
        case 1: return 8;
        case 2: return 8;
        default: return 16;
    }
}

}  // namespace kpmx
