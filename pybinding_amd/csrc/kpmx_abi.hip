// C-ABI of the KPM engine (declared in include/kpmx.h): contexts, resident matrices and starter
// blocks, and the recurrence drivers for each kind of moments.  All heavy work runs in the
// kernels of kpmx_kernels.hip / kpmx_gemm.hip on the context's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <map>
#include <atomic>
#include <cmath>
#include <limits>
#include <random>
#include <cstring>
#include <initializer_list>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/kpmx.h"
#include "kpmx_internal.hpp"

using namespace kpmx;

namespace {

thread_local std::string g_last_error;

struct Error {
    int code;
    std::string msg;
};

#define KPMX_HIP(expr)                                                                        \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) throw Error{KPMX_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)}; \
    } while (0)

#define KPMX_REQUIRE(cond, msg)                          \
    do {                                                 \
        if (!(cond)) throw Error{KPMX_EINVAL, (msg)};    \
    } while (0)

template <class F>
int guarded(F&& f) {
    try {
        f();
        return KPMX_OK;
    } catch (Error const& e) {
        g_last_error = e.msg;
        return e.code;
    } catch (std::bad_alloc const&) {
        g_last_error = "host allocation failed";
        return KPMX_ENOMEM;
    } catch (std::exception const& e) {
        g_last_error = e.what();
        return KPMX_EINVAL;
    }
}

void* dev_alloc(size_t bytes) {
    void* p = nullptr;
    if (bytes == 0) bytes = 256;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error{KPMX_ENOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e)};
    }
    return p;
}

void dev_free(void* p) {
    if (p) (void)hipFree(p);
}

// Grow-only device workspace.
struct Workspace {
    void* ptr = nullptr;
    size_t bytes = 0;
    void* get(size_t need) {
        if (need > bytes) {
            // the old buffer may still be read by queued kernels: drain the device first
            if (ptr) (void)hipDeviceSynchronize();
            dev_free(ptr);
            ptr = nullptr;
            bytes = 0;
            ptr = dev_alloc(need);
            bytes = need;
        }
        return ptr;
    }
    // give the memory back now (queued kernels that read it are drained first)
    void release() {
        if (ptr) (void)hipDeviceSynchronize();
        dev_free(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    ~Workspace() { dev_free(ptr); }
};

bool valid_scalar(int s) { return s >= 0 && s <= 3; }
bool is_complex(int s) { return s >= 2; }

}  // namespace

// Device scratch of the allocation-free diagonal entry point (kpmx_moments_diagonal_noalloc):
// one static array per GPU, part of the library's code object and held until the process ends.
// In host code `scratch` is only a symbol: its device address comes from hipGetSymbolAddress,
// taken once per device with that device current (scratch_address, called by kpmx_ctx_create).
constexpr size_t KPMX_SCRATCH_BYTES = (size_t)64 << 20;
static __device__ __align__(256) char scratch[KPMX_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;  // device address of `scratch`, per device

void* scratch_address(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize((size_t)device + 1, nullptr);
    if (!g_scratch_ptr[(size_t)device]) {
        void* p = nullptr;
        KPMX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        KPMX_REQUIRE(p != nullptr, "no device address for the scratch array");
        g_scratch_ptr[(size_t)device] = p;
    }
    return g_scratch_ptr[(size_t)device];
}
}  // namespace

// Widest column tile of the diagonal recurrence by default: fp64 recurrences run in tiles of 16
// vectors (on the 4.19 M-site workload four 16-column recurrences take 1.25 ms per step of 64
// vectors, one 64-column recurrence 1.29 ms), the other types in one tile of up to 64.
int default_tile_cap(int scalar) { return scalar == KPMX_F64 ? 16 : 64; }

struct kpmx_ctx {
    int device = 0;
    int num_cu = 256;
    char* scratch = nullptr;  // this device's `scratch` array (device address)
    hipStream_t stream = nullptr;
    Workspace partials, sums, work0, work1, work2, aux0, aux1, aux2, panel_l, panel_r;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    int tune[4][7] = {};  // tile-shape variant per (scalar, log2 vw), chosen by kpmx_autotune_diagonal
    // default variant for matrices with the value-table form (-1: `tune`); cleared when a variant
    // is pinned or tuned, which then applies to every matrix
    int tune_coded[4][7] = {{-1, -1, -1, -1, -1, -1, -1}, {-1, -1, -1, -1, -1, -1, -1},
                            {-1, -1, -1, -1, -1, -1, -1}, {-1, -1, -1, -1, -1, -1, -1}};
    // widest column tile the diagonal recurrence runs per scalar type (kpmx_ctx_set_tile_width):
    // wider starter tiles are run as several recurrences over column sub-tiles
    int tile_cap[4] = {64, 64, 64, 64};
    // rows per tile of the two-step diagonal recurrence per scalar type; 0: single steps
    // (kpmx_ctx_set_two_step)
    int two_step[4] = {0, 0, 0, 0};
    int two_step_wg_per_cu[4] = {0, 0, 0, 0};  // persistent grid of the two-step kernel; 0: a workgroup per tile
    int two_step_lds[4] = {0, 0, 0, 0};        // the LDS form (k_step2_lds) of the two-step recurrence
    bool autotune_two_step = true;             // kpmx_autotune_diagonal may choose the LDS two-step form
    bool two_step_auto[4] = {false, false, false, false};  // the two-step settings are the autotuner's
    Workspace work3, work4;  // the third and fourth vector of the two-step recurrence
    int graph_policy = KPMX_GRAPH_NEVER;
    int adjoint_variant = -1;  // adjoint step variant (kpmx_ctx_set_adjoint_variant); -1: the default
    // device memory of the largest finished adjoint tape, reused by the next tape that fits: a
    // stored tape of a 1 M-site, 16-vector, 1024-moment gradient is 65 GB, whose allocation and
    // mapping would otherwise be paid (and vary widely) on every gradient
    struct TapePool {
        char* ptr = nullptr;
        size_t bytes = 0;
        bool busy = false;
        ~TapePool() { dev_free(ptr); }
    } tape_pool;
};

// Host-built arrays of the LDS two-step form for one matrix form: the tile-local position of every
// stored entry's column (lidx) and the tiles' packed halo slices (see k_step2_lds).
struct LdsHalo {
    uint16_t* lidx = nullptr;
    int64_t* hs_first = nullptr;
    int64_t* hs_ptr = nullptr;
    int32_t* hs_width = nullptr;
    int32_t* hrow = nullptr;
    int32_t* hidx = nullptr;
    void* hval = nullptr;
    uint8_t* hcode = nullptr;
    void* hdiag = nullptr;
    void release() {
        for (void* p : {(void*)lidx, (void*)hs_first, (void*)hs_ptr, (void*)hs_width, (void*)hrow, (void*)hidx, hval,
                        (void*)hcode, hdiag})
            dev_free(p);
        *this = LdsHalo{};
    }
};

struct kpmx_mat {
    kpmx_ctx* ctx = nullptr;
    uint64_t serial = 0;  // unique per upload: keys captured graphs (addresses can be reused)
    int scalar = 1;
    int format = KPMX_FMT_CSR;
    int64_t rows = 0, nnz = 0;
    int32_t* indptr = nullptr;
    int32_t* indices = nullptr;
    void* vals = nullptr;
    // SELL-64 copy (format == SELL)
    int64_t nslices = 0;
    int max_width = 0;
    int64_t* slice_ptr = nullptr;
    int32_t* slice_width = nullptr;
    int32_t* sell_indices = nullptr;
    void* sell_vals = nullptr;
    int64_t sell_total = 0;
    std::vector<int64_t> sell_pos;  // SELL storage position of every CSR entry
    // light cone (kpmx_mat_set_step_rows): rows computed by each step of the next recurrences
    std::vector<int64_t> step_rows;
    mutable int64_t* d_sell_pos = nullptr;  // device copy, uploaded by the first adjoint pass
    // halo rows of the two-step recurrence's tiles of halo_tile rows (built on first use): the
    // rows outside tile j that its rows reference are h1_rows[h1_ptr[j] .. h1_ptr[j+1])
    mutable int64_t halo_tile = 0, halo_ntiles = 0, halo_rows = 0;
    mutable int64_t* d_h1_ptr = nullptr;
    mutable int32_t* d_h1_rows = nullptr;
    // the LDS two-step form's arrays (tiles of STEP2_LDS_TILE rows) for the standard SELL and for
    // the value-table SELL, and the most halo rows of a tile
    mutable LdsHalo lds_std, lds_coded;
    mutable int lds_hmax = -1;
    // value-table form (SELL over the off-diagonal entries with 8-bit codes into `ctable`, the
    // diagonal as a dense vector), built at upload when the off-diagonal entries take at most 256
    // distinct values (lattices with a few hopping families); null otherwise
    int64_t* cslice_ptr = nullptr;
    int32_t* cslice_width = nullptr;
    int32_t* cindices = nullptr;
    uint8_t* ccodes = nullptr;
    void* cdiag = nullptr;
    void* ctable = nullptr;
    int cmax_width = 0, ctable_size = 0;
    ~kpmx_mat() {
        for (void* p : {(void*)cslice_ptr, (void*)cslice_width, (void*)cindices, (void*)ccodes, cdiag, ctable}) dev_free(p);
        dev_free(d_h1_ptr);
        dev_free(d_h1_rows);
        lds_std.release();
        lds_coded.release();
        dev_free(d_sell_pos);
        dev_free(indptr);
        dev_free(indices);
        dev_free(vals);
        dev_free(slice_ptr);
        dev_free(slice_width);
        dev_free(sell_indices);
        dev_free(sell_vals);
    }
};

struct Tile {
    int64_t c0 = 0, cols = 0;
    int vw = 1;
    void* data = nullptr;   // n x vw row-major pristine starters
};

struct kpmx_block {
    kpmx_ctx* ctx = nullptr;
    int scalar = 1;
    int64_t n = 0, nvec = 0;
    std::vector<Tile> tiles;
    // raw sums of the last kpmx_moments_diagonal_run, per run tile: (M/2) x (4 vw) doubles
    std::vector<void*> sums;
    int64_t sums_moments = 0;
    int sums_cap = 0;  // column tile cap the sums were run with
    // captured diagonal recurrence and the launch state it was captured with
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    std::vector<uint64_t> graph_key, last_key;
    void drop_graph() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph = nullptr;
        graph_key.clear();
    }
    ~kpmx_block() {
        drop_graph();
        for (auto& t : tiles) dev_free(t.data);
        for (auto* s : sums) dev_free(s);
    }
};

// Batched independent systems (block-diagonal H2, one block per system, each starting on a
// 64-row slice boundary) with their starter vectors and device workspaces.
// Widest column tile of a batch's vectors per scalar type (f32, f64, c64, c128).  One block of
// all vectors by default: on 256 fp64 flakes (31 M sites) 16, 32 and 64 vectors took 306, 638 and
// 1282 ms per 256 moments in one block against 321, 642 and 1285 ms as tiles of 8 (tiles of 4:
// 357, 715, 1430 ms); narrower tiles save work memory (two tile-wide work blocks).
constexpr int BATCH_TILE_DEFAULT[4] = {64, 64, 64, 64};

struct kpmx_batch {
    kpmx_ctx* ctx = nullptr;
    int scalar = 1;
    int64_t num_sys = 0;
    std::vector<int64_t> row_off, pad_off;  // unpadded / padded first row of every system (S + 1)
    int64_t n_pad = 0, nslices = 0;
    std::vector<int32_t> slice_sys_host;  // system of every slice
    int max_width = 0;
    int64_t* slice_ptr = nullptr;
    int32_t* slice_width = nullptr;
    int32_t* slice_sys = nullptr;
    int32_t* sidx = nullptr;
    void* svals = nullptr;
    // the value-table form (build_batch_coded): off-diagonal SELL with 1-byte codes into one table
    // of at most 256 values per system, dense diagonal over the padded rows
    bool has_coded = false, use_coded = false;
    int cmax_width = 0;
    int64_t s_entries = 0;  // stored slots of the standard form (padding included)
    int64_t c_entries = 0;  // stored off-diagonal slots (padding included)
    int64_t table_len = 0;
    int64_t* cslice_ptr = nullptr;
    int32_t* cslice_width = nullptr;
    int32_t* cidx = nullptr;
    uint8_t* ccodes = nullptr;
    void* cdiag = nullptr;
    void* ctable = nullptr;
    int32_t* ctab_off = nullptr;
    int num_wg = 0;
    int walk_chunk = 0;  // BatchWalk of the step kernel
    int64_t num_slots = 0;
    int64_t* tile_begin = nullptr;
    int64_t* slot_begin = nullptr;
    int64_t* slot_index = nullptr;
    int64_t* slot_first = nullptr;
    // starters and workspaces: the vectors run as ntiles column tiles of tile_cols vectors (the
    // last may be narrower), each a row-major n_pad x vw block of r0 and its own recurrence
    int64_t nvec = 0;
    int vw = 4;
    int tile_cap = 0;  // widest tile (0: the default of the scalar type)
    int64_t ntiles = 1, tile_cols = 0;
    void* r0 = nullptr;
    void* w0 = nullptr;
    void* w1 = nullptr;
    double* partials = nullptr;
    double* partials2 = nullptr;     // the persistent run's two partial buffers
    unsigned* barrier = nullptr;     // the persistent run's grid barrier
    double* sums = nullptr;
    int64_t sums_steps = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int64_t graph_moments = 0;
    void drop_graph() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph = nullptr;
        graph_moments = 0;
    }
    ~kpmx_batch() {
        drop_graph();
        for (void* p : {(void*)slice_ptr, (void*)slice_width, (void*)slice_sys, (void*)sidx, svals, (void*)tile_begin,
                        (void*)slot_begin, (void*)slot_index, (void*)slot_first, r0, w0, w1, (void*)partials, (void*)partials2,
                        (void*)barrier, (void*)sums, (void*)cslice_ptr, (void*)cslice_width, (void*)cidx, (void*)ccodes,
                        cdiag, ctable, (void*)ctab_off})
            dev_free(p);
    }
};

// Forward record of an adjoint calculation: every vector v_0..v_S, or checkpoint pairs
// (v_cC, v_cC+1) plus a segment buffer for the recomputed vectors.
struct kpmx_tape {
    const kpmx_mat* h2 = nullptr;
    kpmx_ctx* ctx = nullptr;
    int64_t M = 0, S = 0, C = 0, nvec = 0, nslots = 0, nck = 0;
    int vw = 4;
    size_t blk = 0, slots_bytes = 0;
    bool store_all = true;
    bool pooled = false;  // slots borrowed from the context's tape pool
    char* slots = nullptr;
    char* cks = nullptr;
    void* padded = nullptr;
    char* slot(int64_t j) const { return slots + (size_t)j * blk; }
    ~kpmx_tape() {
        if (pooled) {
            ctx->tape_pool.busy = false;
        } else if (ctx && slots && !ctx->tape_pool.busy && slots_bytes > ctx->tape_pool.bytes) {
            dev_free(ctx->tape_pool.ptr);   // keep the larger buffer for the next tape
            ctx->tape_pool.ptr = slots;
            ctx->tape_pool.bytes = slots_bytes;
        } else {
            dev_free(slots);
        }
        dev_free(cks);
        dev_free(padded);
    }
};

namespace {

int pick_vw(int64_t remaining) {
    int vw = 1;
    while (vw < remaining && vw < 64) vw <<= 1;
    return vw;
}

int grid_size(kpmx_ctx* ctx, int64_t num_tiles) {
    int64_t max_wg = (int64_t)ctx->num_cu * 8;
    int64_t g = std::min<int64_t>(num_tiles, max_wg);
    if (g >= 8) g -= g % 8;
    if (g < 1) g = 1;
    return (int)g;
}

StepLaunch base_launch(const kpmx_mat* h2, int vw, bool use_sell) {
    StepLaunch L;
    L.scalar = h2->scalar;
    L.vw = vw;
    L.n = h2->rows;
    if (use_sell && (vw == 1 || vw >= 4) && h2->format == KPMX_FMT_SELL) {
        L.format = FORMAT_SELL;
        L.indices = h2->sell_indices;
        L.vals = h2->sell_vals;
        L.slice_ptr = h2->slice_ptr;
        L.slice_width = h2->slice_width;
        L.max_width = h2->max_width;
    } else {
        L.format = FORMAT_CSR;
        L.indptr = h2->indptr;
        L.indices = h2->indices;
        L.vals = h2->vals;
    }
    L.num_cu = h2->ctx->num_cu;
    L.num_wg = grid_size(h2->ctx, (h2->rows + step_rows_per_tile(L.format, vw) - 1) / step_rows_per_tile(L.format, vw));
    if (L.format == FORMAT_SELL && h2->ccodes) {
        L.cslice_ptr = h2->cslice_ptr;
        L.cslice_width = h2->cslice_width;
        L.cindices = h2->cindices;
        L.ccodes = h2->ccodes;
        L.cdiag = h2->cdiag;
        L.ctable = h2->ctable;
        L.cmax_width = h2->cmax_width;
        L.ctable_size = h2->ctable_size;
    }
    return L;
}

void launch(const StepLaunch& L, hipStream_t s) {
    if (launch_step(L, s) != 0) throw Error{KPMX_EINVAL, "unsupported step configuration"};
    KPMX_HIP(hipGetLastError());
}

// Run the Chebyshev recurrence r_{k+1} = H2 r_k - r_{k-1} over one tile for `steps` launches
// (k = 0 is the first step r_1 = 0.5 H2 r_0).  `per_step` customises each launch (collection
// pointers) and the partials of reduction modes are summed into `sums` (steps x 4vw doubles).
int log2i(int v) {
    int l = 0;
    while ((1 << l) < v) ++l;
    return l;
}

// The step variant of a (matrix, vector width): the context's choice, or its default for
// matrices with the value-table form.
int tuned_variant(const kpmx_mat* h2, int vw) {
    const kpmx_ctx* ctx = h2->ctx;
    const int lv = log2i(vw);
    if (h2->ccodes && ctx->tune_coded[h2->scalar][lv] >= 0) return ctx->tune_coded[h2->scalar][lv];
    return ctx->tune[h2->scalar][lv];
}

// Steps whose per-workgroup partial sums are buffered before one reduction launch (256 MiB cap).
int64_t partials_chunk(size_t per_step_bytes, int64_t steps) {
    const size_t budget = (size_t)256 << 20;
    return std::max<int64_t>(1, std::min<int64_t>(steps, (int64_t)(budget / per_step_bytes)));
}

// Buffers a caller hands to run_recurrence instead of the context's grow-only workspaces.
struct FixedBuffers {
    void* w0 = nullptr;
    void* w1 = nullptr;
    double* partials = nullptr;
    size_t partials_bytes = 0;
};

template <class PerStep, class PostStep>
void run_recurrence(const kpmx_mat* h2, const Tile& t, int mode, int64_t steps, const void* w, double* sums,
                    PerStep per_step, PostStep post_step, int tune = -1, const FixedBuffers* fixed = nullptr) {
    kpmx_ctx* ctx = h2->ctx;
    const int64_t n = h2->rows;
    const size_t sb = scalar_bytes(h2->scalar);
    void* W0 = fixed ? fixed->w0 : ctx->work0.get((size_t)n * t.vw * sb);
    void* W1 = fixed ? fixed->w1 : ctx->work1.get((size_t)n * t.vw * sb);

    StepLaunch L = base_launch(h2, t.vw, true);
    L.mode = mode;
    L.w = w;
    L.tune = tune >= 0 ? tune : tuned_variant(h2, t.vw);
    L.num_wg = step_grid(L);  // the SELL block kernel runs a resident grid
    const int width = 4 * t.vw;  // raw sums per step
    // the fp64 column-pair kernel writes half-width partials in its own layout; the reduction
    // brings them to the standard one
    const bool pairs = mode == MODE_DIAG && step_uses_pairs(L);
    const int pwidth = pairs ? 2 * t.vw : width;
    int64_t chunk = 1;
    double* partials = nullptr;
    if (mode != MODE_NONE) {
        const size_t per_step_bytes = (size_t)L.num_wg * pwidth * sizeof(double);
        if (fixed) {
            KPMX_REQUIRE(fixed->partials_bytes >= per_step_bytes, "partial-sum buffer too small");
            chunk = std::max<int64_t>(1, std::min<int64_t>(steps, (int64_t)(fixed->partials_bytes / per_step_bytes)));
            partials = fixed->partials;
        } else {
            chunk = partials_chunk(per_step_bytes, steps);
            partials = static_cast<double*>(ctx->partials.get(per_step_bytes * chunk));
        }
    }
    // light cone: step k computes only rows [0, step_rows[k]); rows past it must read as the zeros
    // they are in the exact recurrence, so the first work block written after step 0 starts zeroed
    const std::vector<int64_t>& cone = h2->step_rows;
    if (!cone.empty()) {
        launch_zero(h2->scalar, W0, n * t.vw, ctx->stream);
        KPMX_HIP(hipGetLastError());
    }
    int64_t chunk_start = 0;
    for (int64_t k = 0; k < steps; ++k) {
        L.n = (k > 0 && k < (int64_t)cone.size()) ? std::max<int64_t>(1, std::min<int64_t>(n, cone[(size_t)k])) : n;
        if (k == 0) {
            L.x = t.data;
            L.yin = nullptr;
            L.yout = W1;
            L.hscale = 0.5;
            L.sub_y = 0;
        } else {
            const bool odd = (k % 2) == 1;
            L.x = odd ? W1 : W0;
            L.yout = odd ? W0 : W1;
            L.yin = (k == 1) ? t.data : L.yout;
            L.hscale = 1.0;
            L.sub_y = 1;
        }
        if (partials) L.partials = partials + (size_t)(k - chunk_start) * L.num_wg * pwidth;
        per_step(L, k);
        launch(L, ctx->stream);
        post_step(L, k);
        if (partials && (k - chunk_start + 1 == chunk || k == steps - 1)) {
            launch_reduce_partials(partials, L.num_wg, pwidth, k - chunk_start + 1, sums + (size_t)chunk_start * width,
                                   ctx->stream, pairs ? t.vw : 0);
            KPMX_HIP(hipGetLastError());
            chunk_start = k + 1;
        }
    }
}

// Host-side collector for the diagonal recurrence (collectors.cpp:5-34), from raw sums.
// Step k = 0 gives A = |r0|^2 and B = sum conj(r1) r0 = r1.dot(r0); step k >= 1 gives
// A = |r_k|^2 and B = sum conj(r_{k+1}) r_k, so mu_2k = 2 (A - m0) and mu_2k+1 = 2 B - m1
// with m0 = A_0 / 2 and m1 = B_0 (the doubling identities T_2k = 2 T_k^2 - T_0 and
// T_2k+1 = 2 T_k+1 T_k - T_1).
template <class R>
void diag_collect(const std::vector<double>& s, int vw, int64_t cols, int64_t M, bool cplx, R* mu_base,
                  int64_t c0) {
    const int width = 4 * vw;
    for (int64_t c = 0; c < cols; ++c) {
        R* mu = mu_base + (c0 + c) * M * (cplx ? 2 : 1);
        auto put = [&](int64_t m, double re, double im) {
            if (cplx) {
                mu[2 * m] = (R)re;
                mu[2 * m + 1] = (R)im;
            } else {
                mu[m] = (R)re;
            }
        };
        const double m0 = 0.5 * s[0 * width + 0 * vw + c];
        const double m1r = s[0 * width + 2 * vw + c], m1i = s[0 * width + 3 * vw + c];
        put(0, m0, 0.0);
        put(1, m1r, m1i);
        for (int64_t k = 1; k < M / 2; ++k) {
            const double a = s[k * width + 0 * vw + c];
            const double br = s[k * width + 2 * vw + c], bi = s[k * width + 3 * vw + c];
            put(2 * k, 2.0 * (a - m0), 0.0);
            put(2 * k + 1, 2.0 * br - m1r, 2.0 * bi - m1i);
        }
    }
}

// Halo rows of every tile of P rows: for tile j the sorted distinct columns of its rows' SELL
// entries that lie outside the tile (read back from the device copy once per tile size).
void ensure_halo(const kpmx_mat* h2, int64_t P) {
    if (h2->halo_tile == P && h2->d_h1_ptr) return;
    KPMX_REQUIRE(h2->format == KPMX_FMT_SELL, "the two-step recurrence needs the SELL format");
    kpmx_ctx* ctx = h2->ctx;
    const int64_t n = h2->rows, ns = h2->nslices;
    std::vector<int64_t> sptr(ns + 1);
    std::vector<int32_t> swidth(ns), sidx((size_t)std::max<int64_t>(h2->sell_total, 1));
    KPMX_HIP(hipMemcpyAsync(sptr.data(), h2->slice_ptr, (ns + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
    KPMX_HIP(hipMemcpyAsync(swidth.data(), h2->slice_width, ns * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    if (h2->sell_total)
        KPMX_HIP(hipMemcpyAsync(sidx.data(), h2->sell_indices, h2->sell_total * sizeof(int32_t), hipMemcpyDeviceToHost,
                                ctx->stream));
    KPMX_HIP(hipStreamSynchronize(ctx->stream));
    const int64_t ntiles = (n + P - 1) / P;
    std::vector<int64_t> hptr(ntiles + 1, 0);
    std::vector<int32_t> hrows;
    std::vector<int64_t> stamp((size_t)n, -1);
    std::vector<int32_t> cur;
    for (int64_t j = 0; j < ntiles; ++j) {
        const int64_t t0 = j * P, t1 = std::min(n, t0 + P);
        cur.clear();
        for (int64_t r = t0; r < t1; ++r) {
            const int64_t slc = r >> 6, lane = r & 63;
            for (int32_t k = 0; k < swidth[slc]; ++k) {
                const int32_t c = sidx[(size_t)(sptr[slc] + (int64_t)k * 64 + lane)];
                if ((c < t0 || c >= t1) && stamp[(size_t)c] != j) {
                    stamp[(size_t)c] = j;
                    cur.push_back(c);
                }
            }
        }
        std::sort(cur.begin(), cur.end());
        hrows.insert(hrows.end(), cur.begin(), cur.end());
        hptr[j + 1] = (int64_t)hrows.size();
    }
    KPMX_HIP(hipStreamSynchronize(ctx->stream));
    dev_free(h2->d_h1_ptr);
    dev_free(h2->d_h1_rows);
    h2->d_h1_ptr = nullptr;
    h2->d_h1_rows = nullptr;
    h2->halo_tile = 0;
    h2->d_h1_ptr = static_cast<int64_t*>(dev_alloc(hptr.size() * sizeof(int64_t)));
    h2->d_h1_rows = static_cast<int32_t*>(dev_alloc(std::max<size_t>(hrows.size(), 1) * sizeof(int32_t)));
    KPMX_HIP(hipMemcpyAsync(h2->d_h1_ptr, hptr.data(), hptr.size() * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream));
    if (!hrows.empty())
        KPMX_HIP(hipMemcpyAsync(h2->d_h1_rows, hrows.data(), hrows.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                                ctx->stream));
    KPMX_HIP(hipStreamSynchronize(ctx->stream));
    h2->halo_tile = P;
    h2->halo_ntiles = ntiles;
    h2->halo_rows = (int64_t)hrows.size();
}

// Whether the LDS two-step form can take this matrix: a value-table form the single step would
// read (at most 4 off-diagonal entries per row) or the standard SELL.
bool lds_coded(const kpmx_mat* h2) { return h2->ccodes && h2->cmax_width <= 4; }

// The LDS two-step form's arrays for the matrix form the kernel reads (the value-table SELL when
// lds_coded, else the standard SELL), built once from device copies (tiles of STEP2_LDS_TILE rows):
//  - lidx: every stored entry's column as its tile-local position -- a column inside the tile is
//    its offset in the tile, one outside is STEP2_LDS_TILE + its index in the tile's sorted halo;
//  - the packed halo slices: the tile's halo rows in groups of 64, each row's stored entries (its
//    slice's width, padding included) column-major, so the kernel reads them like a SELL slice.
void ensure_lidx(const kpmx_mat* h2) {
    constexpr int64_t P = STEP2_LDS_TILE;
    ensure_halo(h2, P);
    const bool coded = lds_coded(h2);
    LdsHalo& L = coded ? h2->lds_coded : h2->lds_std;
    if (L.lidx) return;
    kpmx_ctx* ctx = h2->ctx;
    const int64_t n = h2->rows, ns = h2->nslices, ntiles = h2->halo_ntiles;
    std::vector<int64_t> hptr(ntiles + 1), sptr(ns + 1);
    std::vector<int32_t> hrows((size_t)std::max<int64_t>(h2->halo_rows, 1)), swidth(ns);
    auto down = [&](void* dst, const void* src, size_t bytes) {
        if (bytes) KPMX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    };
    down(hptr.data(), h2->d_h1_ptr, hptr.size() * sizeof(int64_t));
    down(hrows.data(), h2->d_h1_rows, h2->halo_rows * sizeof(int32_t));
    down(sptr.data(), coded ? h2->cslice_ptr : h2->slice_ptr, sptr.size() * sizeof(int64_t));
    down(swidth.data(), coded ? h2->cslice_width : h2->slice_width, swidth.size() * sizeof(int32_t));
    KPMX_HIP(hipStreamSynchronize(ctx->stream));
    const int64_t total = sptr[ns];
    const size_t sb = scalar_bytes(h2->scalar);  // values are copied as bytes
    std::vector<int32_t> idx((size_t)std::max<int64_t>(total, 1));
    std::vector<char> val(coded ? 1 : (size_t)std::max<int64_t>(total, 1) * sb), diag(coded ? (size_t)n * sb : 1);
    std::vector<uint8_t> code(coded ? (size_t)std::max<int64_t>(total, 1) : 1);
    std::vector<char> table(coded ? (size_t)std::max(h2->ctable_size, 1) * sb : 1, 0);
    down(idx.data(), coded ? h2->cindices : h2->sell_indices, total * sizeof(int32_t));
    if (coded) {
        down(code.data(), h2->ccodes, total);
        down(diag.data(), h2->cdiag, n * sb);
        down(table.data(), h2->ctable, h2->ctable_size * sb);
    } else {
        down(val.data(), h2->sell_vals, total * sb);
    }
    KPMX_HIP(hipStreamSynchronize(ctx->stream));
    int zcode = -1;  // the value table's zero (padding of the packed halo)
    if (coded) {
        const std::vector<char> zero(sb, 0);
        for (int c = 0; c < h2->ctable_size && zcode < 0; ++c)
            if (std::memcmp(&table[(size_t)c * sb], zero.data(), sb) == 0) zcode = c;
        KPMX_REQUIRE(zcode >= 0, "internal error: the value table has no zero");
    }
    int hmax = 0;
    for (int64_t j = 0; j < ntiles; ++j) hmax = std::max<int>(hmax, (int)(hptr[j + 1] - hptr[j]));
    KPMX_REQUIRE(P + hmax < 65536, "the LDS two-step form needs fewer halo rows per tile");
    // tile-local positions
    std::vector<uint16_t> loc((size_t)std::max<int64_t>(total, 1), 0);
    for (int64_t slc = 0; slc < ns; ++slc) {
        const int64_t j = slc * 64 / P, t0 = j * P, t1 = std::min(n, t0 + P);
        const int32_t* h = hrows.data() + hptr[j];
        const int64_t nh = hptr[j + 1] - hptr[j];
        for (int64_t e = sptr[slc]; e < sptr[slc + 1]; ++e) {
            const int64_t c = idx[(size_t)e];
            int64_t pos;
            if (c >= t0 && c < t1) {
                pos = c - t0;
            } else {
                const int32_t* it = std::lower_bound(h, h + nh, (int32_t)c);
                KPMX_REQUIRE(it != h + nh && *it == c, "internal error: a column outside its tile's halo");
                pos = P + (it - h);
            }
            loc[(size_t)e] = (uint16_t)pos;
        }
    }
    // packed halo slices
    std::vector<int64_t> hs_first(ntiles + 1, 0), hs_ptr(1, 0);
    std::vector<int32_t> hs_width, hrow, hidx;
    std::vector<char> hval, hdiag;
    std::vector<uint8_t> hcode;
    for (int64_t j = 0; j < ntiles; ++j) {
        const int64_t t0 = j * P, nh = hptr[j + 1] - hptr[j];
        const int32_t* h = hrows.data() + hptr[j];
        for (int64_t q0 = 0; q0 < nh; q0 += 64) {
            int32_t w = 0;
            for (int64_t r = q0; r < std::min(nh, q0 + 64); ++r) w = std::max(w, swidth[(size_t)(h[r] >> 6)]);
            const size_t eb = hidx.size();
            hidx.resize(eb + (size_t)w * 64);
            if (coded) hcode.resize(eb + (size_t)w * 64, (uint8_t)zcode);
            else hval.resize((eb + (size_t)w * 64) * sb, 0);
            const size_t db = hdiag.size();
            if (coded) hdiag.resize(db + 64 * sb, 0);
            for (int64_t r = 0; r < 64; ++r) {
                const bool real = q0 + r < nh;
                const int64_t row = real ? h[q0 + r] : t0;
                hrow.push_back((int32_t)row);
                if (coded && real) std::memcpy(&hdiag[db + (size_t)r * sb], &diag[(size_t)row * sb], sb);
                const int64_t slc = row >> 6, lane = row & 63;
                for (int32_t k = 0; k < w; ++k) {
                    const size_t dst = eb + (size_t)k * 64 + r;
                    if (real && k < swidth[(size_t)slc]) {
                        const size_t src = (size_t)(sptr[slc] + (int64_t)k * 64 + lane);
                        hidx[dst] = idx[src];
                        if (coded) hcode[dst] = code[src];
                        else std::memcpy(&hval[dst * sb], &val[src * sb], sb);
                    } else {
                        hidx[dst] = (int32_t)row;
                    }
                }
            }
            hs_width.push_back(w);
            hs_ptr.push_back((int64_t)hidx.size());
        }
        hs_first[j + 1] = (int64_t)hs_width.size();
    }
    auto up = [&](const void* src, size_t bytes) -> void* {
        void* d = dev_alloc(std::max<size_t>(bytes, 8));
        if (bytes) KPMX_HIP(hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        return d;
    };
    L.lidx = static_cast<uint16_t*>(up(loc.data(), loc.size() * sizeof(uint16_t)));
    L.hs_first = static_cast<int64_t*>(up(hs_first.data(), hs_first.size() * sizeof(int64_t)));
    L.hs_ptr = static_cast<int64_t*>(up(hs_ptr.data(), hs_ptr.size() * sizeof(int64_t)));
    L.hs_width = static_cast<int32_t*>(up(hs_width.data(), hs_width.size() * sizeof(int32_t)));
    L.hrow = static_cast<int32_t*>(up(hrow.data(), hrow.size() * sizeof(int32_t)));
    L.hidx = static_cast<int32_t*>(up(hidx.data(), hidx.size() * sizeof(int32_t)));
    if (coded) {
        L.hcode = static_cast<uint8_t*>(up(hcode.data(), hcode.size()));
        L.hdiag = up(hdiag.data(), hdiag.size());
    } else {
        L.hval = up(hval.data(), hval.size());
    }
    KPMX_HIP(hipStreamSynchronize(ctx->stream));
    h2->lds_hmax = hmax;
}

// Whether the diagonal recurrence of this column tile runs two steps per launch; the LDS form
// also needs the tile's rows and its largest halo to fit one workgroup's LDS (else single steps).
bool use_two_step(const kpmx_mat* h2, const Tile& t) {
    const kpmx_ctx* ctx = h2->ctx;
    if (ctx->two_step[h2->scalar] <= 0 || h2->format != KPMX_FMT_SELL || !h2->step_rows.empty() || h2->rows < 2 * 64)
        return false;
    if (!ctx->two_step_lds[h2->scalar]) return step2_supported(h2->scalar, t.vw);
    if (!step2_lds_supported(h2->scalar, t.vw)) return false;
    ensure_lidx(h2);
    return step2_lds_bytes(h2->scalar, t.vw, h2->lds_hmax) <= STEP2_LDS_MAX_BYTES;
}

// The launch description of the two-step recurrence of this column tile (matrix, halo, grid);
// vectors and partials are the caller's.
Step2Launch two_step_plan(const kpmx_mat* h2, const Tile& t) {
    kpmx_ctx* ctx = h2->ctx;
    const bool lds = ctx->two_step_lds[h2->scalar] != 0;
    const int64_t P = lds ? STEP2_LDS_TILE : ctx->two_step[h2->scalar];
    if (lds) ensure_lidx(h2);
    else ensure_halo(h2, P);
    Step2Launch S;
    S.scalar = h2->scalar;
    S.vw = t.vw;
    S.max_width = h2->max_width;
    S.indices = h2->sell_indices;
    S.vals = h2->sell_vals;
    S.slice_ptr = h2->slice_ptr;
    S.slice_width = h2->slice_width;
    S.h1_ptr = h2->d_h1_ptr;
    S.h1_rows = h2->d_h1_rows;
    S.n = h2->rows;
    S.ntiles = h2->halo_ntiles;
    S.tile_rows = P;
    if (!lds) {
        S.num_wg = step2_grid(h2->halo_ntiles, ctx->two_step_wg_per_cu[h2->scalar] * ctx->num_cu);
        return S;
    }
    S.hmax = h2->lds_hmax;
    if (lds_coded(h2)) {
        S.max_width = h2->cmax_width;
        S.indices = h2->cindices;
        S.vals = nullptr;
        S.slice_ptr = h2->cslice_ptr;
        S.slice_width = h2->cslice_width;
        S.codes = h2->ccodes;
        S.diag = h2->cdiag;
        S.table = h2->ctable;
        S.table_size = h2->ctable_size;
    }
    const LdsHalo& L = lds_coded(h2) ? h2->lds_coded : h2->lds_std;
    S.lidx = L.lidx;
    S.hs_first = L.hs_first;
    S.hs_ptr = L.hs_ptr;
    S.hs_width = L.hs_width;
    S.hrow = L.hrow;
    S.hidx = L.hidx;
    S.hval = L.hval;
    S.hcode = L.hcode;
    S.hdiag = L.hdiag;
    // workgroups per CU: the caller's (at most 6), else as many as LDS allows up to 4 (tighter
    // register budgets spill: 5 and 6 measured 1.5-2.4x slower); the kernel's registers are
    // then bounded so that many fit
    const int by_lds = (int)std::min<size_t>(4, (size_t)160 * 1024 / (step2_lds_bytes(h2->scalar, t.vw, S.hmax) + 4096));
    const int asked = ctx->two_step_wg_per_cu[h2->scalar];
    S.min_waves = std::max(1, asked > 0 ? std::min(6, asked) : by_lds);
    const int per_cu = step2_lds_blocks_per_cu(S);
    KPMX_REQUIRE(per_cu > 0, "the LDS two-step form does not fit this matrix's halo (" + std::to_string(S.hmax) +
                                 " rows per tile) in LDS");
    const int want = asked > 0 ? asked : per_cu;
    const int64_t all = (h2->halo_ntiles + 7) / 8 * 8;
    S.num_wg = (int)std::max<int64_t>(8, std::min<int64_t>(all, (int64_t)want * ctx->num_cu / 8 * 8));
    return S;
}

// Device memory the two-step recurrence of this tile uses (sized before a graph capture).
void size_two_step(const kpmx_mat* h2, const Tile& t, int64_t steps) {
    kpmx_ctx* ctx = h2->ctx;
    const Step2Launch plan = two_step_plan(h2, t);
    const size_t blk = (size_t)h2->rows * t.vw * scalar_bytes(h2->scalar);
    ctx->work0.get(blk);
    ctx->work1.get(blk);
    ctx->work3.get(blk);
    ctx->work4.get(blk);
    const int grid = plan.num_wg;
    const size_t per_step = (size_t)grid * step2_partial_width(h2->scalar, t.vw) * sizeof(double);
    int64_t chunk = std::max<int64_t>(2, partials_chunk(per_step, steps));
    StepLaunch L = base_launch(h2, t.vw, true);
    L.mode = MODE_DIAG;
    L.tune = tuned_variant(h2, t.vw);
    L.num_wg = step_grid(L);
    const size_t single = (size_t)L.num_wg * 4 * t.vw * sizeof(double);
    ctx->partials.get(std::max(per_step * (size_t)chunk, single));
}

// The diagonal recurrence with two steps per launch (k_step2_sell over tiles of
// ctx->two_step[scalar] rows), the last step of an odd count by the single-step kernel.  r_j
// (j >= 1) lives in the j % 4-th of four work vectors: a launch reads r_{k-1}, r_k and writes
// r_{k+1}, r_{k+2}, so nothing it reads is overwritten.  Same raw sums as run_recurrence.
void run_recurrence2(const kpmx_mat* h2, const Tile& t, int64_t steps, double* sums) {
    kpmx_ctx* ctx = h2->ctx;
    size_two_step(h2, t, steps);
    const int64_t n = h2->rows;
    const size_t blk = (size_t)n * t.vw * scalar_bytes(h2->scalar);
    void* W[4] = {ctx->work0.get(blk), ctx->work1.get(blk), ctx->work3.get(blk), ctx->work4.get(blk)};
    auto r = [&](int64_t j) -> void* { return j == 0 ? t.data : W[j % 4]; };
    Step2Launch S = two_step_plan(h2, t);
    const int grid = S.num_wg;
    const int pw = step2_partial_width(h2->scalar, t.vw);
    const int width = 4 * t.vw;
    const int pair_cols = h2->scalar == KPMX_F64 ? t.vw : 0;
    const size_t per_step = (size_t)grid * pw * sizeof(double);
    int64_t chunk = std::max<int64_t>(2, partials_chunk(per_step, steps));
    chunk -= chunk % 2;
    double* partials = static_cast<double*>(ctx->partials.get(per_step * (size_t)chunk));
    int64_t k = 0, chunk_start = 0;
    for (; k + 1 < steps; k += 2) {
        S.xprev = k == 0 ? nullptr : r(k - 1);
        S.x = r(k);
        S.out1 = r(k + 1);
        S.out2 = r(k + 2);
        S.hscaleA = k == 0 ? 0.5 : 1.0;
        S.partialsA = partials + (size_t)(k - chunk_start) * grid * pw;
        S.partialsB = partials + (size_t)(k + 1 - chunk_start) * grid * pw;
        if (launch_step2(S, ctx->stream) != 0) throw Error{KPMX_EINVAL, "unsupported two-step configuration"};
        KPMX_HIP(hipGetLastError());
        if (k + 2 - chunk_start == chunk || k + 2 >= steps - 1) {
            launch_reduce_partials(partials, grid, pw, k + 2 - chunk_start, sums + (size_t)chunk_start * width,
                                   ctx->stream, pair_cols);
            KPMX_HIP(hipGetLastError());
            chunk_start = k + 2;
        }
    }
    if (k < steps) {  // an odd step count ends with one single step
        StepLaunch L = base_launch(h2, t.vw, true);
        L.mode = MODE_DIAG;
        L.tune = tuned_variant(h2, t.vw);
        L.num_wg = step_grid(L);
        L.x = r(k);
        L.yin = k == 0 ? nullptr : r(k - 1);
        L.yout = r(k + 1);
        L.hscale = k == 0 ? 0.5 : 1.0;
        L.sub_y = k == 0 ? 0 : 1;
        L.partials = partials;
        launch(L, ctx->stream);
        const bool pairs = step_uses_pairs(L);
        launch_reduce_partials(partials, L.num_wg, pairs ? 2 * t.vw : width, 1, sums + (size_t)k * width, ctx->stream,
                               pairs ? t.vw : 0);
        KPMX_HIP(hipGetLastError());
    }
}

void check_pair(const kpmx_mat* h2, const kpmx_block* r0) {
    KPMX_REQUIRE(h2 && r0, "null matrix or starter block");
    KPMX_REQUIRE(h2->ctx == r0->ctx, "matrix and starter block belong to different contexts");
    KPMX_REQUIRE(h2->scalar == r0->scalar, "matrix and starter block have different scalar types");
    KPMX_REQUIRE(h2->rows == r0->n, "Size mismatch between the matrix and the starter vectors");
}

// The column tiles the diagonal recurrence runs over: every starter tile no wider than `cap`
// as it is, wider ones as sub-tiles of `cap` columns, each copied out of its tile (into the
// context's work2 buffer) right before its own recurrence.  Narrower column tiles keep the
// gathered rows of one XCD's band of rows in its L2 and move less per row; the matrix is read
// once per sub-tile and step.
struct RunTile {
    size_t tile = 0;    // starter tile
    int c = 0;          // first column inside it
    bool copy = false;  // sub-tile: columns [c, c + sub.vw) copied to work2 first
    Tile sub;
};
std::vector<RunTile> run_tiles(const kpmx_block* r0, int cap) {
    std::vector<RunTile> out;
    for (size_t i = 0; i < r0->tiles.size(); ++i) {
        const Tile& t = r0->tiles[i];
        if (t.vw <= cap) {
            out.push_back({i, 0, false, t});
            continue;
        }
        for (int c = 0; c < t.cols; c += cap) {
            RunTile rt{i, c, true, t};
            rt.sub.c0 = t.c0 + c;
            rt.sub.cols = std::min<int64_t>(cap, t.cols - c);
            rt.sub.vw = cap;
            rt.sub.data = nullptr;  // work2, set when enqueued
            out.push_back(rt);
        }
    }
    return out;
}
int diag_cap(const kpmx_mat* h2) { return h2->ctx->tile_cap[h2->scalar]; }

// copies a sub-tile's columns into `dst` (default: work2, sized beforehand) and returns the tile
// to run
Tile stage_run_tile(const kpmx_mat* h2, const kpmx_block* r0, const RunTile& rt, void* dst = nullptr) {
    if (!rt.copy) return rt.sub;
    kpmx_ctx* ctx = h2->ctx;
    const Tile& t = r0->tiles[rt.tile];
    Tile sub = rt.sub;
    sub.data = dst ? dst : ctx->work2.get((size_t)h2->rows * sub.vw * scalar_bytes(h2->scalar));
    launch_copy_cols(h2->scalar, t.data, h2->rows, t.vw, rt.c, sub.vw, sub.data, ctx->stream);
    KPMX_HIP(hipGetLastError());
    return sub;
}

void diag_enqueue(const kpmx_mat* h2, const kpmx_block* r0, int64_t steps) {
    const auto rts = run_tiles(r0, r0->sums_cap);
    for (size_t i = 0; i < rts.size(); ++i) {
        const Tile sub = stage_run_tile(h2, r0, rts[i]);
        if (use_two_step(h2, sub))
            run_recurrence2(h2, sub, steps, static_cast<double*>(r0->sums[i]));
        else
            run_recurrence(h2, sub, MODE_DIAG, steps, nullptr, static_cast<double*>(r0->sums[i]),
                           [](StepLaunch&, int64_t) {}, [](StepLaunch const&, int64_t) {});
    }
}

// Small systems are launch-bound (a step of a 60k-site flake is a few microseconds of device
// work), so their whole recurrence is captured once and replayed as one graph.
bool diag_graph_candidate(const kpmx_mat* h2, const kpmx_block* r0) {
    if (!h2->step_rows.empty()) return false;  // per-call row limits are not part of the graph key
    const int policy = h2->ctx->graph_policy;
    if (policy == KPMX_GRAPH_NEVER) return false;
    if (policy == KPMX_GRAPH_ALWAYS) return true;
    return (size_t)r0->n * r0->nvec * scalar_bytes(h2->scalar) <= ((size_t)64 << 20);
}

void diag_run(const kpmx_mat* h2, const kpmx_block* r0c, int64_t M) {
    auto* r0 = const_cast<kpmx_block*>(r0c);
    check_pair(h2, r0);
    KPMX_REQUIRE(M >= 2 && M % 2 == 0, "num_moments must be even and >= 2");
    kpmx_ctx* ctx = h2->ctx;
    const int64_t steps = M / 2;
    const int cap = diag_cap(h2);
    const auto rts = run_tiles(r0, cap);
    if (r0->sums_moments != M || r0->sums_cap != cap) {
        if (!r0->sums.empty()) KPMX_HIP(hipStreamSynchronize(ctx->stream));
        r0->drop_graph();
        for (auto* s : r0->sums) dev_free(s);
        r0->sums.clear();
        for (auto& rt : rts) r0->sums.push_back(dev_alloc((size_t)steps * 4 * rt.sub.vw * sizeof(double)));
        r0->sums_moments = M;
        r0->sums_cap = cap;
    }
    if (!diag_graph_candidate(h2, r0)) {
        diag_enqueue(h2, r0, steps);
        return;
    }
    // Size the shared workspaces for every tile before capturing (nothing is allocated while the
    // stream captures) and key the graph on everything its launches read.
    const size_t sb = scalar_bytes(h2->scalar);
    size_t work = 0, part = 0, copy = 0;
    std::vector<uint64_t> key = {h2->serial, (uint64_t)M, (uint64_t)cap};
    for (const RunTile& rt : rts) {
        const Tile& t = rt.sub;
        work = std::max(work, (size_t)h2->rows * t.vw * sb);
        if (rt.copy) copy = std::max(copy, (size_t)h2->rows * t.vw * sb);
        StepLaunch L = base_launch(h2, t.vw, true);
        L.mode = MODE_DIAG;
        L.tune = tuned_variant(h2, t.vw);
        L.num_wg = step_grid(L);
        const size_t per_step_bytes = (size_t)L.num_wg * 4 * t.vw * sizeof(double);
        part = std::max(part, per_step_bytes * partials_chunk(per_step_bytes, steps));
        key.push_back((uint64_t)L.tune);
        if (use_two_step(h2, t)) {
            size_two_step(h2, t, steps);
            key.push_back((uint64_t)ctx->two_step[h2->scalar]);
            key.push_back((uint64_t)ctx->two_step_wg_per_cu[h2->scalar]);
            key.push_back((uint64_t)ctx->two_step_lds[h2->scalar]);
            key.push_back((uint64_t)(uintptr_t)ctx->work3.get(work));
            key.push_back((uint64_t)(uintptr_t)ctx->work4.get(work));
            part = std::max(part, ctx->partials.bytes);
        }
    }
    key.push_back((uint64_t)(uintptr_t)ctx->work0.get(work));
    key.push_back((uint64_t)(uintptr_t)ctx->work1.get(work));
    key.push_back((uint64_t)(uintptr_t)ctx->partials.get(part));
    if (copy) key.push_back((uint64_t)(uintptr_t)ctx->work2.get(copy));
    if (ctx->graph_policy == KPMX_GRAPH_AUTO && r0->graph_key != key && r0->last_key != key) {
        // automatic mode captures on the second run of the same recurrence: a one-off run
        // does not pay for capture and instantiation
        r0->last_key = key;
        diag_enqueue(h2, r0, steps);
        return;
    }
    if (!r0->graph_exec || r0->graph_key != key) {
        r0->drop_graph();
        KPMX_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        try {
            diag_enqueue(h2, r0, steps);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(ctx->stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        KPMX_HIP(hipStreamEndCapture(ctx->stream, &r0->graph));
        KPMX_HIP(hipGraphInstantiate(&r0->graph_exec, r0->graph, nullptr, nullptr, 0));
        r0->graph_key = key;
    }
    KPMX_HIP(hipGraphLaunch(r0->graph_exec, ctx->stream));
}

void diag_fetch(const kpmx_mat* h2, const kpmx_block* r0, int64_t M, void* mu) {
    check_pair(h2, r0);
    KPMX_REQUIRE(r0->sums_moments == M, "no finished run for this number of moments");
    kpmx_ctx* ctx = h2->ctx;
    const bool cplx = is_complex(h2->scalar);
    const int64_t steps = M / 2;
    const auto rts = run_tiles(r0, r0->sums_cap);
    for (size_t i = 0; i < rts.size(); ++i) {
        const Tile& t = rts[i].sub;
        std::vector<double> s((size_t)steps * 4 * t.vw);
        KPMX_HIP(hipMemcpyAsync(s.data(), r0->sums[i], s.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        KPMX_HIP(hipStreamSynchronize(ctx->stream));
        if (h2->scalar == 0 || h2->scalar == 2)
            diag_collect<float>(s, t.vw, t.cols, M, cplx, static_cast<float*>(mu), t.c0);
        else
            diag_collect<double>(s, t.vw, t.cols, M, cplx, static_cast<double*>(mu), t.c0);
    }
}

// op^H beta on the host: w_j = sum_i conj(op_ij) beta_i
template <class R>
void op_adjoint_apply(int64_t n, bool cplx, int64_t nnz, const int32_t* indptr, const int32_t* indices,
                      const R* vals, const R* beta, std::vector<R>& w) {
    const int c = cplx ? 2 : 1;
    w.assign((size_t)n * c, R(0));
    if (nnz == 0 || !vals) {
        std::copy(beta, beta + n * c, w.begin());
        return;
    }
    for (int64_t i = 0; i < n; ++i) {
        for (int32_t k = indptr[i]; k < indptr[i + 1]; ++k) {
            const int64_t j = indices[k];
            if (cplx) {
                const R ar = vals[2 * k], ai = -vals[2 * k + 1];
                const R br = beta[2 * i], bi = beta[2 * i + 1];
                w[2 * j] += ar * br - ai * bi;
                w[2 * j + 1] += ar * bi + ai * br;
            } else {
                w[j] += vals[k] * beta[i];
            }
        }
    }
}


// Extreme eigenvalues of the symmetric tridiagonal matrix (diag a[0..k), off-diagonal b[0..k-1))
// by Sturm-sequence bisection (Golub & Van Loan, sec. 8.4.1) started from the Gershgorin interval.
int sturm_count(const std::vector<double>& a, const std::vector<double>& b, size_t k, double x) {
    int cnt = 0;
    double d = 1.0;
    for (size_t i = 0; i < k; ++i) {
        const double off = (i > 0) ? b[i - 1] * b[i - 1] : 0.0;
        d = a[i] - x - ((i > 0) ? off / d : 0.0);
        if (d == 0.0) d = -1e-300;
        if (d < 0) ++cnt;
    }
    return cnt;
}

void tridiagonal_minmax(const std::vector<double>& a, const std::vector<double>& b, size_t k, double* mn, double* mx) {
    double lo = a[0], hi = a[0];
    for (size_t i = 0; i < k; ++i) {
        const double r = ((i > 0) ? std::abs(b[i - 1]) : 0.0) + ((i + 1 < k) ? std::abs(b[i]) : 0.0);
        lo = std::min(lo, a[i] - r);
        hi = std::max(hi, a[i] + r);
    }
    const double span = std::max(hi - lo, 1e-300);
    // smallest eigenvalue: smallest x with count(x) >= 1
    double l = lo - 1e-12 * span, h = hi + 1e-12 * span;
    for (int it = 0; it < 200 && h - l > 0; ++it) {
        const double m = 0.5 * (l + h);
        if (m <= l || m >= h) break;
        if (sturm_count(a, b, k, m) >= 1) h = m; else l = m;
    }
    *mn = 0.5 * (l + h);
    // largest eigenvalue: largest x with count(x) <= k - 1
    l = lo - 1e-12 * span;
    h = hi + 1e-12 * span;
    for (int it = 0; it < 200 && h - l > 0; ++it) {
        const double m = 0.5 * (l + h);
        if (m <= l || m >= h) break;
        if (sturm_count(a, b, k, m) <= (int)k - 1) l = m; else h = m;
    }
    *mx = 0.5 * (l + h);
}

}  // namespace

// ==================================================================================== C-ABI
namespace {
// Walker w of the batched step (one per wave, BATCH_WALKERS_PER_WG per workgroup) walks its
// slices along BatchWalk (chunk 0: the contiguous slices [tile_begin[w], tile_begin[w+1]);
// chunk > 0: the XCD-aware round-robin runs) and meets each system in one stretch.  Replaying the
// walks numbers the stretches (walker w's first is slot_begin[w]) and gives each its partial
// slot, slot_index[stretch]: the slots of one system are consecutive, in walker order, from
// slot_first[s].  Rebuilding the partition drops the captured graph and the partial-sum buffer
// (both depend on it).
void batch_partition(kpmx_batch* b, int num_wg, int chunk) {
    kpmx_ctx* ctx = b->ctx;
    const int64_t ns_tot = b->nslices;
    constexpr int PER = BATCH_WALKERS_PER_WG;
    if (chunk > 0) {
        num_wg = std::max(8, num_wg / 8 * 8);
    } else {
        num_wg = (int)std::max<int64_t>(1, std::min<int64_t>((ns_tot + PER - 1) / PER, num_wg));
    }
    const int walkers = num_wg * PER;
    std::vector<int64_t> tb(walkers + 1, 0), sbeg(walkers);
    if (chunk <= 0)
        for (int w = 0; w <= walkers; ++w) tb[w] = ns_tot * w / walkers;
    std::vector<int32_t> stretch_sys;  // system of every stretch, in walker order
    const auto& ssys = b->slice_sys_host;
    for (int w = 0; w < walkers; ++w) {
        sbeg[w] = (int64_t)stretch_sys.size();
        int prev = -1;
        for (BatchWalk walk(ns_tot, walkers, w, chunk, tb.data()); walk.valid(); walk.next()) {
            if (ssys[walk.s] != prev) stretch_sys.push_back(ssys[walk.s]);
            prev = ssys[walk.s];
        }
    }
    const int64_t num_slots = (int64_t)stretch_sys.size();
    // counting sort of the stretches by system (stable: walker order inside a system)
    std::vector<int64_t> sfirst(b->num_sys + 1, 0), sidx(std::max<int64_t>(num_slots, 1));
    for (int32_t sy : stretch_sys) ++sfirst[sy + 1];
    for (int64_t sy = 0; sy < b->num_sys; ++sy) sfirst[sy + 1] += sfirst[sy];
    {
        std::vector<int64_t> fill(sfirst.begin(), sfirst.end() - 1);
        for (int64_t p = 0; p < num_slots; ++p) sidx[p] = fill[stretch_sys[p]]++;
    }
    KPMX_HIP(hipStreamSynchronize(ctx->stream));
    b->drop_graph();
    for (void** p : {(void**)&b->tile_begin, (void**)&b->slot_begin, (void**)&b->slot_index, (void**)&b->slot_first,
                     (void**)&b->partials, (void**)&b->partials2}) {
        dev_free(*p);
        *p = nullptr;
    }
    b->num_wg = num_wg;
    b->walk_chunk = chunk;
    b->num_slots = num_slots;
    b->tile_begin = static_cast<int64_t*>(dev_alloc(tb.size() * sizeof(int64_t)));
    b->slot_begin = static_cast<int64_t*>(dev_alloc(sbeg.size() * sizeof(int64_t)));
    b->slot_index = static_cast<int64_t*>(dev_alloc(sidx.size() * sizeof(int64_t)));
    b->slot_first = static_cast<int64_t*>(dev_alloc(sfirst.size() * sizeof(int64_t)));
    KPMX_HIP(hipMemcpyAsync(b->tile_begin, tb.data(), tb.size() * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream));
    KPMX_HIP(hipMemcpyAsync(b->slot_begin, sbeg.data(), sbeg.size() * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream));
    KPMX_HIP(hipMemcpyAsync(b->slot_index, sidx.data(), sidx.size() * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream));
    KPMX_HIP(hipMemcpyAsync(b->slot_first, sfirst.data(), sfirst.size() * sizeof(int64_t), hipMemcpyHostToDevice,
                            ctx->stream));
    KPMX_HIP(hipStreamSynchronize(ctx->stream));
}

// Upload pieces into one workspace, each at a 256-byte aligned offset; returns their device addresses.
std::vector<char*> upload_pieces(kpmx_ctx* ctx, Workspace& ws, std::initializer_list<std::pair<const void*, size_t>> parts) {
    size_t total = 0;
    for (auto& p : parts) total += (p.second + 255) / 256 * 256;
    char* base = static_cast<char*>(ws.get(std::max<size_t>(total, 256)));
    std::vector<char*> out;
    size_t off = 0;
    for (auto& p : parts) {
        out.push_back(base + off);
        if (p.first && p.second)  // a null source reserves space only
            KPMX_HIP(hipMemcpyAsync(base + off, p.first, p.second, hipMemcpyHostToDevice, ctx->stream));
        off += (p.second + 255) / 256 * 256;
    }
    return out;
}

void cheb_series(kpmx_ctx* ctx, int real, bool exp_series, int64_t M, int64_t cols, const void* mu, int64_t E,
                 const void* theta, const void* pref, double* out) {
    KPMX_REQUIRE(ctx && mu && theta && pref && out, "null argument");
    KPMX_REQUIRE(real == KPMX_F32 || real == KPMX_F64, "real must be KPMX_F32 or KPMX_F64");
    KPMX_REQUIRE(M >= 1 && cols >= 1 && E >= 0, "empty series");
    KPMX_HIP(hipSetDevice(ctx->device));
    if (E == 0) return;
    const size_t rb = real == KPMX_F32 ? 4 : 8, cw = exp_series ? 2 : 1;
    auto d = upload_pieces(ctx, ctx->aux0,
                           {{mu, (size_t)(cols * M) * cw * rb}, {theta, (size_t)E * rb}, {pref, (size_t)E * cw * rb}});
    double* dout = static_cast<double*>(ctx->aux1.get((size_t)(E * cols) * cw * sizeof(double)));
    if (launch_cheb_series(real, exp_series, d[0], M, cols, E, d[1], d[2], dout, ctx->stream) != 0)
        throw Error{KPMX_EINVAL, "too many energies or columns for one call"};
    KPMX_HIP(hipGetLastError());
    KPMX_HIP(hipMemcpyAsync(out, dout, (size_t)(E * cols) * cw * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    KPMX_HIP(hipStreamSynchronize(ctx->stream));
}
}  // namespace

extern "C" {

const char* kpmx_last_error(void) { return g_last_error.c_str(); }

int kpmx_version(void) { return 101; }

int kpmx_device_count(int* count) {
    return guarded([&] {
        KPMX_REQUIRE(count, "null count");
        int c = 0;
        hipError_t e = hipGetDeviceCount(&c);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            c = 0;
        }
        *count = c;
    });
}

int kpmx_ctx_create(int device, kpmx_ctx** out) {
    return guarded([&] {
        KPMX_REQUIRE(out, "null output");
        int count = 0;
        KPMX_HIP(hipGetDeviceCount(&count));
        KPMX_REQUIRE(device >= 0 && device < count, "invalid device index");
        auto* ctx = new kpmx_ctx();
        ctx->device = device;
        try {
            KPMX_HIP(hipSetDevice(device));
            hipDeviceProp_t prop;
            KPMX_HIP(hipGetDeviceProperties(&prop, device));
            ctx->num_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
            KPMX_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
            ctx->scratch = static_cast<char*>(scratch_address(device));
            // default step variants: fp64 blocks of 16 or more vectors as column pairs (the tuned
            // choice on the 4.19 M-site 64-vector and 1 M-site 16-vector workloads)
            const int pairs = 2 * NUM_TILE_SHAPES * NUM_GRID_VARIANTS;
            ctx->tune[KPMX_F64][4] = pairs + 1;                     // 16 vectors: tile shape 1, resident grid
            ctx->tune[KPMX_F64][5] = pairs + 1;                     // 32 vectors
            ctx->tune[KPMX_F64][6] = pairs + 1 + NUM_TILE_SHAPES;   // 64 vectors: tile shape 1, half grid
            // on matrices with the value-table form: its column-pair variant with tile shape 2 on
            // the resident grid (16 columns: 1.207 ms per 64-vector step against 1.270 ms)
            ctx->tune_coded[KPMX_F64][4] = pairs + NUM_TILE_SHAPES * NUM_GRID_VARIANTS + 2;
            ctx->tile_cap[KPMX_F64] = default_tile_cap(KPMX_F64);
        } catch (...) {
            delete ctx;
            throw;
        }
        *out = ctx;
    });
}

int kpmx_ctx_destroy(kpmx_ctx* ctx) {
    return guarded([&] {
        if (!ctx) return;
        (void)hipSetDevice(ctx->device);
        if (ctx->stream) {
            (void)hipStreamSynchronize(ctx->stream);
            (void)hipStreamDestroy(ctx->stream);
        }
        if (ctx->t0) (void)hipEventDestroy(ctx->t0);
        if (ctx->t1) (void)hipEventDestroy(ctx->t1);
        delete ctx;
    });
}

int kpmx_ctx_synchronize(kpmx_ctx* ctx) {
    return guarded([&] {
        KPMX_REQUIRE(ctx, "null context");
        KPMX_HIP(hipStreamSynchronize(ctx->stream));
    });
}

namespace {
// The value-table form of a SELL matrix (see kpmx_mat): built when every row holds at most one
// diagonal entry and the off-diagonal entries take at most 256 distinct values (compared bit for
// bit; the value 0 of padded slots included).  Off-diagonal entries keep their CSR order.
void build_coded_sell(kpmx_mat* m, const int32_t* indptr, const int32_t* indices, const char* v, size_t sb) {
    const int64_t rows = m->rows;
    if (rows == 0 || m->nnz == 0) return;
    std::vector<std::string> table;
    std::map<std::string, int> code_of;
    const std::string zero(sb, '\0');
    auto code = [&](const char* p) -> int {
        const std::string key(p, sb);
        auto it = code_of.find(key);
        if (it != code_of.end()) return it->second;
        if (table.size() >= 256) return -1;
        code_of[key] = (int)table.size();
        table.push_back(key);
        return (int)table.size() - 1;
    };
    const int zcode = code(zero.data());
    std::vector<char> diag((size_t)rows * sb, 0);
    std::vector<int32_t> offcount(rows, 0);
    for (int64_t r = 0; r < rows; ++r) {
        int ndiag = 0;
        for (int32_t k = indptr[r]; k < indptr[r + 1]; ++k) {
            if (indices[k] == r) {
                if (++ndiag > 1) return;
                std::memcpy(&diag[(size_t)r * sb], v + (size_t)k * sb, sb);
            } else {
                if (code(v + (size_t)k * sb) < 0) return;
                ++offcount[r];
            }
        }
    }
    const int64_t ns = (rows + 63) / 64;
    std::vector<int64_t> sptr(ns + 1, 0);
    std::vector<int32_t> swidth(ns, 0);
    for (int64_t sl = 0; sl < ns; ++sl) {
        int32_t w = 0;
        for (int64_t r = sl * 64; r < std::min<int64_t>(rows, (sl + 1) * 64); ++r) w = std::max(w, offcount[r]);
        swidth[sl] = w;
        sptr[sl + 1] = sptr[sl] + (int64_t)w * 64;
    }
    const int64_t total = sptr[ns];
    std::vector<int32_t> cidx((size_t)std::max<int64_t>(total, 1));
    std::vector<uint8_t> ccode((size_t)std::max<int64_t>(total, 1), (uint8_t)zcode);
    for (int64_t sl = 0; sl < ns; ++sl)
        for (int lane = 0; lane < 64; ++lane) {
            const int64_t r = sl * 64 + lane;
            int32_t k2 = 0;
            if (r < rows)
                for (int32_t k = indptr[r]; k < indptr[r + 1]; ++k) {
                    if (indices[k] == r) continue;
                    const int64_t dst = sptr[sl] + (int64_t)k2 * 64 + lane;
                    cidx[(size_t)dst] = indices[k];
                    ccode[(size_t)dst] = (uint8_t)code(v + (size_t)k * sb);
                    ++k2;
                }
            for (; k2 < swidth[sl]; ++k2) {
                const int64_t dst = sptr[sl] + (int64_t)k2 * 64 + lane;
                cidx[(size_t)dst] = (int32_t)std::min<int64_t>(r, rows - 1);
                ccode[(size_t)dst] = (uint8_t)zcode;
            }
        }
    std::vector<char> tab(table.size() * sb);
    for (size_t i = 0; i < table.size(); ++i) std::memcpy(&tab[i * sb], table[i].data(), sb);
    kpmx_ctx* ctx = m->ctx;
    m->cslice_ptr = static_cast<int64_t*>(dev_alloc((ns + 1) * sizeof(int64_t)));
    m->cslice_width = static_cast<int32_t*>(dev_alloc(ns * sizeof(int32_t)));
    m->cindices = static_cast<int32_t*>(dev_alloc(cidx.size() * sizeof(int32_t)));
    m->ccodes = static_cast<uint8_t*>(dev_alloc(ccode.size()));
    m->cdiag = dev_alloc(diag.size());
    m->ctable = dev_alloc(tab.size());
    KPMX_HIP(hipMemcpyAsync(m->cslice_ptr, sptr.data(), sptr.size() * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream));
    KPMX_HIP(hipMemcpyAsync(m->cslice_width, swidth.data(), swidth.size() * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    KPMX_HIP(hipMemcpyAsync(m->cindices, cidx.data(), cidx.size() * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    KPMX_HIP(hipMemcpyAsync(m->ccodes, ccode.data(), ccode.size(), hipMemcpyHostToDevice, ctx->stream));
    KPMX_HIP(hipMemcpyAsync(m->cdiag, diag.data(), diag.size(), hipMemcpyHostToDevice, ctx->stream));
    KPMX_HIP(hipMemcpyAsync(m->ctable, tab.data(), tab.size(), hipMemcpyHostToDevice, ctx->stream));
    KPMX_HIP(hipStreamSynchronize(ctx->stream));
    m->cmax_width = ns ? *std::max_element(swidth.begin(), swidth.end()) : 0;
    m->ctable_size = (int)table.size();
}
}  // namespace

int kpmx_mat_upload(kpmx_ctx* ctx, int scalar, int format, int64_t rows, int64_t nnz, const int32_t* indptr,
                    const int32_t* indices, const void* values, kpmx_mat** out) {
    return guarded([&] {
        KPMX_REQUIRE(ctx && out, "null context or output");
        KPMX_REQUIRE(valid_scalar(scalar), "invalid scalar code");
        KPMX_REQUIRE(format == KPMX_FMT_CSR || format == KPMX_FMT_SELL, "invalid matrix format");
        KPMX_REQUIRE(rows >= 0 && nnz >= 0 && rows < (int64_t)1 << 31 && nnz < (int64_t)1 << 31,
                     "matrix too large for 32-bit indices");
        KPMX_REQUIRE(indptr && (nnz == 0 || (indices && values)), "null CSR arrays");
        KPMX_REQUIRE(indptr[0] == 0 && indptr[rows] == nnz, "inconsistent CSR row pointers");
        for (int64_t i = 0; i < rows; ++i) KPMX_REQUIRE(indptr[i] <= indptr[i + 1], "decreasing CSR row pointers");
        for (int64_t k = 0; k < nnz; ++k) KPMX_REQUIRE(indices[k] >= 0 && indices[k] < rows, "CSR column index out of range");
        KPMX_HIP(hipSetDevice(ctx->device));
        const size_t sb = scalar_bytes(scalar);
        auto m = std::make_unique<kpmx_mat>();
        m->ctx = ctx;
        static std::atomic<uint64_t> next_serial{1};
        m->serial = next_serial++;
        m->scalar = scalar;
        m->format = format;
        m->rows = rows;
        m->nnz = nnz;
        m->indptr = static_cast<int32_t*>(dev_alloc((rows + 1) * sizeof(int32_t)));
        m->indices = static_cast<int32_t*>(dev_alloc(nnz * sizeof(int32_t)));
        m->vals = dev_alloc(nnz * sb);
        KPMX_HIP(hipMemcpyAsync(m->indptr, indptr, (rows + 1) * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
        if (nnz) {
            KPMX_HIP(hipMemcpyAsync(m->indices, indices, nnz * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
            KPMX_HIP(hipMemcpyAsync(m->vals, values, nnz * sb, hipMemcpyHostToDevice, ctx->stream));
        }
        if (format == KPMX_FMT_SELL) {
            // Slices of 64 rows; inside a slice, column-major with the slice's max row length.
            // Padding: value 0, column index = the row itself (a valid gather of data the row
            // touches anyway).
            const int64_t ns = (rows + 63) / 64;
            std::vector<int64_t> sptr(ns + 1, 0);
            std::vector<int32_t> swidth(ns, 0);
            for (int64_t s = 0; s < ns; ++s) {
                int32_t w = 0;
                for (int64_t r = s * 64; r < std::min<int64_t>(rows, (s + 1) * 64); ++r)
                    w = std::max<int32_t>(w, indptr[r + 1] - indptr[r]);
                swidth[s] = w;
                sptr[s + 1] = sptr[s] + (int64_t)w * 64;
            }
            const int64_t total = sptr[ns];
            std::vector<int32_t> sidx((size_t)total);
            std::vector<char> svals((size_t)total * sb, 0);
            const char* v = static_cast<const char*>(values);
            for (int64_t s = 0; s < ns; ++s) {
                for (int lane = 0; lane < 64; ++lane) {
                    const int64_t r = s * 64 + lane;
                    for (int32_t k = 0; k < swidth[s]; ++k) {
                        const int64_t dst = sptr[s] + (int64_t)k * 64 + lane;
                        if (r < rows && indptr[r] + k < indptr[r + 1]) {
                            sidx[dst] = indices[indptr[r] + k];
                            std::memcpy(&svals[dst * sb], v + (size_t)(indptr[r] + k) * sb, sb);
                        } else {
                            sidx[dst] = (int32_t)std::min<int64_t>(r, rows - 1);
                        }
                    }
                }
            }
            m->sell_pos.assign((size_t)nnz, 0);
            for (int64_t r = 0; r < rows; ++r)
                for (int32_t k = indptr[r]; k < indptr[r + 1]; ++k)
                    m->sell_pos[(size_t)k] = sptr[r / 64] + (int64_t)(k - indptr[r]) * 64 + (r % 64);
            m->sell_total = total;
            m->nslices = ns;
            m->max_width = ns ? *std::max_element(swidth.begin(), swidth.end()) : 0;
            m->slice_ptr = static_cast<int64_t*>(dev_alloc((ns + 1) * sizeof(int64_t)));
            m->slice_width = static_cast<int32_t*>(dev_alloc(ns * sizeof(int32_t)));
            m->sell_indices = static_cast<int32_t*>(dev_alloc(total * sizeof(int32_t)));
            m->sell_vals = dev_alloc(total * sb);
            KPMX_HIP(hipMemcpyAsync(m->slice_ptr, sptr.data(), (ns + 1) * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream));
            KPMX_HIP(hipMemcpyAsync(m->slice_width, swidth.data(), ns * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
            if (total) {
                KPMX_HIP(hipMemcpyAsync(m->sell_indices, sidx.data(), total * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
                KPMX_HIP(hipMemcpyAsync(m->sell_vals, svals.data(), total * sb, hipMemcpyHostToDevice, ctx->stream));
            }
            KPMX_HIP(hipStreamSynchronize(ctx->stream));
            build_coded_sell(m.get(), indptr, indices, v, sb);
        }
        KPMX_HIP(hipStreamSynchronize(ctx->stream));
        *out = m.release();
    });
}

int kpmx_mat_coded_bytes(const kpmx_mat* m, int64_t* bytes) {
    return guarded([&] {
        KPMX_REQUIRE(m && bytes, "null argument");
        *bytes = 0;
        if (!m->ccodes) return;
        const int64_t ns = m->nslices, sb = (int64_t)scalar_bytes(m->scalar);
        std::vector<int64_t> cptr(2);
        KPMX_HIP(hipMemcpy(cptr.data(), m->cslice_ptr + ns, sizeof(int64_t), hipMemcpyDeviceToHost));
        const int64_t entries = cptr[0];
        *bytes = (ns + 1) * 8 + ns * 4 + entries * (4 + 1) + m->rows * sb + (int64_t)m->ctable_size * sb;
    });
}

int kpmx_mat_destroy(kpmx_mat* m) {
    return guarded([&] {
        if (!m) return;
        (void)hipSetDevice(m->ctx->device);
        (void)hipStreamSynchronize(m->ctx->stream);
        delete m;
    });
}

int kpmx_mat_set_step_rows(kpmx_mat* h2, int64_t num_steps, const int64_t* rows) {
    return guarded([&] {
        KPMX_REQUIRE(h2, "null matrix");
        KPMX_REQUIRE(num_steps >= 0 && (num_steps == 0 || rows), "null step rows");
        for (int64_t k = 0; k < num_steps; ++k)
            KPMX_REQUIRE(rows[k] >= 1 && rows[k] <= h2->rows, "step rows must lie in [1, matrix rows]");
        h2->step_rows.assign(rows, rows + num_steps);
    });
}

int kpmx_block_upload(kpmx_ctx* ctx, int scalar, int64_t n, int64_t num_vectors, const void* r0, kpmx_block** out) {
    return guarded([&] {
        KPMX_REQUIRE(ctx && out, "null context or output");
        KPMX_REQUIRE(valid_scalar(scalar), "invalid scalar code");
        KPMX_REQUIRE(n > 0 && num_vectors > 0 && r0, "empty starter block");
        KPMX_HIP(hipSetDevice(ctx->device));
        const size_t sb = scalar_bytes(scalar);
        auto b = std::make_unique<kpmx_block>();
        b->ctx = ctx;
        b->scalar = scalar;
        b->n = n;
        b->nvec = num_vectors;
        void* staging = ctx->aux0.get((size_t)n * num_vectors * sb);
        KPMX_HIP(hipMemcpyAsync(staging, r0, (size_t)n * num_vectors * sb, hipMemcpyHostToDevice, ctx->stream));
        for (int64_t c0 = 0; c0 < num_vectors;) {
            Tile t;
            t.c0 = c0;
            t.vw = pick_vw(num_vectors - c0);
            t.cols = std::min<int64_t>(t.vw, num_vectors - c0);
            t.data = dev_alloc((size_t)n * t.vw * sb);
            b->tiles.push_back(t);
            launch_gather_block(scalar, staging, n, t.c0, t.cols, t.vw, t.data, ctx->stream);
            KPMX_HIP(hipGetLastError());
            c0 += t.cols;
        }
        KPMX_HIP(hipStreamSynchronize(ctx->stream));
        *out = b.release();
    });
}

int kpmx_block_destroy(kpmx_block* b) {
    return guarded([&] {
        if (!b) return;
        (void)hipSetDevice(b->ctx->device);
        (void)hipStreamSynchronize(b->ctx->stream);
        delete b;
    });
}

int kpmx_ctx_set_tile_width(kpmx_ctx* ctx, int scalar, int width) {
    return guarded([&] {
        KPMX_REQUIRE(ctx, "null context");
        KPMX_REQUIRE(valid_scalar(scalar), "invalid scalar code");
        KPMX_REQUIRE(width >= 1 && width <= 64 && (width & (width - 1)) == 0, "tile width must be 1, 2, 4, ..., 64");
        ctx->tile_cap[scalar] = width;
        ctx->two_step_auto[scalar] = false;
    });
}

int kpmx_ctx_get_tile_width(kpmx_ctx* ctx, int scalar, int* width) {
    return guarded([&] {
        KPMX_REQUIRE(ctx && width, "null argument");
        KPMX_REQUIRE(valid_scalar(scalar), "invalid scalar code");
        *width = ctx->tile_cap[scalar];
    });
}

int kpmx_ctx_set_two_step(kpmx_ctx* ctx, int scalar, int tile_rows, int workgroups_per_cu) {
    return guarded([&] {
        KPMX_REQUIRE(ctx, "null context");
        KPMX_REQUIRE(valid_scalar(scalar), "invalid scalar code");
        KPMX_REQUIRE(tile_rows == 0 || (tile_rows >= 64 && tile_rows % 64 == 0 && tile_rows <= (1 << 20)),
                     "two-step tile rows must be 0 or a multiple of 64 up to 1048576");
        KPMX_REQUIRE(workgroups_per_cu >= 0 && workgroups_per_cu <= 64, "workgroups per CU must lie in [0, 64]");
        ctx->two_step[scalar] = tile_rows;
        ctx->two_step_auto[scalar] = false;
        ctx->two_step_wg_per_cu[scalar] = workgroups_per_cu;
        ctx->two_step_lds[scalar] = 0;
    });
}

int kpmx_ctx_get_two_step_lds(kpmx_ctx* ctx, int scalar, int* enabled, int* workgroups_per_cu) {
    return guarded([&] {
        KPMX_REQUIRE(ctx, "null context");
        KPMX_REQUIRE(valid_scalar(scalar), "invalid scalar code");
        if (enabled) *enabled = ctx->two_step_lds[scalar] && ctx->two_step[scalar] > 0;
        if (workgroups_per_cu) *workgroups_per_cu = ctx->two_step_wg_per_cu[scalar];
    });
}

int kpmx_ctx_set_autotune_two_step(kpmx_ctx* ctx, int enable) {
    return guarded([&] {
        KPMX_REQUIRE(ctx, "null context");
        ctx->autotune_two_step = enable != 0;
    });
}

int kpmx_ctx_set_two_step_lds(kpmx_ctx* ctx, int scalar, int workgroups_per_cu) {
    return guarded([&] {
        KPMX_REQUIRE(ctx, "null context");
        KPMX_REQUIRE(scalar == KPMX_F64 || scalar == KPMX_C128, "the LDS two-step form is fp64 or complex128 only");
        KPMX_REQUIRE(workgroups_per_cu >= 0 && workgroups_per_cu <= 64, "workgroups per CU must lie in [0, 64]");
        ctx->two_step[scalar] = STEP2_LDS_TILE;
        ctx->two_step_auto[scalar] = false;
        ctx->two_step_wg_per_cu[scalar] = workgroups_per_cu;
        ctx->two_step_lds[scalar] = 1;
    });
}

int kpmx_ctx_get_two_step(kpmx_ctx* ctx, int scalar, int* tile_rows) {
    return guarded([&] {
        KPMX_REQUIRE(ctx && tile_rows, "null argument");
        KPMX_REQUIRE(valid_scalar(scalar), "invalid scalar code");
        *tile_rows = ctx->two_step[scalar];
    });
}

int kpmx_ctx_set_graph_policy(kpmx_ctx* ctx, int policy) {
    return guarded([&] {
        KPMX_REQUIRE(ctx, "null context");
        KPMX_REQUIRE(policy == KPMX_GRAPH_NEVER || policy == KPMX_GRAPH_AUTO || policy == KPMX_GRAPH_ALWAYS,
                     "invalid graph policy");
        ctx->graph_policy = policy;
    });
}

int kpmx_moments_diagonal_run(const kpmx_mat* h2, const kpmx_block* r0, int64_t num_moments) {
    return guarded([&] {
        KPMX_REQUIRE(h2 && r0, "null matrix or starter block");
        KPMX_HIP(hipSetDevice(h2->ctx->device));
        diag_run(h2, r0, num_moments);
    });
}

int kpmx_moments_diagonal_fetch(const kpmx_mat* h2, const kpmx_block* r0, int64_t num_moments, void* mu) {
    return guarded([&] {
        KPMX_REQUIRE(h2 && r0 && mu, "null argument");
        KPMX_HIP(hipSetDevice(h2->ctx->device));
        diag_fetch(h2, r0, num_moments, mu);
    });
}

int kpmx_moments_diagonal(const kpmx_mat* h2, const kpmx_block* r0, int64_t num_moments, void* mu) {
    return guarded([&] {
        KPMX_REQUIRE(h2 && r0 && mu, "null argument");
        KPMX_HIP(hipSetDevice(h2->ctx->device));
        diag_run(h2, r0, num_moments);
        diag_fetch(h2, r0, num_moments, mu);
    });
}

// kpmx_moments_diagonal without device allocation: the work vectors, partial sums, raw sums and
// collected moments of the recurrence live in this device's `scratch` array (layout below, each
// piece 256-byte aligned); nothing in it needs clearing (every launch overwrites what it reads
// later).  Concurrent calls on one device share the array and race.
int kpmx_moments_diagonal_noalloc(const kpmx_mat* h2, const kpmx_block* r0, int64_t num_moments, void* mu) {
    return guarded([&] {
        KPMX_REQUIRE(h2 && r0 && mu, "null argument");
        check_pair(h2, r0);
        const int64_t M = num_moments;
        KPMX_REQUIRE(M >= 2 && M % 2 == 0, "num_moments must be even and >= 2");
        kpmx_ctx* ctx = h2->ctx;
        KPMX_HIP(hipSetDevice(ctx->device));
        KPMX_REQUIRE(ctx->scratch != nullptr, "no scratch array on this device");
        const int64_t steps = M / 2;
        const size_t sb = scalar_bytes(h2->scalar);
        auto align = [](size_t b) { return (b + 255) / 256 * 256; };
        // the same column tiles as kpmx_moments_diagonal (sub-tiles of wide starter tiles)
        const auto rts = run_tiles(r0, diag_cap(h2));
        int vw_max = 1;
        bool any_copy = false;
        size_t part_step = 0;
        for (const RunTile& rt : rts) {
            const Tile& t = rt.sub;
            vw_max = std::max(vw_max, t.vw);
            any_copy = any_copy || rt.copy;
            StepLaunch L = base_launch(h2, t.vw, true);
            L.mode = MODE_DIAG;
            L.tune = tuned_variant(h2, t.vw);
            L.num_wg = step_grid(L);
            part_step = std::max(part_step, (size_t)L.num_wg * 4 * t.vw * sizeof(double));
        }
        // layout: [W0 | W1 | sub-tile copy (if any) | raw sums (steps x 4 vw_max) | moments (nvec x M) |
        //          partial sums (rest)]
        const size_t work = align((size_t)h2->rows * vw_max * sb);
        const size_t copy = any_copy ? work : 0;
        const size_t sums = align((size_t)steps * 4 * vw_max * sizeof(double));
        const size_t out = align((size_t)r0->nvec * M * sb);
        const size_t fixed_part = 2 * work + copy + sums + out;
        if (fixed_part + part_step > KPMX_SCRATCH_BYTES)
            throw Error{KPMX_EINVAL, "kpmx_moments_diagonal_noalloc: this matrix, block and moment count need " +
                                         std::to_string(fixed_part + part_step) + " bytes of device scratch; the limit is " +
                                         std::to_string(KPMX_SCRATCH_BYTES) + " bytes"};
        FixedBuffers fb;
        fb.w0 = ctx->scratch;
        fb.w1 = ctx->scratch + work;
        char* d_copy = ctx->scratch + 2 * work;
        double* d_sums = reinterpret_cast<double*>(ctx->scratch + 2 * work + copy);
        char* d_mu = ctx->scratch + 2 * work + copy + sums;
        fb.partials = reinterpret_cast<double*>(ctx->scratch + fixed_part);
        fb.partials_bytes = KPMX_SCRATCH_BYTES - fixed_part;
        for (const RunTile& rt : rts) {
            const Tile t = stage_run_tile(h2, r0, rt, d_copy);
            run_recurrence(h2, t, MODE_DIAG, steps, nullptr, d_sums, [](StepLaunch&, int64_t) {},
                           [](StepLaunch const&, int64_t) {}, -1, &fb);
            launch_diag_collect(h2->scalar, d_sums, t.vw, t.cols, M, t.c0, d_mu, ctx->stream);
            KPMX_HIP(hipGetLastError());
        }
        KPMX_HIP(hipMemcpyAsync(mu, d_mu, (size_t)r0->nvec * M * sb, hipMemcpyDeviceToHost, ctx->stream));
        KPMX_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int kpmx_moments_generic(const kpmx_mat* h2, const kpmx_block* r0, int64_t M, const void* beta, int64_t op_nnz,
                         const int32_t* op_indptr, const int32_t* op_indices, const void* op_values, void* mu) {
    return guarded([&] {
        check_pair(h2, r0);
        KPMX_REQUIRE(beta && mu, "null beta or output");
        KPMX_REQUIRE(M >= 2, "num_moments must be >= 2");
        KPMX_REQUIRE(op_nnz == 0 || (op_indptr && op_indices && op_values), "null operator arrays");
        kpmx_ctx* ctx = h2->ctx;
        KPMX_HIP(hipSetDevice(ctx->device));
        const int64_t n = h2->rows;
        const bool cplx = is_complex(h2->scalar);
        const size_t sb = scalar_bytes(h2->scalar);
        // w = op^H beta (host), uploaded once
        std::vector<char> wbuf((size_t)n * sb);
        if (h2->scalar == 0 || h2->scalar == 2) {
            std::vector<float> w;
            op_adjoint_apply<float>(n, cplx, op_nnz, op_indptr, op_indices, static_cast<const float*>(op_values),
                                    static_cast<const float*>(beta), w);
            std::memcpy(wbuf.data(), w.data(), wbuf.size());
        } else {
            std::vector<double> w;
            op_adjoint_apply<double>(n, cplx, op_nnz, op_indptr, op_indices, static_cast<const double*>(op_values),
                                     static_cast<const double*>(beta), w);
            std::memcpy(wbuf.data(), w.data(), wbuf.size());
        }
        void* dw = ctx->aux1.get(wbuf.size());
        KPMX_HIP(hipMemcpyAsync(dw, wbuf.data(), wbuf.size(), hipMemcpyHostToDevice, ctx->stream));
        // one-vector tile view of the first starter
        const Tile& first = r0->tiles.front();
        Tile t;
        t.vw = 1;
        t.cols = 1;
        if (first.vw == 1) {
            t.data = first.data;
        } else {
            void* tmp = ctx->aux2.get((size_t)n * sb);
            launch_scatter_block(h2->scalar, first.data, n, first.vw, 0, 1, tmp, ctx->stream);
            KPMX_HIP(hipGetLastError());
            t.data = tmp;
        }
        const int64_t steps = M - 1;
        double* sums = static_cast<double*>(ctx->sums.get((size_t)steps * 4 * sizeof(double)));
        run_recurrence(h2, t, MODE_DOT, steps, dw, sums, [](StepLaunch&, int64_t) {}, [](StepLaunch const&, int64_t) {});
        std::vector<double> s((size_t)steps * 4);
        KPMX_HIP(hipMemcpyAsync(s.data(), sums, s.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        KPMX_HIP(hipStreamSynchronize(ctx->stream));
        auto put = [&](int64_t m, double re, double im) {
            if (h2->scalar == 0) static_cast<float*>(mu)[m] = (float)re;
            else if (h2->scalar == 1) static_cast<double*>(mu)[m] = re;
            else if (h2->scalar == 2) { static_cast<float*>(mu)[2 * m] = (float)re; static_cast<float*>(mu)[2 * m + 1] = (float)im; }
            else { static_cast<double*>(mu)[2 * m] = re; static_cast<double*>(mu)[2 * m + 1] = im; }
        };
        put(0, 0.5 * s[0], 0.5 * s[1]);
        for (int64_t k = 0; k < steps; ++k) put(k + 1, s[k * 4 + 2], s[k * 4 + 3]);
    });
}

int kpmx_moments_multi_unit(const kpmx_mat* h2, const kpmx_block* r0, int64_t M, int64_t ndest, const int32_t* dest,
                            void* mu) {
    return guarded([&] {
        check_pair(h2, r0);
        KPMX_REQUIRE(mu && (ndest == 0 || dest), "null output or dest");
        KPMX_REQUIRE(M >= 2, "num_moments must be >= 2");
        kpmx_ctx* ctx = h2->ctx;
        KPMX_HIP(hipSetDevice(ctx->device));
        const int64_t n = h2->rows;
        const size_t sb = scalar_bytes(h2->scalar);
        // unique destinations -> slots
        std::vector<int32_t> mark((size_t)n, -1);
        std::vector<int32_t> slot_of((size_t)ndest);
        std::vector<int32_t> uniq;
        for (int64_t i = 0; i < ndest; ++i) {
            KPMX_REQUIRE(dest[i] >= 0 && dest[i] < n, "destination index out of range");
            if (mark[dest[i]] < 0) {
                mark[dest[i]] = (int32_t)uniq.size();
                uniq.push_back(dest[i]);
            }
            slot_of[i] = mark[dest[i]];
        }
        const int64_t nu = (int64_t)uniq.size();
        int32_t* dmark = static_cast<int32_t*>(ctx->aux1.get((size_t)n * sizeof(int32_t)));
        KPMX_HIP(hipMemcpyAsync(dmark, mark.data(), (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
        char* dmu = static_cast<char*>(ctx->aux2.get((size_t)std::max<int64_t>(nu, 1) * M * sb));
        // a step writes the slots of the rows it computes; under a light cone a destination the
        // recurrence has not reached keeps the zero it has in the exact recurrence
        launch_zero(h2->scalar, dmu, std::max<int64_t>(nu, 1) * M, ctx->stream);
        KPMX_HIP(hipGetLastError());
        const Tile& first = r0->tiles.front();
        Tile t;
        t.vw = 1;
        t.cols = 1;
        void* vec;
        if (first.vw == 1) {
            vec = first.data;
        } else {
            vec = ctx->aux0.get((size_t)n * sb);
            launch_scatter_block(h2->scalar, first.data, n, first.vw, 0, 1, vec, ctx->stream);
            KPMX_HIP(hipGetLastError());
        }
        t.data = vec;
        run_recurrence(h2, t, MODE_NONE, M - 1, nullptr, nullptr, [&](StepLaunch& L, int64_t k) {
            L.mark = dmark;
            L.mark_out = dmu + (size_t)(k + 1) * sb;
            L.mark_stride = M;
        }, [](StepLaunch const&, int64_t) {});
        std::vector<char> host((size_t)nu * M * sb);
        if (nu) KPMX_HIP(hipMemcpyAsync(host.data(), dmu, host.size(), hipMemcpyDeviceToHost, ctx->stream));
        std::vector<char> r0host((size_t)nu * sb);
        for (int64_t i = 0; i < nu; ++i)
            KPMX_HIP(hipMemcpyAsync(&r0host[i * sb], static_cast<char*>(vec) + (size_t)uniq[i] * sb, sb,
                                    hipMemcpyDeviceToHost, ctx->stream));
        KPMX_HIP(hipStreamSynchronize(ctx->stream));
        char* out = static_cast<char*>(mu);
        for (int64_t i = 0; i < ndest; ++i) {
            const int32_t s = slot_of[i];
            std::memcpy(out + (size_t)i * M * sb, &host[(size_t)s * M * sb], (size_t)M * sb);
            // moment 0: 0.5 * r0[dest]
            if (h2->scalar == 0 || h2->scalar == 2) {
                float* o = reinterpret_cast<float*>(out + (size_t)i * M * sb);
                const float* v = reinterpret_cast<const float*>(&r0host[(size_t)s * sb]);
                o[0] = v[0] * 0.5f;
                if (h2->scalar == 2) o[1] = v[1] * 0.5f;
            } else {
                double* o = reinterpret_cast<double*>(out + (size_t)i * M * sb);
                const double* v = reinterpret_cast<const double*>(&r0host[(size_t)s * sb]);
                o[0] = v[0] * 0.5;
                if (h2->scalar == 3) o[1] = v[1] * 0.5;
            }
        }
    });
}


int kpmx_moments_dense(const kpmx_mat* h2, const kpmx_block* r0, int64_t M, const kpmx_mat* op, void* out) {
    return guarded([&] {
        check_pair(h2, r0);
        KPMX_REQUIRE(out, "null output");
        KPMX_REQUIRE(M >= 2, "num_moments must be >= 2");
        if (op) {
            KPMX_REQUIRE(op->ctx == h2->ctx && op->scalar == h2->scalar && op->rows == h2->rows,
                         "operator does not match the Hamiltonian");
        }
        kpmx_ctx* ctx = h2->ctx;
        KPMX_HIP(hipSetDevice(ctx->device));
        const int64_t n = h2->rows;
        const size_t sb = scalar_bytes(h2->scalar);
        for (const Tile& t : r0->tiles) {
            // device panel: vw x M x n (vector, moment, site)
            char* D = static_cast<char*>(ctx->aux2.get((size_t)t.vw * M * n * sb));
            StepLaunch O;
            if (op) {
                O = base_launch(op, t.vw, false);
                O.mode = MODE_NONE;
                O.yout = nullptr;
                O.sub_y = 0;
                O.copy_col_stride = M * n;
                // row 0: op (0.5 r0)
                O.x = t.data;
                O.hscale = 0.5;
                O.copy_out = D;
                launch(O, ctx->stream);
            } else {
                launch_scatter_scaled(h2->scalar, t.data, n, t.vw, t.vw, 0.5, D, M * n, ctx->stream);
                KPMX_HIP(hipGetLastError());
            }
            run_recurrence(
                h2, t, MODE_NONE, M - 1, nullptr, nullptr,
                [&](StepLaunch& L, int64_t k) {
                    if (!op) {
                        L.copy_out = D + (size_t)(k + 1) * n * sb;
                        L.copy_col_stride = M * n;
                    }
                },
                [&](StepLaunch const& L, int64_t k) {
                    if (op) {
                        O.x = L.yout;
                        O.hscale = 1.0;
                        O.copy_out = D + (size_t)(k + 1) * n * sb;
                        launch(O, ctx->stream);
                    }
                });
            for (int64_t c = 0; c < t.cols; ++c)
                KPMX_HIP(hipMemcpyAsync(static_cast<char*>(out) + (size_t)(t.c0 + c) * M * n * sb, D + (size_t)c * M * n * sb,
                                        (size_t)M * n * sb, hipMemcpyDeviceToHost, ctx->stream));
            KPMX_HIP(hipStreamSynchronize(ctx->stream));
        }
    });
}

int kpmx_kubo_moments(const kpmx_mat* h2, const kpmx_mat* vl, const kpmx_mat* vr, const kpmx_block* r0, int64_t M,
                      int mode, void* mu_nm) {
    return guarded([&] {
        check_pair(h2, r0);
        KPMX_REQUIRE(vl && vr && mu_nm, "null argument");
        KPMX_REQUIRE(mode == KPMX_KUBO_FP64 || mode == KPMX_KUBO_FP32_SPLIT || mode == KPMX_KUBO_BF16X3,
                     "invalid contraction mode");
        for (const kpmx_mat* v : {vl, vr})
            KPMX_REQUIRE(v->ctx == h2->ctx && v->scalar == h2->scalar && v->rows == h2->rows,
                         "velocity operator does not match the Hamiltonian");
        KPMX_REQUIRE(M >= 2, "num_moments must be >= 2");
        kpmx_ctx* ctx = h2->ctx;
        KPMX_HIP(hipSetDevice(ctx->device));
        const int64_t n = h2->rows;
        const int sc = h2->scalar;
        const size_t sb = scalar_bytes(sc);
        const bool cplx = is_complex(sc);
        const int64_t cw = cplx ? 2 : 1;

        // group width g: both panels (2 g M n scalars) within 80% of the device memory that is free
        // or held by this context's panel workspaces (they are given back before growing), so that
        // a 288 GB device advances as many vectors per sweep as the panels allow: every extra
        // vector in a group shares the sweep's matrix reads
        size_t free_b = 0, total_b = 0;
        KPMX_HIP(hipMemGetInfo(&free_b, &total_b));
        const double budget = 0.8 * (double)(free_b + ctx->panel_l.bytes + ctx->panel_r.bytes);
        // bf16x3: panels are written by the sweeps as bf16 pieces (3 per real part), rows padded to
        // a multiple of 8 elements with a zero tail
        const bool split = mode == KPMX_KUBO_BF16X3;
        const int64_t ldp = split ? (n + 7) / 8 * 8 : n;
        const size_t plane = (size_t)M * ldp;
        const size_t panel1 = split ? (size_t)cw * 3 * plane * sizeof(uint16_t) : (size_t)M * n * sb;
        KPMX_REQUIRE(2.0 * panel1 <= budget, "Kubo moment panels do not fit in device memory");

        double* C = static_cast<double*>(ctx->sums.get((size_t)M * M * cw * sizeof(double)));
        launch_zero_f64(C, M * M * cw, ctx->stream);
        KPMX_HIP(hipGetLastError());

        for (const Tile& full : r0->tiles) {
            int g = full.vw;
            while (g > 1 && 2.0 * g * panel1 > budget) g >>= 1;
            // a g-wide sweep writing split panels stores to 6 g bf16 planes (each M x n, about 1 GB
            // apart at the 1 M-site workload) at once; measured at 512 moments on 1.05 M sites,
            // 16 vectors: 0.70 s (g = 2), 0.56 s (4), 0.90 s (8), 3.18 s (16)
            if (split) g = std::min(g, 4);
            for (int64_t c0 = 0; c0 < full.cols; c0 += g) {
                const int64_t cols = std::min<int64_t>(g, full.cols - c0);
                // this group's starters r (n x g), and vl r (n x g)
                void* Pr = ctx->aux0.get((size_t)n * g * sb);
                if (g == full.vw) {
                    KPMX_HIP(hipMemcpyAsync(Pr, full.data, (size_t)n * g * sb, hipMemcpyDeviceToDevice, ctx->stream));
                } else {
                    // pull columns [c0, c0+g) of the tile out (via a vector-major bounce)
                    void* vm = ctx->aux1.get((size_t)n * full.vw * sb);
                    launch_scatter_block(sc, full.data, n, full.vw, 0, full.vw, vm, ctx->stream);
                    launch_gather_block(sc, vm, n, c0, cols, g, Pr, ctx->stream);
                    KPMX_HIP(hipGetLastError());
                }
                if (ctx->panel_l.bytes < (size_t)g * panel1 || ctx->panel_r.bytes < (size_t)g * panel1) {
                    ctx->panel_l.release();
                    ctx->panel_r.release();
                }
                char* Lpan = static_cast<char*>(ctx->panel_l.get((size_t)g * panel1));
                char* Rpan = static_cast<char*>(ctx->panel_r.get((size_t)g * panel1));
                uint16_t* Lsp = reinterpret_cast<uint16_t*>(Lpan);
                uint16_t* Rsp = reinterpret_cast<uint16_t*>(Rpan);
                if (split && ldp > n) {
                    launch_zero_tails(Lsp, (int64_t)g * cw * 3 * M, ldp, n, ctx->stream);
                    launch_zero_tails(Rsp, (int64_t)g * cw * 3 * M, ldp, n, ctx->stream);
                }
                void* L0 = ctx->aux2.get((size_t)n * g * sb);

                // L0 = vl r
                StepLaunch V = base_launch(vl, g, true);   // SELL when the operator was uploaded so
                V.mode = MODE_NONE;
                V.sub_y = 0;
                V.hscale = 1.0;
                V.x = Pr;
                V.yout = L0;
                launch(V, ctx->stream);
                // L panel rows: T_n(H~) L0 (row 0 halved)
                Tile tl;
                tl.vw = g;
                tl.cols = cols;
                tl.data = L0;
                if (split)
                    launch_scatter_split(sc, L0, n, g, g, 0.5, Lsp, (int64_t)plane, ctx->stream);
                else
                    launch_scatter_scaled(sc, L0, n, g, g, 0.5, Lpan, M * n, ctx->stream);
                KPMX_HIP(hipGetLastError());
                run_recurrence(
                    h2, tl, MODE_NONE, M - 1, nullptr, nullptr,
                    [&](StepLaunch& L, int64_t k) {
                        if (split) {
                            L.copy_split = Lsp + (size_t)(k + 1) * ldp;
                            L.copy_col_stride = (int64_t)plane;
                        } else {
                            L.copy_out = Lpan + (size_t)(k + 1) * n * sb;
                            L.copy_col_stride = M * n;
                        }
                    },
                    [](StepLaunch const&, int64_t) {});
                // R panel rows: vr T_m(H~) r (row 0 halved)
                StepLaunch O = base_launch(vr, g, true);
                O.mode = MODE_NONE;
                O.sub_y = 0;
                O.yout = nullptr;
                O.copy_col_stride = split ? (int64_t)plane : M * n;
                O.x = Pr;
                O.hscale = 0.5;
                if (split)
                    O.copy_split = Rsp;
                else
                    O.copy_out = Rpan;
                launch(O, ctx->stream);
                Tile tr;
                tr.vw = g;
                tr.cols = cols;
                tr.data = Pr;
                run_recurrence(
                    h2, tr, MODE_NONE, M - 1, nullptr, nullptr, [](StepLaunch&, int64_t) {},
                    [&](StepLaunch const& L, int64_t k) {
                        O.x = L.yout;
                        O.hscale = 1.0;
                        if (split)
                            O.copy_split = Rsp + (size_t)(k + 1) * ldp;
                        else
                            O.copy_out = Rpan + (size_t)(k + 1) * n * sb;
                        launch(O, ctx->stream);
                    });
                // mu_nm += sum_c L_c R_c^H
                const int64_t wd = gemm_workspace_doubles(M, n, sc, mode, ctx->num_cu);
                double* work = static_cast<double*>(ctx->partials.get((size_t)wd * sizeof(double)));
                for (int64_t c = 0; c < cols; ++c) {
                    const int rc = split ? launch_gemm_nc_planes(cplx, Lsp + (size_t)c * cw * 3 * plane,
                                                                 Rsp + (size_t)c * cw * 3 * plane, ldp, (int64_t)plane,
                                                                 M, n, C, work, ctx->num_cu, ctx->stream)
                                         : launch_gemm_nc(sc, mode, Lpan + (size_t)c * M * n * sb, n,
                                                          Rpan + (size_t)c * M * n * sb, n, M, n, C, work, ctx->num_cu,
                                                          ctx->stream);
                    if (rc != 0) throw Error{KPMX_EINVAL, "unsupported contraction"};
                    KPMX_HIP(hipGetLastError());
                }
            }
        }
        std::vector<double> h((size_t)M * M * cw);
        KPMX_HIP(hipMemcpyAsync(h.data(), C, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        KPMX_HIP(hipStreamSynchronize(ctx->stream));
        const double inv = 1.0 / (double)r0->nvec;
        for (size_t e = 0; e < h.size(); ++e) {
            if (sc == 0 || sc == 2) static_cast<float*>(mu_nm)[e] = (float)(h[e] * inv);
            else static_cast<double*>(mu_nm)[e] = h[e] * inv;
        }
    });
}


int kpmx_lanczos_bounds(const kpmx_mat* h, double precision_percent, double* min_energy, double* max_energy,
                        int* loops) {
    return guarded([&] {
        KPMX_REQUIRE(h && min_energy && max_energy && loops, "null argument");
        KPMX_REQUIRE(h->rows > 0, "empty matrix");
        kpmx_ctx* ctx = h->ctx;
        KPMX_HIP(hipSetDevice(ctx->device));
        const int64_t n = h->rows;
        const int sc = h->scalar;
        const size_t sb = scalar_bytes(sc);
        // v1 = uniform [0, 1) reals from a default-seeded mt19937, normalized; v0 = 0
        std::vector<char> hv((size_t)n * sb, 0);
        std::mt19937 gen;
        if (sc == 0 || sc == 2) {
            std::uniform_real_distribution<float> dist;
            std::vector<float> v((size_t)n);
            for (auto& x : v) x = dist(gen);
            float nrm = 0;
            for (auto x : v) nrm += x * x;
            nrm = std::sqrt(nrm);
            float* o = reinterpret_cast<float*>(hv.data());
            for (int64_t i = 0; i < n; ++i) o[(sc == 2 ? 2 : 1) * i] = v[(size_t)i] / nrm;
        } else {
            std::uniform_real_distribution<double> dist;
            std::vector<double> v((size_t)n);
            for (auto& x : v) x = dist(gen);
            double nrm = 0;
            for (auto x : v) nrm += x * x;
            nrm = std::sqrt(nrm);
            double* o = reinterpret_cast<double*>(hv.data());
            for (int64_t i = 0; i < n; ++i) o[(sc == 3 ? 2 : 1) * i] = v[(size_t)i] / nrm;
        }
        void* v1 = ctx->work1.get((size_t)n * sb);
        void* v0 = ctx->work0.get((size_t)n * sb);
        KPMX_HIP(hipMemcpyAsync(v1, hv.data(), hv.size(), hipMemcpyHostToDevice, ctx->stream));
        launch_zero(sc, v0, n, ctx->stream);
        const int nblocks = (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, (int64_t)ctx->num_cu * 4));
        double* partials = static_cast<double*>(ctx->partials.get((size_t)nblocks * sizeof(double)));
        double* red = static_cast<double*>(ctx->sums.get(4 * sizeof(double)));
        std::vector<double> alpha, beta;
        const double precision = precision_percent / 100;
        double prev_min = std::numeric_limits<double>::max(), prev_max = std::numeric_limits<double>::lowest();
        for (int i = 0; i < 1000; ++i) {
            const double b_prev = beta.empty() ? 0.0 : beta.back();
            launch_lanczos_spmv(sc, h->indptr, h->indices, h->vals, v1, v0, n, b_prev, partials, nblocks, ctx->stream);
            launch_reduce_partials(partials, nblocks, 1, 1, red, ctx->stream);
            double a = 0;
            KPMX_HIP(hipMemcpyAsync(&a, red, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
            KPMX_HIP(hipStreamSynchronize(ctx->stream));
            launch_lanczos_axpy(sc, v1, v0, n, a, partials, nblocks, ctx->stream);
            launch_reduce_partials(partials, nblocks, 1, 1, red, ctx->stream);
            double b2 = 0;
            KPMX_HIP(hipMemcpyAsync(&b2, red, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
            KPMX_HIP(hipStreamSynchronize(ctx->stream));
            const double b = std::sqrt(b2);
            alpha.push_back(a);
            beta.push_back(b);
            if (b > 0) launch_scale(sc, v0, n, 1.0 / b, ctx->stream);
            KPMX_HIP(hipGetLastError());
            std::swap(v0, v1);
            double mn, mx;
            tridiagonal_minmax(alpha, beta, alpha.size(), &mn, &mx);
            const bool conv_min = std::abs((prev_min - mn) / mn) < precision;
            const bool conv_max = std::abs((prev_max - mx) / mx) < precision;
            if ((conv_min && conv_max) || b == 0) {
                *min_energy = mn;
                *max_energy = mx;
                *loops = i;
                return;
            }
            prev_min = mn;
            prev_max = mx;
        }
        throw Error{KPMX_EINVAL, "Lanczos algorithm did not converge for the min/max eigenvalues."};
    });
}

int kpmx_ctx_timer_start(kpmx_ctx* ctx) {
    return guarded([&] {
        KPMX_REQUIRE(ctx, "null context");
        KPMX_HIP(hipSetDevice(ctx->device));
        if (!ctx->t0) KPMX_HIP(hipEventCreate(&ctx->t0));
        if (!ctx->t1) KPMX_HIP(hipEventCreate(&ctx->t1));
        KPMX_HIP(hipEventRecord(ctx->t0, ctx->stream));
    });
}

int kpmx_ctx_timer_stop(kpmx_ctx* ctx, double* ms) {
    return guarded([&] {
        KPMX_REQUIRE(ctx && ms, "null argument");
        KPMX_REQUIRE(ctx->t0 && ctx->t1, "timer was not started");
        KPMX_HIP(hipEventRecord(ctx->t1, ctx->stream));
        KPMX_HIP(hipEventSynchronize(ctx->t1));
        float f = 0;
        KPMX_HIP(hipEventElapsedTime(&f, ctx->t0, ctx->t1));
        *ms = (double)f;
    });
}

int kpmx_ctx_set_adjoint_variant(kpmx_ctx* ctx, int variant) {
    return guarded([&] {
        KPMX_REQUIRE(ctx, "null context");
        KPMX_REQUIRE(variant >= -1 && variant < NUM_TUNE_VARIANTS, "variant out of range");
        ctx->adjoint_variant = variant;
    });
}

int kpmx_ctx_set_step_variant(kpmx_ctx* ctx, int scalar, int vector_width, int variant) {
    return guarded([&] {
        KPMX_REQUIRE(ctx, "null context");
        KPMX_REQUIRE(valid_scalar(scalar), "invalid scalar code");
        KPMX_REQUIRE(vector_width >= 1 && vector_width <= 64 && (vector_width & (vector_width - 1)) == 0,
                     "vector width must be a power of two up to 64");
        KPMX_REQUIRE(variant >= 0 && variant < NUM_TUNE_VARIANTS, "variant out of range");
        ctx->tune[scalar][log2i(vector_width)] = variant;
        ctx->tune_coded[scalar][log2i(vector_width)] = -1;
    });
}

namespace {
// The LDS two-step form against single steps (fp64 and complex128 SELL matrices): column tiles of
// 64-byte rows (8 fp64 / 4 complex128 vectors) with 4 workgroups per CU and of 128-byte rows with
// 3, each timed on the first column tile it would run;
// the fastest per vector is kept in the context (tile width and two-step settings) when it beats
// the single steps' `single_per_vector` milliseconds per vector and step by 2 %, else the
// settings are left as they were.
void autotune_two_step(const kpmx_mat* h2, const kpmx_block* r0, double single_per_vector) {
    kpmx_ctx* ctx = h2->ctx;
    const int sc = h2->scalar;
    if ((sc != KPMX_F64 && sc != KPMX_C128) || h2->format != KPMX_FMT_SELL || !h2->step_rows.empty() ||
        h2->rows < 2 * 64)
        return;
    const int saved[4] = {ctx->two_step[sc], ctx->two_step_wg_per_cu[sc], ctx->two_step_lds[sc], ctx->tile_cap[sc]};
    auto set = [&](int rows, int wg, int lds, int cap) {
        ctx->two_step[sc] = rows;
        ctx->two_step_wg_per_cu[sc] = wg;
        ctx->two_step_lds[sc] = lds;
        ctx->tile_cap[sc] = cap;
    };
    // about 4 GB of vector traffic per timed sample of 64-byte rows (16 to 256 steps, even)
    int64_t st = std::max<int64_t>(16, std::min<int64_t>(256, (int64_t)(4e9 / (3.0 * h2->rows * 64))));
    st -= st % 2;
    struct Cand {
        int cap, wg;
        bool ok;
        float ms;
    };
    // tiles of 64-byte rows on 4 workgroups per CU and of 128-byte rows on 3
    const int narrow = sc == KPMX_C128 ? 4 : 8;
    std::vector<Cand> cands{{narrow, 4, false, 0.f}, {2 * narrow, 3, false, 0.f}};
    // one untimed run of every candidate (code loading, workspaces), then three rounds that
    // alternate the candidates; each keeps its fastest sample
    for (int round = 0; round < 4; ++round) {
        for (Cand& c : cands) {
            set(STEP2_LDS_TILE, c.wg, 1, c.cap);
            const Tile t = stage_run_tile(h2, r0, run_tiles(r0, c.cap).front());
            if (round == 0) c.ok = use_two_step(h2, t);
            if (!c.ok) continue;
            double* sums = static_cast<double*>(ctx->sums.get((size_t)st * 4 * t.vw * sizeof(double)));
            KPMX_HIP(hipEventRecord(ctx->t0, ctx->stream));
            run_recurrence2(h2, t, st, sums);
            KPMX_HIP(hipEventRecord(ctx->t1, ctx->stream));
            KPMX_HIP(hipEventSynchronize(ctx->t1));
            float ms = 0;
            KPMX_HIP(hipEventElapsedTime(&ms, ctx->t0, ctx->t1));
            const float per = ms / (float)t.vw;  // per vector
            if (round == 1 || (round > 1 && per < c.ms)) c.ms = per;
        }
    }
    double best = 0.98 * single_per_vector;
    int best_cap = 0, best_wg = 0;
    for (const Cand& c : cands) {
        const double per_vector = c.ms / (double)st;
        if (c.ok && per_vector < best) {
            best = per_vector;
            best_cap = c.cap;
            best_wg = c.wg;
        }
    }
    if (best_cap) {
        set(STEP2_LDS_TILE, best_wg, 1, best_cap);
        ctx->two_step_auto[sc] = true;
    } else {
        set(saved[0], saved[1], saved[2], saved[3]);
    }
}
}  // namespace

int kpmx_autotune_diagonal(const kpmx_mat* h2, const kpmx_block* r0, int* variant) {
    return guarded([&] {
        check_pair(h2, r0);
        kpmx_ctx* ctx = h2->ctx;
        KPMX_HIP(hipSetDevice(ctx->device));
        if (ctx->two_step_auto[h2->scalar]) {  // an earlier autotuner choice: start from the defaults
            ctx->two_step[h2->scalar] = ctx->two_step_wg_per_cu[h2->scalar] = ctx->two_step_lds[h2->scalar] = 0;
            ctx->tile_cap[h2->scalar] = default_tile_cap(h2->scalar);
            ctx->two_step_auto[h2->scalar] = false;
        }
        // the first column tile the diagonal recurrence runs (a sub-tile of a wider starter tile)
        const Tile t = stage_run_tile(h2, r0, run_tiles(r0, diag_cap(h2)).front());
        // About 2 GB of vector traffic per timed sample (at least 9 steps, at most 255): small
        // systems get enough steps that launch jitter does not decide the choice.
        const double step_bytes = 3.0 * (double)h2->rows * t.vw * scalar_bytes(h2->scalar);
        const int64_t steps = std::max<int64_t>(9, std::min<int64_t>(255, (int64_t)(2e9 / step_bytes)));
        double* sums = static_cast<double*>(ctx->sums.get((size_t)steps * 4 * t.vw * sizeof(double)));
        if (!ctx->t0) KPMX_HIP(hipEventCreate(&ctx->t0));
        if (!ctx->t1) KPMX_HIP(hipEventCreate(&ctx->t1));
        int best = 0;
        float best_ms = 0;
        for (int v = 0; v < NUM_TUNE_VARIANTS; ++v) {
            // the column-pair variants exist for fp64 blocks of 8 or more vectors only (elsewhere
            // they run the wave-staged kernel, already timed), the value-table ones only where the
            // matrix has that form (elsewhere they run as the column-pair ones)
            if (v >= 2 * NUM_TILE_SHAPES * NUM_GRID_VARIANTS && !(h2->scalar == KPMX_F64 && t.vw >= 8)) continue;
            if (v >= 3 * NUM_TILE_SHAPES * NUM_GRID_VARIANTS && !(h2->ccodes && h2->cmax_width <= 4)) continue;
            float ms_v = 0;
            for (int rep = 0; rep < 4; ++rep) {  // the first repetition warms caches and code
                KPMX_HIP(hipEventRecord(ctx->t0, ctx->stream));
                run_recurrence(h2, t, MODE_DIAG, steps, nullptr, sums, [](StepLaunch&, int64_t) {},
                               [](StepLaunch const&, int64_t) {}, v);
                KPMX_HIP(hipEventRecord(ctx->t1, ctx->stream));
                KPMX_HIP(hipEventSynchronize(ctx->t1));
                float ms = 0;
                KPMX_HIP(hipEventElapsedTime(&ms, ctx->t0, ctx->t1));
                if (rep == 1 || (rep > 1 && ms < ms_v)) ms_v = ms;
            }
            if (v == 0 || ms_v < best_ms) {
                best_ms = ms_v;
                best = v;
            }
        }
        ctx->tune[h2->scalar][log2i(t.vw)] = best;
        ctx->tune_coded[h2->scalar][log2i(t.vw)] = -1;
        if (variant) *variant = best;
        if (ctx->autotune_two_step) autotune_two_step(h2, r0, best_ms / ((double)steps * t.vw));
    });
}

namespace {
// The value-table form of a batch (the step kernel's CODED path).  Per system, the distinct
// off-diagonal values (compared bit for bit, code 0 is zero: padding and explicit zeros) make a
// table of at most 256 entries; the diagonal entries go to a dense vector over the padded rows.
// Returns without building when a system has more than 255 distinct off-diagonal values or a row
// holds two diagonal entries.  The SELL slices are the standard form's (same slices, same
// systems); only the off-diagonal widths change.
void build_batch_coded(kpmx_batch* b, const int64_t* row_offsets, const int32_t* indptr, const int32_t* indices,
                       const void* values) {
    const size_t sb = scalar_bytes(b->scalar);
    const int64_t S = b->num_sys, ns_tot = b->nslices;
    const char* v = static_cast<const char*>(values);
    struct Key {
        uint64_t lo = 0, hi = 0;
        bool operator==(const Key& o) const { return lo == o.lo && hi == o.hi; }
    };
    auto key_of = [&](int64_t k) {
        Key key;
        char buf[16] = {};
        std::memcpy(buf, v + (size_t)k * sb, sb);
        std::memcpy(&key.lo, buf, 8);
        std::memcpy(&key.hi, buf + 8, 8);
        return key;
    };
    struct KeyHash {
        size_t operator()(const Key& k) const { return std::hash<uint64_t>()(k.lo * 0x9E3779B97F4A7C15ull ^ k.hi); }
    };
    std::vector<int32_t> tab_off(S + 1, 0);
    std::vector<char> table;
    std::vector<uint8_t> code_of_entry((size_t)indptr[row_offsets[S]]);
    std::vector<int32_t> cwidth(ns_tot, 0);
    std::vector<int64_t> diag_entry(b->n_pad, -1);  // CSR position of every padded row's diagonal
    for (int64_t sy = 0; sy < S; ++sy) {
        std::unordered_map<Key, int, KeyHash> codes;
        std::vector<Key> keys{Key{}};
        codes.emplace(Key{}, 0);
        Key last{};
        int last_code = 0;
        const int64_t shift = b->pad_off[sy] - row_offsets[sy];
        for (int64_t r = row_offsets[sy]; r < row_offsets[sy + 1]; ++r) {
            int32_t off = 0;
            for (int32_t k = indptr[r]; k < indptr[r + 1]; ++k) {
                if (indices[k] == r) {
                    if (diag_entry[r + shift] >= 0) return;  // two diagonal entries in one row
                    diag_entry[r + shift] = k;
                    continue;
                }
                ++off;
                const Key key = key_of(k);
                if (!(key == last)) {
                    auto it = codes.find(key);
                    if (it == codes.end()) {
                        if (keys.size() >= 256) return;  // more values than one byte codes
                        it = codes.emplace(key, (int)keys.size()).first;
                        keys.push_back(key);
                    }
                    last = key;
                    last_code = it->second;
                }
                code_of_entry[k] = (uint8_t)last_code;
            }
            const int64_t sl = (r + shift) / 64;
            cwidth[sl] = std::max(cwidth[sl], off);
        }
        tab_off[sy + 1] = tab_off[sy] + (int32_t)keys.size();
        for (const Key& key : keys) {
            char buf[16];
            std::memcpy(buf, &key.lo, 8);
            std::memcpy(buf + 8, &key.hi, 8);
            table.insert(table.end(), buf, buf + sb);
        }
    }
    std::vector<int64_t> cptr(ns_tot + 1, 0);
    for (int64_t sl = 0; sl < ns_tot; ++sl) cptr[sl + 1] = cptr[sl] + (int64_t)cwidth[sl] * 64;
    const int64_t total = cptr[ns_tot];
    std::vector<int32_t> cidx((size_t)std::max<int64_t>(total, 1));
    std::vector<uint8_t> ccodes((size_t)std::max<int64_t>(total, 1), 0);
    std::vector<char> diag((size_t)b->n_pad * sb, 0);
    for (int64_t sy = 0; sy < S; ++sy) {
        const int64_t shift = b->pad_off[sy] - row_offsets[sy];
        for (int64_t pr = b->pad_off[sy]; pr < b->pad_off[sy + 1]; ++pr) {
            const int64_t sl = pr / 64, lane = pr % 64, r = pr - shift;
            int32_t k = 0;
            if (r < row_offsets[sy + 1]) {
                for (int32_t e = indptr[r]; e < indptr[r + 1]; ++e) {
                    if (e == diag_entry[pr]) continue;
                    const int64_t dst = cptr[sl] + (int64_t)k * 64 + lane;
                    cidx[dst] = (int32_t)(indices[e] + shift);
                    ccodes[dst] = code_of_entry[e];
                    ++k;
                }
                if (diag_entry[pr] >= 0) std::memcpy(&diag[(size_t)pr * sb], v + (size_t)diag_entry[pr] * sb, sb);
            }
            for (; k < cwidth[sl]; ++k) cidx[cptr[sl] + (int64_t)k * 64 + lane] = (int32_t)pr;  // code 0: zero
        }
    }
    auto up = [&](void* dst, const void* src, size_t bytes) {
        if (bytes) KPMX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, b->ctx->stream));
    };
    b->cslice_ptr = static_cast<int64_t*>(dev_alloc((ns_tot + 1) * sizeof(int64_t)));
    b->cslice_width = static_cast<int32_t*>(dev_alloc(ns_tot * sizeof(int32_t)));
    b->cidx = static_cast<int32_t*>(dev_alloc(cidx.size() * sizeof(int32_t)));
    b->ccodes = static_cast<uint8_t*>(dev_alloc(ccodes.size()));
    b->cdiag = dev_alloc(diag.size());
    b->ctable = dev_alloc(table.size());
    b->ctab_off = static_cast<int32_t*>(dev_alloc(tab_off.size() * sizeof(int32_t)));
    up(b->cslice_ptr, cptr.data(), cptr.size() * sizeof(int64_t));
    up(b->cslice_width, cwidth.data(), cwidth.size() * sizeof(int32_t));
    up(b->cidx, cidx.data(), cidx.size() * sizeof(int32_t));
    up(b->ccodes, ccodes.data(), ccodes.size());
    up(b->cdiag, diag.data(), diag.size());
    up(b->ctable, table.data(), table.size());
    up(b->ctab_off, tab_off.data(), tab_off.size() * sizeof(int32_t));
    KPMX_HIP(hipStreamSynchronize(b->ctx->stream));  // the host vectors go out of scope
    b->cmax_width = ns_tot ? *std::max_element(cwidth.begin(), cwidth.end()) : 0;
    b->c_entries = total;
    b->table_len = tab_off[S];
    b->has_coded = b->use_coded = true;
}
}  // namespace

int kpmx_batch_upload(kpmx_ctx* ctx, int scalar, int64_t num_systems, const int64_t* row_offsets, int64_t nnz,
                      const int32_t* indptr, const int32_t* indices, const void* values, kpmx_batch** out) {
    return guarded([&] {
        KPMX_REQUIRE(ctx && out && row_offsets && indptr, "null argument");
        KPMX_REQUIRE(valid_scalar(scalar), "invalid scalar code");
        KPMX_REQUIRE(num_systems > 0, "no systems");
        KPMX_REQUIRE(row_offsets[0] == 0, "row offsets must start at 0");
        const int64_t n = row_offsets[num_systems];
        KPMX_REQUIRE(n > 0 && n < ((int64_t)1 << 31) && indptr[n] == nnz && (nnz == 0 || (indices && values)),
                     "inconsistent CSR arrays");
        KPMX_HIP(hipSetDevice(ctx->device));
        const size_t sb = scalar_bytes(scalar);
        auto b = std::make_unique<kpmx_batch>();
        b->ctx = ctx;
        b->scalar = scalar;
        b->num_sys = num_systems;
        b->row_off.assign(row_offsets, row_offsets + num_systems + 1);
        b->pad_off.assign(num_systems + 1, 0);
        for (int64_t sy = 0; sy < num_systems; ++sy) {
            const int64_t ns = row_offsets[sy + 1] - row_offsets[sy];
            KPMX_REQUIRE(ns > 0, "empty system");
            b->pad_off[sy + 1] = b->pad_off[sy] + (ns + 63) / 64 * 64;
        }
        b->n_pad = b->pad_off[num_systems];
        const int64_t ns_tot = b->n_pad / 64;
        b->nslices = ns_tot;
        // SELL-64 over padded rows; column indices remapped into the padded numbering
        std::vector<int64_t> sptr(ns_tot + 1, 0);
        std::vector<int32_t> swidth(ns_tot, 0), ssys(ns_tot, 0);
        std::vector<int64_t> slice_first_row(ns_tot);  // unpadded row of lane 0 (or -1)
        for (int64_t sy = 0; sy < num_systems; ++sy) {
            for (int64_t pr = b->pad_off[sy]; pr < b->pad_off[sy + 1]; pr += 64) {
                const int64_t sl = pr / 64;
                ssys[sl] = (int32_t)sy;
                int32_t w = 0;
                for (int lane = 0; lane < 64; ++lane) {
                    const int64_t r = row_offsets[sy] + (pr - b->pad_off[sy]) + lane;
                    if (r < row_offsets[sy + 1]) {
                        for (int32_t k = indptr[r]; k < indptr[r + 1]; ++k)
                            KPMX_REQUIRE(indices[k] >= row_offsets[sy] && indices[k] < row_offsets[sy + 1],
                                         "the matrix is not block diagonal over the given systems");
                        w = std::max<int32_t>(w, indptr[r + 1] - indptr[r]);
                    }
                }
                swidth[sl] = w;
            }
        }
        for (int64_t sl = 0; sl < ns_tot; ++sl) sptr[sl + 1] = sptr[sl] + (int64_t)swidth[sl] * 64;
        const int64_t total = sptr[ns_tot];
        std::vector<int32_t> sidx((size_t)std::max<int64_t>(total, 1));
        std::vector<char> svals((size_t)std::max<int64_t>(total, 1) * sb, 0);
        const char* v = static_cast<const char*>(values);
        for (int64_t sy = 0; sy < num_systems; ++sy) {
            const int64_t shift = b->pad_off[sy] - row_offsets[sy];
            for (int64_t pr = b->pad_off[sy]; pr < b->pad_off[sy + 1]; ++pr) {
                const int64_t sl = pr / 64, lane = pr % 64;
                const int64_t r = pr - shift;
                const bool real_row = r < row_offsets[sy + 1];
                for (int32_t k = 0; k < swidth[sl]; ++k) {
                    const int64_t dst = sptr[sl] + (int64_t)k * 64 + lane;
                    if (real_row && indptr[r] + k < indptr[r + 1]) {
                        sidx[dst] = (int32_t)(indices[indptr[r] + k] + shift);
                        std::memcpy(&svals[dst * sb], v + (size_t)(indptr[r] + k) * sb, sb);
                    } else {
                        sidx[dst] = (int32_t)pr;
                    }
                }
            }
        }
        b->max_width = ns_tot ? *std::max_element(swidth.begin(), swidth.end()) : 0;
        b->s_entries = total;
        b->slice_sys_host = ssys;
        auto up = [&](void* dst, const void* src, size_t bytes) {
            if (bytes) KPMX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        };
        b->slice_ptr = static_cast<int64_t*>(dev_alloc((ns_tot + 1) * sizeof(int64_t)));
        b->slice_width = static_cast<int32_t*>(dev_alloc(ns_tot * sizeof(int32_t)));
        b->slice_sys = static_cast<int32_t*>(dev_alloc(ns_tot * sizeof(int32_t)));
        b->sidx = static_cast<int32_t*>(dev_alloc(sidx.size() * sizeof(int32_t)));
        b->svals = dev_alloc(svals.size());
        up(b->slice_ptr, sptr.data(), sptr.size() * sizeof(int64_t));
        up(b->slice_width, swidth.data(), swidth.size() * sizeof(int32_t));
        up(b->slice_sys, ssys.data(), ssys.size() * sizeof(int32_t));
        up(b->sidx, sidx.data(), sidx.size() * sizeof(int32_t));
        up(b->svals, svals.data(), svals.size());
        build_batch_coded(b.get(), row_offsets, indptr, indices, values);
        KPMX_HIP(hipStreamSynchronize(ctx->stream));
        batch_partition(b.get(), ctx->num_cu * 4, 0);
        *out = b.release();
    });
}

int kpmx_batch_starters(kpmx_batch* b, int64_t num_vectors, const void* r0) {
    return guarded([&] {
        KPMX_REQUIRE(b && r0, "null argument");
        KPMX_REQUIRE(num_vectors >= 1 && num_vectors <= 64, "1 to 64 vectors per system");
        kpmx_ctx* ctx = b->ctx;
        KPMX_HIP(hipSetDevice(ctx->device));
        const size_t sb = scalar_bytes(b->scalar);
        const int64_t cap = b->tile_cap > 0 ? b->tile_cap : BATCH_TILE_DEFAULT[b->scalar];
        const int64_t tile_cols = std::min<int64_t>(num_vectors, cap);
        const int64_t ntiles = (num_vectors + tile_cols - 1) / tile_cols;
        const int vw = std::max(4, pick_vw(tile_cols));
        if (b->vw != vw || b->ntiles != ntiles || !b->r0) {
            KPMX_HIP(hipStreamSynchronize(ctx->stream));
            dev_free(b->r0);
            dev_free(b->w0);
            dev_free(b->w1);
            dev_free(b->sums);
            b->r0 = b->w0 = b->w1 = nullptr;
            b->sums = nullptr;
            b->sums_steps = 0;
            b->drop_graph();
            b->r0 = dev_alloc((size_t)ntiles * b->n_pad * vw * sb);
            b->w0 = dev_alloc((size_t)b->n_pad * vw * sb);
            b->w1 = dev_alloc((size_t)b->n_pad * vw * sb);
        }
        b->vw = vw;
        b->ntiles = ntiles;
        b->tile_cols = tile_cols;
        b->nvec = num_vectors;
        // padded vector-major copy on the host, then one gather into the row-major block
        const int64_t n = b->row_off.back();
        std::vector<char> padded((size_t)num_vectors * b->n_pad * sb, 0);
        const char* src = static_cast<const char*>(r0);
        for (int64_t c = 0; c < num_vectors; ++c)
            for (int64_t sy = 0; sy < b->num_sys; ++sy)
                std::memcpy(&padded[((size_t)c * b->n_pad + b->pad_off[sy]) * sb], src + ((size_t)c * n + b->row_off[sy]) * sb,
                            (size_t)(b->row_off[sy + 1] - b->row_off[sy]) * sb);
        void* staging = ctx->aux0.get(padded.size());
        KPMX_HIP(hipMemcpyAsync(staging, padded.data(), padded.size(), hipMemcpyHostToDevice, ctx->stream));
        for (int64_t t = 0; t < ntiles; ++t)
            launch_gather_block(b->scalar, staging, b->n_pad, t * tile_cols,
                                std::min<int64_t>(tile_cols, num_vectors - t * tile_cols), vw,
                                static_cast<char*>(b->r0) + (size_t)t * b->n_pad * vw * sb, ctx->stream);
        KPMX_HIP(hipGetLastError());
        KPMX_HIP(hipStreamSynchronize(ctx->stream));
    });
}

namespace {
// The step launch of a batch in its current form; `standard` asks for the standard form whatever
// is selected.
void batch_launch(const kpmx_batch* b, BatchedLaunch& B, bool standard = false) {
    B = BatchedLaunch{};
    B.scalar = b->scalar;
    B.vw = b->vw;
    B.num_wg = b->num_wg;
    B.chunk = b->walk_chunk;
    B.nslices = b->nslices;
    B.max_width = b->max_width;
    B.indices = b->sidx;
    B.vals = b->svals;
    B.slice_ptr = b->slice_ptr;
    B.slice_width = b->slice_width;
    B.slice_sys = b->slice_sys;
    B.tile_begin = b->tile_begin;
    B.slot_begin = b->slot_begin;
    B.slot_index = b->slot_index;
    B.partials = b->partials;
    B.n = b->n_pad;
    if (b->use_coded && !standard) {
        B.max_width = b->cmax_width;
        B.indices = b->cidx;
        B.vals = nullptr;
        B.slice_ptr = b->cslice_ptr;
        B.slice_width = b->cslice_width;
        B.codes = b->ccodes;
        B.diag = b->cdiag;
        B.table = b->ctable;
        B.tab_off = b->ctab_off;
    }
}

// Bytes of the per-system sums of a run of `steps` steps: every column tile's steps x systems x 4 vw.
size_t batch_sums_bytes(const kpmx_batch* b, int64_t steps) {
    return (size_t)b->ntiles * steps * b->num_sys * 4 * b->vw * sizeof(double);
}

// The recurrence of column tile t: sums of step k at sums + ((t * steps + k) * num_sys) * 4 vw.
void batch_enqueue(kpmx_batch* b, int64_t steps, int64_t t = 0) {
    kpmx_ctx* ctx = b->ctx;
    const int width = 4 * b->vw;
    const void* r0 = static_cast<const char*>(b->r0) + (size_t)t * b->n_pad * b->vw * scalar_bytes(b->scalar);
    double* sums = b->sums + (size_t)t * steps * b->num_sys * width;
    BatchedLaunch B;
    batch_launch(b, B);
    for (int64_t k = 0; k < steps; ++k) {
        if (k == 0) {
            B.x = r0;
            B.yin = nullptr;
            B.yout = b->w1;
            B.hscale = 0.5;
            B.sub_y = 0;
        } else {
            const bool odd = (k % 2) == 1;
            B.x = odd ? b->w1 : b->w0;
            B.yout = odd ? b->w0 : b->w1;
            B.yin = (k == 1) ? r0 : B.yout;
            B.hscale = 1.0;
            B.sub_y = 1;
        }
        if (launch_batched_step(B, ctx->stream) != 0) throw Error{KPMX_EINVAL, "unsupported batched configuration"};
        launch_reduce_segments(b->partials, b->slot_first, b->num_sys, width, sums + (size_t)k * b->num_sys * width,
                               ctx->stream);
    }
}
}  // namespace

namespace {
// All steps in one persistent launch (launch_batched_persistent), then a check of its barrier.
void batch_run_persistent(kpmx_batch* b, int64_t steps) {
    kpmx_ctx* ctx = b->ctx;
    BatchedLaunch B;
    batch_launch(b, B);
    const int per_cu = batched_persistent_blocks_per_cu(B);
    if ((int64_t)b->num_wg > (int64_t)per_cu * ctx->num_cu)
        throw Error{KPMX_EINVAL, "persistent batched run: the partition's " + std::to_string(b->num_wg) +
                                     " workgroups exceed the resident grid of " + std::to_string(per_cu * ctx->num_cu) +
                                     "; choose fewer workgroups (kpmx_batch_set_partition)"};
    const int width = 4 * b->vw;
    if (!b->partials2)
        b->partials2 = static_cast<double*>(dev_alloc((size_t)2 * std::max<int64_t>(b->num_slots, 1) * 256 * sizeof(double)));
    if (!b->barrier) b->barrier = static_cast<unsigned*>(dev_alloc(4 * sizeof(unsigned)));
    B.w0 = b->w0;
    B.w1 = b->w1;
    B.partials = b->partials2;
    B.slot_first = b->slot_first;
    B.num_slots = b->num_slots;
    B.num_sys = b->num_sys;
    B.steps = steps;
    B.width = width;
    B.barrier = b->barrier;
    for (int64_t t = 0; t < b->ntiles; ++t) {  // one launch per column tile
        B.r0 = static_cast<const char*>(b->r0) + (size_t)t * b->n_pad * b->vw * scalar_bytes(b->scalar);
        B.sums = b->sums + (size_t)t * steps * b->num_sys * width;
        if (launch_batched_persistent(B, ctx->stream) != 0) throw Error{KPMX_EINVAL, "unsupported batched configuration"};
        KPMX_HIP(hipGetLastError());
        unsigned flags[4] = {0, 0, 0, 0};
        KPMX_HIP(hipMemcpyAsync(flags, b->barrier, sizeof(flags), hipMemcpyDeviceToHost, ctx->stream));
        KPMX_HIP(hipStreamSynchronize(ctx->stream));
        if (flags[2] != 0)
            throw Error{KPMX_EINVAL, "persistent batched run: the grid was not co-resident (barrier wait expired)"};
    }
}
}  // namespace

int kpmx_batch_run(kpmx_batch* b, int64_t M, int use_graph) {
    return guarded([&] {
        KPMX_REQUIRE(b, "null batch");
        KPMX_REQUIRE(b->r0, "no starters uploaded");
        KPMX_REQUIRE(M >= 2 && M % 2 == 0, "num_moments must be even and >= 2");
        KPMX_REQUIRE(use_graph >= 0 && use_graph <= 2, "run mode must be 0, 1 or 2");
        kpmx_ctx* ctx = b->ctx;
        KPMX_HIP(hipSetDevice(ctx->device));
        const int64_t steps = M / 2;
        if (!b->partials) b->partials = static_cast<double*>(dev_alloc((size_t)b->num_slots * 64 * 4 * sizeof(double)));
        if (b->sums_steps != steps || !b->sums) {
            KPMX_HIP(hipStreamSynchronize(ctx->stream));
            dev_free(b->sums);
            b->sums = static_cast<double*>(dev_alloc(batch_sums_bytes(b, steps)));
            b->sums_steps = steps;
            b->drop_graph();
        }
        if (use_graph == 2) {
            batch_run_persistent(b, steps);
            return;
        }
        if (!use_graph) {
            for (int64_t t = 0; t < b->ntiles; ++t) batch_enqueue(b, steps, t);
            KPMX_HIP(hipGetLastError());
            return;
        }
        if (b->graph_moments != M || !b->graph_exec) {
            b->drop_graph();
            KPMX_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            try {
                for (int64_t t = 0; t < b->ntiles; ++t) batch_enqueue(b, steps, t);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(ctx->stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            KPMX_HIP(hipStreamEndCapture(ctx->stream, &b->graph));
            KPMX_HIP(hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0));
            b->graph_moments = M;
        }
        KPMX_HIP(hipGraphLaunch(b->graph_exec, ctx->stream));
    });
}

int kpmx_batch_set_partition(kpmx_batch* b, int num_workgroups, int walk_chunk) {
    return guarded([&] {
        KPMX_REQUIRE(b, "null batch");
        KPMX_REQUIRE(num_workgroups >= 1 && walk_chunk >= 0, "invalid partition");
        KPMX_REQUIRE(walk_chunk == 0 || (num_workgroups >= 8 && num_workgroups % 8 == 0),
                     "the XCD-aware walk needs a multiple of 8 workgroups");
        KPMX_HIP(hipSetDevice(b->ctx->device));
        batch_partition(b, num_workgroups, walk_chunk);
    });
}

int kpmx_batch_set_tile_width(kpmx_batch* b, int width) {
    return guarded([&] {
        KPMX_REQUIRE(b, "null batch");
        KPMX_REQUIRE(width >= 0 && width <= 64, "tile width must be 0 (default) or 1 to 64 vectors");
        b->tile_cap = width;
    });
}

int kpmx_batch_tiles(const kpmx_batch* b, int64_t* num_tiles, int64_t* tile_vectors) {
    return guarded([&] {
        KPMX_REQUIRE(b, "null batch");
        KPMX_REQUIRE(b->r0, "no starters uploaded");
        if (num_tiles) *num_tiles = b->ntiles;
        if (tile_vectors) *tile_vectors = b->tile_cols;
    });
}

int kpmx_batch_set_value_table(kpmx_batch* b, int enable) {
    return guarded([&] {
        KPMX_REQUIRE(b, "null batch");
        KPMX_REQUIRE(enable == 0 || enable == 1, "enable must be 0 or 1");
        KPMX_REQUIRE(!enable || b->has_coded,
                     "this batch has no value-table form (a system has more than 255 distinct off-diagonal values or "
                     "a row with two diagonal entries)");
        if (b->use_coded != (enable == 1)) {
            KPMX_HIP(hipSetDevice(b->ctx->device));
            KPMX_HIP(hipStreamSynchronize(b->ctx->stream));
            b->drop_graph();
            b->use_coded = enable == 1;
        }
    });
}

int kpmx_batch_format(const kpmx_batch* b, int* available, int* enabled, int64_t* matrix_bytes) {
    return guarded([&] {
        KPMX_REQUIRE(b, "null batch");
        const int64_t sb = (int64_t)scalar_bytes(b->scalar);
        if (available) *available = b->has_coded ? 1 : 0;
        if (enabled) *enabled = b->use_coded ? 1 : 0;
        if (matrix_bytes)
            *matrix_bytes = b->use_coded ? b->c_entries * 5 + b->n_pad * sb + b->table_len * sb
                                         : b->s_entries * (4 + sb);
    });
}

int kpmx_batch_autotune(kpmx_batch* b, int* num_workgroups, int* walk_chunk) {
    return guarded([&] {
        KPMX_REQUIRE(b, "null batch");
        KPMX_REQUIRE(b->r0, "no starters uploaded");
        kpmx_ctx* ctx = b->ctx;
        KPMX_HIP(hipSetDevice(ctx->device));
        if (!ctx->t0) KPMX_HIP(hipEventCreate(&ctx->t0));
        if (!ctx->t1) KPMX_HIP(hipEventCreate(&ctx->t1));
        // about 1 GB of vector traffic per timed sample (at least 8 steps)
        const double step_bytes = 3.0 * (double)b->n_pad * b->vw * scalar_bytes(b->scalar);
        const int64_t steps = std::max<int64_t>(8, std::min<int64_t>(64, (int64_t)(1e9 / step_bytes)));
        KPMX_HIP(hipStreamSynchronize(ctx->stream));
        dev_free(b->sums);
        b->sums = static_cast<double*>(dev_alloc((size_t)steps * b->num_sys * 64 * 4 * sizeof(double)));
        b->sums_steps = -1;  // the timing runs leave no result to fetch
        // candidates: contiguous ranges over 1 ... 256 workgroups per CU (in steps of half a
        // workgroup per CU up to 4; the grids beyond the resident one leave the balance of uneven
        // ranges to the hardware's dispatch of workgroups as others finish), and the XCD-aware
        // walk on 1/2 ... 32 times the resident grid with runs of 1, 2 or 4 slices or one run per
        // walker (contiguous ranges inside each XCD's region)
        std::vector<std::pair<int, int>> cand;  // (workgroups, chunk)
        for (int half_per_cu : {2, 3, 4, 5, 6, 7, 8, 12, 16, 32, 64, 128, 256, 512}) {
            const int wg = (int)std::min<int64_t>((b->nslices + BATCH_WALKERS_PER_WG - 1) / BATCH_WALKERS_PER_WG,
                                                  (int64_t)ctx->num_cu * half_per_cu / 2);
            if (!cand.empty() && wg == cand.back().first) break;  // no more slices to spread
            cand.push_back({wg, 0});
        }
        BatchedLaunch probe;
        batch_launch(b, probe);
        const int resident = ctx->num_cu * batched_blocks_per_cu(probe);
        for (int64_t wg : {(int64_t)resident / 2, (int64_t)resident, 2 * (int64_t)resident, 4 * (int64_t)resident,
                           8 * (int64_t)resident, 16 * (int64_t)resident, 32 * (int64_t)resident}) {
            const int64_t walkers = wg * BATCH_WALKERS_PER_WG;
            if (wg < 8 || wg % 8 != 0 || walkers > b->nslices || wg > (1 << 24)) continue;
            const int one_run = (int)((b->nslices + walkers - 1) / walkers);
            for (int chunk : {1, 2, 4, one_run})  // runs longer than one_run are one run too
                if (chunk <= one_run && std::find(cand.begin(), cand.end(), std::make_pair((int)wg, chunk)) == cand.end())
                    cand.push_back({(int)wg, chunk});
        }
        int best = -1;
        float best_ms = 0;
        for (size_t c = 0; c < cand.size(); ++c) {
            batch_partition(b, cand[c].first, cand[c].second);
            b->partials = static_cast<double*>(dev_alloc((size_t)b->num_slots * 64 * 4 * sizeof(double)));
            float ms_w = 0;
            for (int rep = 0; rep < 3; ++rep) {  // the first repetition warms caches and code
                KPMX_HIP(hipEventRecord(ctx->t0, ctx->stream));
                batch_enqueue(b, steps);
                KPMX_HIP(hipGetLastError());
                KPMX_HIP(hipEventRecord(ctx->t1, ctx->stream));
                KPMX_HIP(hipEventSynchronize(ctx->t1));
                float ms = 0;
                KPMX_HIP(hipEventElapsedTime(&ms, ctx->t0, ctx->t1));
                if (rep == 1 || (rep > 1 && ms < ms_w)) ms_w = ms;
            }
            if (best < 0 || ms_w < best_ms) {
                best_ms = ms_w;
                best = (int)c;
            }
        }
        if (b->num_wg != cand[best].first || b->walk_chunk != cand[best].second)
            batch_partition(b, cand[best].first, cand[best].second);
        if (walk_chunk) *walk_chunk = b->walk_chunk;
        if (num_workgroups) *num_workgroups = b->num_wg;
    });
}

int kpmx_batch_fetch(kpmx_batch* b, int64_t M, void* mu) {
    return guarded([&] {
        KPMX_REQUIRE(b && mu, "null argument");
        KPMX_REQUIRE(b->sums && b->sums_steps == M / 2, "no finished run for this number of moments");
        kpmx_ctx* ctx = b->ctx;
        KPMX_HIP(hipSetDevice(ctx->device));
        const int width = 4 * b->vw;
        const int64_t steps = M / 2;
        std::vector<double> s(batch_sums_bytes(b, steps) / sizeof(double));
        KPMX_HIP(hipMemcpyAsync(s.data(), b->sums, s.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        KPMX_HIP(hipStreamSynchronize(ctx->stream));
        const bool cplx = is_complex(b->scalar);
        // per system and column tile: the diagonal collector over that system's sums; mu is S x V x M
        std::vector<double> per((size_t)steps * width);
        for (int64_t sy = 0; sy < b->num_sys; ++sy) {
            const size_t off = (size_t)sy * b->nvec * M * (cplx ? 2 : 1);
            for (int64_t t = 0; t < b->ntiles; ++t) {
                const double* st = s.data() + (size_t)t * steps * b->num_sys * width;
                for (int64_t k = 0; k < steps; ++k)
                    std::copy(st + ((size_t)k * b->num_sys + sy) * width, st + ((size_t)k * b->num_sys + sy) * width + width,
                              &per[(size_t)k * width]);
                const int64_t c0 = t * b->tile_cols, cols = std::min<int64_t>(b->tile_cols, b->nvec - c0);
                if (b->scalar == 0 || b->scalar == 2)
                    diag_collect<float>(per, b->vw, cols, M, cplx, static_cast<float*>(mu) + off, c0);
                else
                    diag_collect<double>(per, b->vw, cols, M, cplx, static_cast<double*>(mu) + off, c0);
            }
        }
    });
}

int kpmx_batch_destroy(kpmx_batch* b) {
    return guarded([&] {
        if (!b) return;
        (void)hipSetDevice(b->ctx->device);
        (void)hipStreamSynchronize(b->ctx->stream);
        delete b;
    });
}

int kpmx_adjoint_forward(const kpmx_mat* h2, const kpmx_block* r0, int64_t M, int64_t checkpoint, void* mu,
                         kpmx_tape** out) {
    return guarded([&] {
        check_pair(h2, r0);
        KPMX_REQUIRE(out, "null output");
        KPMX_REQUIRE(M >= 4 && M % 2 == 0, "num_moments must be even and >= 4");
        KPMX_REQUIRE(h2->scalar == KPMX_F32 || h2->scalar == KPMX_F64, "the adjoint supports real Hamiltonians only");
        KPMX_REQUIRE(h2->format == KPMX_FMT_SELL, "the adjoint needs the matrix uploaded in SELL format");
        KPMX_REQUIRE(r0->tiles.size() == 1, "the adjoint takes at most 64 vectors per block");
        kpmx_ctx* ctx = h2->ctx;
        KPMX_HIP(hipSetDevice(ctx->device));
        const int64_t n = h2->rows;
        const size_t sb = scalar_bytes(h2->scalar);
        auto tp = std::make_unique<kpmx_tape>();
        tp->h2 = h2;
        tp->M = M;
        tp->nvec = r0->nvec;
        const Tile& t0 = r0->tiles.front();
        Tile t = t0;
        if (t0.vw < 4) {  // the adjoint kernel works on blocks of >= 4 columns: pad with zero columns
            t.vw = 4;
            tp->padded = dev_alloc((size_t)n * 4 * sb);
            void* vm = ctx->aux0.get((size_t)n * t0.vw * sb);
            launch_scatter_block(h2->scalar, t0.data, n, t0.vw, 0, t0.vw, vm, ctx->stream);
            launch_gather_block(h2->scalar, vm, n, 0, t0.vw, 4, tp->padded, ctx->stream);
            KPMX_HIP(hipGetLastError());
            t.data = tp->padded;
        }
        const int vw = t.vw;
        tp->vw = vw;
        const size_t blk = (size_t)n * vw * sb;
        tp->blk = blk;
        const int64_t S = M / 2;  // vectors v_0 .. v_S
        tp->S = S;
        tp->C = (checkpoint <= 0 || checkpoint >= S) ? S : std::max<int64_t>(2, checkpoint);
        const int64_t C = tp->C;
        tp->store_all = (C >= S);
        // checkpointed: 3 rotating slots for the vectors the backward pass rebuilds, then v_S-1, v_S
        tp->nslots = tp->store_all ? S + 1 : 5;
        tp->ctx = ctx;
        tp->slots_bytes = (size_t)tp->nslots * blk;
        auto& pool = ctx->tape_pool;
        if (!pool.busy && pool.bytes >= tp->slots_bytes) {
            tp->slots = pool.ptr;
            tp->pooled = true;
            pool.busy = true;
        } else {
            if (!pool.busy && pool.ptr) {  // too small for this tape: give it back first
                KPMX_HIP(hipStreamSynchronize(ctx->stream));
                dev_free(pool.ptr);
                pool.ptr = nullptr;
                pool.bytes = 0;
            }
            tp->slots = static_cast<char*>(dev_alloc(tp->slots_bytes));
        }
        tp->nck = tp->store_all ? 0 : (S + C - 1) / C;
        if (tp->nck) tp->cks = static_cast<char*>(dev_alloc((size_t)tp->nck * 2 * blk));
        double* sums = static_cast<double*>(ctx->sums.get((size_t)S * 4 * vw * sizeof(double)));
        kpmx_tape* T = tp.get();
        if (T->store_all) {
            KPMX_HIP(hipMemcpyAsync(T->slot(0), t.data, blk, hipMemcpyDeviceToDevice, ctx->stream));
            run_recurrence(
                h2, t, MODE_DIAG, S, nullptr, sums,
                [&](StepLaunch& L, int64_t k) {
                    L.x = T->slot(k);
                    L.yin = (k == 0) ? nullptr : T->slot(k - 1);
                    L.yout = T->slot(k + 1);
                },
                [](StepLaunch const&, int64_t) {});
        } else {
            KPMX_HIP(hipMemcpyAsync(T->cks, t.data, blk, hipMemcpyDeviceToDevice, ctx->stream));
            run_recurrence(h2, t, MODE_DIAG, S, nullptr, sums, [](StepLaunch&, int64_t) {},
                           [&](StepLaunch const& L, int64_t k) {
                               // after step k, L.yout holds v_{k+1}: keep the pairs (v_cC, v_cC+1)
                               const int64_t j = k + 1;
                               if (j % C == 0 && j / C < T->nck)
                                   KPMX_HIP(hipMemcpyAsync(T->cks + (size_t)(j / C) * 2 * blk, L.yout, blk,
                                                           hipMemcpyDeviceToDevice, ctx->stream));
                               if (j % C == 1 && j / C < T->nck)
                                   KPMX_HIP(hipMemcpyAsync(T->cks + ((size_t)(j / C) * 2 + 1) * blk, L.yout, blk,
                                                           hipMemcpyDeviceToDevice, ctx->stream));
                               if (k == S - 1) {  // the last pair starts the backward pass
                                   KPMX_HIP(hipMemcpyAsync(T->slot(3), L.x, blk, hipMemcpyDeviceToDevice, ctx->stream));
                                   KPMX_HIP(hipMemcpyAsync(T->slot(4), L.yout, blk, hipMemcpyDeviceToDevice,
                                                           ctx->stream));
                               }
                           });
        }
        std::vector<double> hs((size_t)S * 4 * vw);
        KPMX_HIP(hipMemcpyAsync(hs.data(), sums, hs.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        KPMX_HIP(hipStreamSynchronize(ctx->stream));
        if (mu) {
            if (h2->scalar == KPMX_F32) diag_collect<float>(hs, vw, r0->nvec, M, false, static_cast<float*>(mu), 0);
            else diag_collect<double>(hs, vw, r0->nvec, M, false, static_cast<double*>(mu), 0);
        }
        *out = tp.release();
    });
}

int kpmx_adjoint_backward(kpmx_tape* T, const double* w, double* grad) {
    return guarded([&] {
        KPMX_REQUIRE(T && w && grad, "null argument");
        const kpmx_mat* h2 = T->h2;
        kpmx_ctx* ctx = h2->ctx;
        KPMX_HIP(hipSetDevice(ctx->device));
        const int64_t n = h2->rows, S = T->S, C = T->C;
        const size_t blk = T->blk;
        const int vw = T->vw;
        // L = sum_j alpha_j A_j + beta_j B_j with A_j = |v_j|^2 and B_j = <v_j+1|v_j>, from the
        // collector relations mu_0 = A_0/2, mu_1 = B_0, mu_2k = 2 A_k - A_0, mu_2k+1 = 2 B_k - B_0,
        // averaged over the R vectors of the block
        const double inv_r = 1.0 / (double)T->nvec;
        std::vector<double> alpha((size_t)S + 1, 0.0), beta((size_t)S + 1, 0.0);
        alpha[0] = 0.5 * w[0];
        beta[0] = w[1];
        for (int64_t k = 1; k < S; ++k) {
            alpha[0] -= w[2 * k];
            beta[0] -= w[2 * k + 1];
            alpha[(size_t)k] = 2.0 * w[2 * k];
            beta[(size_t)k] = 2.0 * w[2 * k + 1];
        }
        for (auto& v : alpha) v *= inv_r;
        for (auto& v : beta) v *= inv_r;

        const int64_t gsize = std::max<int64_t>(h2->sell_total, 1);
        double* dgrad = static_cast<double*>(ctx->aux2.get((size_t)gsize * sizeof(double)));
        launch_zero_f64(dgrad, gsize, ctx->stream);
        void* lamA = ctx->work0.get(blk);  // lambda_{j+2}, overwritten with lambda_j
        void* lamB = ctx->work1.get(blk);  // lambda_{j+1}
        launch_zero(h2->scalar, lamA, n * vw, ctx->stream);
        launch_zero(h2->scalar, lamB, n * vw, ctx->stream);
        AdjointLaunch A;
        A.scalar = h2->scalar;
        A.vw = vw;
        A.num_cu = ctx->num_cu;
        A.indices = h2->sell_indices;
        A.vals = h2->sell_vals;
        A.slice_ptr = h2->slice_ptr;
        A.slice_width = h2->slice_width;
        A.grad = dgrad;
        A.n = n;
        A.max_width = h2->max_width;
        // default: wave-staged entries on twice the resident grid (155.8 ms against 160.5 ms for
        // the 1 M-site checkpointed gradient with LDS staging on the resident grid), fp64 blocks of
        // 8 or more vectors as column pairs
        const int staging = (h2->scalar == KPMX_F64 && vw >= 8) ? 2 : 1;
        A.tune = ctx->adjoint_variant >= 0 ? ctx->adjoint_variant
                                           : staging * NUM_TILE_SHAPES * NUM_GRID_VARIANTS + 3 * NUM_TILE_SHAPES;
        // With C_j = <v_j|H2|v_j>, B_0 = C_0 / 2 and B_j = C_j - B_j-1 (j >= 1, from
        // v_j+1 = H2 v_j - v_j-1), so L = sum_j alpha_j A_j + sum_{j>=1} gamma_j C_j + gamma_0 C_0 / 2
        // with gamma_j = sum_{i>=j} (-1)^(i-j) beta_i.  v_S no longer enters, and
        //   lambda_j = 2 alpha_j v_j + 2 gamma_j H2 v_j + H2 lambda_j+1 - lambda_j+2     (j = S-1 .. 1)
        //   dL/dH2   = sum_{j>=1} (lambda_j+1 + gamma_j v_j) v_j^T + (lambda_1 + gamma_0 v_0) v_0^T / 2
        std::vector<double> gamma((size_t)S + 1, 0.0);
        for (int64_t j = S - 1; j >= 0; --j) gamma[(size_t)j] = beta[(size_t)j] - gamma[(size_t)j + 1];

        // Checkpointed tapes rebuild v_j backwards inside the adjoint steps: step j forms H2 v_j
        // anyway, and v_j-1 = H2 v_j - v_j+1 holds for every j >= 1 (v_2 = H2 v_1 - v_0 included).
        // The rebuilt vectors restart from the stored ones wherever the forward pass kept v_j
        // (the last pair and the checkpoint pairs), so rounding drift spans at most one segment.
        auto stored = [&](int64_t j) -> char* {
            if (T->store_all) return T->slot(j);
            if (j == S) return T->slot(4);
            if (j == S - 1) return T->slot(3);
            if (j / C < T->nck && (j % C == 0 || j % C == 1))
                return T->cks + ((size_t)(j / C) * 2 + (size_t)(j % C)) * blk;
            return nullptr;
        };
        auto vec = [&](int64_t j) -> char* {  // v_j: stored, or rebuilt in rotating slot j mod 3
            char* p = stored(j);
            return p ? p : T->slot(j % 3);
        };
        // adjoint steps j = S-1 .. 1 (lambda_S = lambda_S+1 = 0)
        for (int64_t j = S - 1; j >= 1; --j) {
            A.lam1 = lamB;
            A.lam2 = lamA;
            A.lam_out = lamA;
            A.v_j = vec(j);
            A.v_next = nullptr;
            A.v_prev = nullptr;
            if (!stored(j - 1)) {
                A.v_next = vec(j + 1);
                A.v_prev = T->slot((j - 1) % 3);
            }
            A.ca = 2.0 * alpha[(size_t)j];
            A.chv = 2.0 * gamma[(size_t)j];
            A.cg = gamma[(size_t)j];
            A.gscale = 1.0;
            A.update = 1;
            if (launch_adjoint_step(A, ctx->stream) != 0) throw Error{KPMX_EINVAL, "unsupported adjoint configuration"};
            KPMX_HIP(hipGetLastError());
            std::swap(lamA, lamB);  // lamB <- lambda_j, lamA <- lambda_{j+1}
        }
        // v_1 = 0.5 H2 v_0 and C_0 / 2:  grad += 0.5 (lambda_1 + gamma_0 v_0) v_0^T   (lamB holds lambda_1)
        A.v_next = nullptr;
        A.v_prev = nullptr;
        A.lam1 = lamB;
        A.lam2 = lamA;
        A.lam_out = lamA;
        A.v_j = vec(0);
        A.ca = A.chv = 0.0;
        A.cg = gamma[0];
        A.gscale = 0.5;
        A.update = 0;
        if (launch_adjoint_step(A, ctx->stream) != 0) throw Error{KPMX_EINVAL, "unsupported adjoint configuration"};
        KPMX_HIP(hipGetLastError());
        // SELL order -> CSR order on the device, then one download
        if (!h2->d_sell_pos && h2->nnz) {
            h2->d_sell_pos = static_cast<int64_t*>(dev_alloc((size_t)h2->nnz * sizeof(int64_t)));
            KPMX_HIP(hipMemcpyAsync(h2->d_sell_pos, h2->sell_pos.data(), (size_t)h2->nnz * sizeof(int64_t),
                                    hipMemcpyHostToDevice, ctx->stream));
        }
        double* csr_grad = static_cast<double*>(ctx->aux1.get((size_t)std::max<int64_t>(h2->nnz, 1) * sizeof(double)));
        launch_permute_f64(dgrad, h2->d_sell_pos, h2->nnz, csr_grad, ctx->stream);
        KPMX_HIP(hipGetLastError());
        KPMX_HIP(hipMemcpyAsync(grad, csr_grad, (size_t)h2->nnz * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        KPMX_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int kpmx_tape_destroy(kpmx_tape* T) {
    return guarded([&] {
        if (!T) return;
        (void)hipSetDevice(T->h2->ctx->device);
        (void)hipStreamSynchronize(T->h2->ctx->stream);
        delete T;
    });
}

int kpmx_adjoint_diagonal(const kpmx_mat* h2, const kpmx_block* r0, int64_t M, const double* w, int64_t checkpoint,
                          double* grad, void* mu) {
    kpmx_tape* T = nullptr;
    int rc = kpmx_adjoint_forward(h2, r0, M, checkpoint, mu, &T);
    if (rc != KPMX_OK) return rc;
    rc = kpmx_adjoint_backward(T, w, grad);
    const std::string err = g_last_error;
    kpmx_tape_destroy(T);
    if (rc != KPMX_OK) g_last_error = err;
    return rc;
}


int kpmx_chebyshev_cos_series(kpmx_ctx* ctx, int real, int64_t num_moments, int64_t cols, const void* mu,
                              int64_t num_energy, const void* theta, const void* pref, double* out) {
    return guarded([&] { cheb_series(ctx, real, false, num_moments, cols, mu, num_energy, theta, pref, out); });
}

int kpmx_chebyshev_exp_series(kpmx_ctx* ctx, int real, int64_t num_moments, int64_t cols, const void* mu,
                              int64_t num_energy, const void* theta, const void* pref, double* out) {
    return guarded([&] { cheb_series(ctx, real, true, num_moments, cols, mu, num_energy, theta, pref, out); });
}

int kpmx_kubo_bastin_sums(kpmx_ctx* ctx, int64_t num_moments, const double* mu_nm, int64_t num_samples,
                          const double* samples, double* out) {
    return guarded([&] {
        KPMX_REQUIRE(ctx && mu_nm && samples && out, "null argument");
        KPMX_REQUIRE(num_moments >= 1 && num_samples >= 0, "empty moments");
        KPMX_REQUIRE(num_samples <= 0x7fffffff, "too many energy samples for one call");
        KPMX_HIP(hipSetDevice(ctx->device));
        if (num_samples == 0) return;
        const int64_t M = num_moments, P = num_samples;
        const int groups = kubo_sums_groups(M, P);
        auto d = upload_pieces(ctx, ctx->aux0, {{mu_nm, (size_t)(M * M) * 2 * sizeof(double)},
                                                {samples, (size_t)P * sizeof(double)}});
        auto w = upload_pieces(ctx, ctx->aux1, {{nullptr, (size_t)(P * M) * sizeof(double)},
                                                {nullptr, (size_t)(P * M) * 2 * sizeof(double)},
                                                {nullptr, (size_t)(P * groups) * 2 * sizeof(double)},
                                                {nullptr, (size_t)P * 2 * sizeof(double)}});
        double* dout = reinterpret_cast<double*>(w[3]);
        launch_kubo_sums(reinterpret_cast<const double*>(d[0]), M, reinterpret_cast<const double*>(d[1]), P,
                         reinterpret_cast<double*>(w[0]), reinterpret_cast<double*>(w[1]),
                         reinterpret_cast<double*>(w[2]), groups, dout, ctx->stream);
        KPMX_HIP(hipGetLastError());
        KPMX_HIP(hipMemcpyAsync(out, dout, (size_t)P * 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        KPMX_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int kpmx_stream_probe(kpmx_ctx* ctx, int64_t bytes, double* copy_gbps, double* read_gbps, double* triad_gbps) {
    return guarded([&] {
        KPMX_REQUIRE(ctx && copy_gbps && read_gbps, "null argument");
        KPMX_REQUIRE(bytes >= (1 << 20) && bytes % 16 == 0, "bytes must be a multiple of 16, at least 1 MiB");
        KPMX_HIP(hipSetDevice(ctx->device));
        void* a = dev_alloc((size_t)bytes);
        void* b = dev_alloc((size_t)bytes);
        struct Free {
            void *a, *b;
            ~Free() {
                dev_free(a);
                dev_free(b);
            }
        } guard{a, b};
        KPMX_HIP(hipMemsetAsync(a, 0, (size_t)bytes, ctx->stream));
        KPMX_HIP(hipMemsetAsync(b, 0, (size_t)bytes, ctx->stream));
        if (!ctx->t0) KPMX_HIP(hipEventCreate(&ctx->t0));
        if (!ctx->t1) KPMX_HIP(hipEventCreate(&ctx->t1));
        double best[3] = {0, 0, 0};
        for (int kind = 0; kind < (triad_gbps ? 3 : 2); ++kind)
            for (int per_cu : {2, 4, 8, 16})
                for (int rep = 0; rep < 3; ++rep) {
                    KPMX_HIP(hipEventRecord(ctx->t0, ctx->stream));
                    launch_probe(kind, a, b, bytes, per_cu * ctx->num_cu, ctx->stream);
                    KPMX_HIP(hipGetLastError());
                    KPMX_HIP(hipEventRecord(ctx->t1, ctx->stream));
                    KPMX_HIP(hipEventSynchronize(ctx->t1));
                    float ms = 0;
                    KPMX_HIP(hipEventElapsedTime(&ms, ctx->t0, ctx->t1));
                    const double moved = (kind == 0 ? 2.0 : kind == 1 ? 1.0 : 3.0) * (double)bytes;
                    if (rep > 0 && ms > 0) best[kind] = std::max(best[kind], moved / (ms * 1e-3) / 1e9);
                }
        *copy_gbps = best[0];
        *read_gbps = best[1];
        if (triad_gbps) *triad_gbps = best[2];
    });
}

}  // extern "C"
