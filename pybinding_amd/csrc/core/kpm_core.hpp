// Host C++ core of the KPM engine: the parts of the reference's kpm::Core that run on the host
// (scaling, optimized Hamiltonian with BFS reorder, starters, damping kernels, collectors,
// reconstruction, stats), with a compute backend `HipCompute` that calls the engine's C-ABI
// (include/kpmx.h).  There is no CPU compute backend: every moment is computed on the GPU.
#pragma once
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <memory>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/kpmx.h"

namespace kpmx_host {

using idx_t = int64_t;
using cd = std::complex<double>;

// constant::pi, constant::kb in float (cppcore/include/numeric/constant.hpp:8,20)
constexpr float PI_F = 3.14159265358979323846f;
constexpr float KB_F = 8.6173303e-5f;

template <class T> struct real_of { using type = T; };
template <class R> struct real_of<std::complex<R>> { using type = R; };
template <class T> using real_t = typename real_of<T>::type;
template <class T> constexpr bool is_cplx = !std::is_same<T, real_t<T>>::value;

template <class T> constexpr int scalar_code();
template <> constexpr int scalar_code<float>() { return KPMX_F32; }
template <> constexpr int scalar_code<double>() { return KPMX_F64; }
template <> constexpr int scalar_code<std::complex<float>>() { return KPMX_C64; }
template <> constexpr int scalar_code<std::complex<double>>() { return KPMX_C128; }

template <class T> T from_cd(cd v);
template <> inline float from_cd<float>(cd v) { return (float)v.real(); }
template <> inline double from_cd<double>(cd v) { return v.real(); }
template <> inline std::complex<float> from_cd<std::complex<float>>(cd v) { return std::complex<float>((float)v.real(), (float)v.imag()); }
template <> inline cd from_cd<cd>(cd v) { return v; }

inline double re(double v) { return v; }
inline double re(float v) { return v; }
template <class R> inline double re(std::complex<R> v) { return v.real(); }

inline void check(int rc) {
    if (rc != KPMX_OK) throw std::runtime_error(std::string("kpmx: ") + kpmx_last_error());
}

// ------------------------------------------------------------------------------------- CSR
template <class T>
struct Csr {
    idx_t rows = 0, cols = 0;
    std::vector<int32_t> indptr{0};
    std::vector<int32_t> indices;
    std::vector<T> data;
    idx_t nnz() const { return (idx_t)data.size(); }
    bool empty() const { return rows == 0; }
};

template <class T>
idx_t max_nnz_per_row(Csr<T> const& m) {
    idx_t mx = 0;
    for (idx_t r = 0; r < m.rows; ++r) mx = std::max<idx_t>(mx, m.indptr[r + 1] - m.indptr[r]);
    return mx;
}

// ------------------------------------------------------------------------------------- Scale
// Scale<real_t> (cppcore/include/kpm/Bounds.hpp:10-33)
template <class R = double>
struct Scale {
    R a = 0, b = 0;
    Scale() = default;
    Scale(R min_energy, R max_energy)
        : a(R(0.5f * (max_energy - min_energy) * (1 + 0.01f))), b(R(0.5f * (max_energy + min_energy))) {
        if (std::abs(b / a) < 0.01f * 0.01f) b = 0;
    }
    template <class U>
    explicit Scale(Scale<U> const& o) : a(static_cast<R>(o.a)), b(static_cast<R>(o.b)) {}
    R operator()(R v) const { return (v - b) / a; }
};

// ------------------------------------------------------------------------------------- Kernel
// round_num_moments, Jackson / Lorentz / Dirichlet (cppcore/include/kpm/Kernel.hpp:9-57,
// cppcore/src/kpm/Kernel.cpp:6-49)
inline idx_t round_num_moments(idx_t n) {
    if (n < 2) return 2;
    while ((n - 2) % 4 != 0) ++n;
    return n;
}

struct Kernel {
    std::function<std::vector<double>(idx_t)> damping_coefficients;
    std::function<idx_t(double)> required_num_moments;
};

inline Kernel jackson_kernel() {
    return {[](idx_t num_moments) {
                const double N = (double)num_moments, Np = N + 1;
                const double pi = (double)PI_F;
                std::vector<double> g((size_t)num_moments);
                for (idx_t i = 0; i < num_moments; ++i) {
                    const double n = (double)i;
                    g[i] = ((Np - n) * std::cos(pi * n / Np) + std::sin(pi * n / Np) / std::tan(pi / Np)) / Np;
                }
                return g;
            },
            [](double scaled_broadening) {
                return round_num_moments(static_cast<idx_t>(PI_F / scaled_broadening) + 1);
            }};
}

inline Kernel lorentz_kernel(double lambda = 4.0) {
    if (lambda <= 0) throw std::invalid_argument("Lorentz kernel: lambda must be positive.");
    return {[=](idx_t num_moments) {
                const double N = (double)num_moments;
                std::vector<double> g((size_t)num_moments);
                for (idx_t i = 0; i < num_moments; ++i) g[i] = std::sinh(lambda * (1 - (double)i / N)) / std::sinh(lambda);
                return g;
            },
            [=](double scaled_broadening) {
                return round_num_moments(static_cast<idx_t>(lambda / scaled_broadening) + 1);
            }};
}

inline Kernel dirichlet_kernel() {
    return {[](idx_t num_moments) { return std::vector<double>((size_t)num_moments, 1.0); },
            [](double scaled_broadening) {
                return round_num_moments(static_cast<idx_t>(PI_F / scaled_broadening) + 1);
            }};
}

// ------------------------------------------------------------------------------------- Config
enum class MatrixFormat { CSR, ELL };

struct AlgorithmConfig {
    bool optimal_size = true;
    bool interleaved = true;
    bool reorder() const { return optimal_size || interleaved; }
};

struct Config {
    float min_energy = 0.0f;
    float max_energy = 0.0f;
    Kernel kernel = jackson_kernel();
    MatrixFormat matrix_format = MatrixFormat::ELL;
    AlgorithmConfig algorithm;
    float lanczos_precision = 0.002f;
    int device = 0;
    bool autotune = false;  // time the diagonal step variants on the first block of each width
};

// ------------------------------------------------------------------------------------- Indices
struct Indices {
    std::vector<int32_t> src, dest;
    Indices() = default;
    Indices(idx_t s, idx_t d) : src{(int32_t)s}, dest{(int32_t)d} {}
    Indices(idx_t s, std::vector<idx_t> const& d) : src{(int32_t)s} {
        for (auto v : d) dest.push_back((int32_t)v);
    }
    Indices(std::vector<idx_t> const& s, std::vector<idx_t> const& d) {
        for (auto v : s) src.push_back((int32_t)v);
        for (auto v : d) dest.push_back((int32_t)v);
    }
    bool is_diagonal() const { return src == dest; }
    bool operator==(Indices const& o) const { return src == o.src && dest == o.dest; }
};

// SliceMap (cppcore/include/kpm/OptimizedHamiltonian.hpp:54-96, OptimizedHamiltonian.cpp:5-19)
class SliceMap {
public:
    std::vector<int32_t> data;
    idx_t src_offset = 0, dest_offset = 0;

    SliceMap() = default;
    explicit SliceMap(idx_t system_size) { data = {(int32_t)system_size}; }
    SliceMap(std::vector<int32_t> borders, Indices const& optimized_idx) : data(std::move(borders)) {
        auto find_offset = [&](std::vector<int32_t> const& idx) {
            const auto mx = *std::max_element(idx.begin(), idx.end());
            auto it = std::find_if(data.begin(), data.end(), [&](int32_t v) { return v > mx; });
            return (idx_t)(it - data.begin());
        };
        src_offset = find_offset(optimized_idx.src);
        dest_offset = find_offset(optimized_idx.dest);
    }
    idx_t last_index() const { return (idx_t)data.size() - 1; }
    idx_t index(idx_t n, idx_t num_moments) const {
        const idx_t mid = (num_moments - 1 + dest_offset - src_offset) / 2;
        const idx_t mx = std::min(last_index(), mid + src_offset);
        if (n < mid) return std::min(mx, n + src_offset);
        return std::min(mx, num_moments - 1 - n + dest_offset);
    }
    idx_t optimal_size(idx_t n, idx_t num_moments) const { return data[(size_t)index(n, num_moments)]; }
    bool uses_full_system(idx_t num_moments) const { return (idx_t)data.size() < num_moments / 2; }
};

// ------------------------------------------------------------------------------------- timers
struct Chrono {
    double elapsed = 0;
    std::chrono::steady_clock::time_point t0;
    void tic() { t0 = std::chrono::steady_clock::now(); }
    void toc() { elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
    void toc_accumulate() { elapsed += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

// ------------------------------------------------------------------------------------- OptimizedHamiltonian
// Scaled (and optionally BFS-reordered) H2 = (H - b) * 2/a
// (cppcore/include/kpm/OptimizedHamiltonian.hpp:98-196, cppcore/src/kpm/OptimizedHamiltonian.cpp:21-242)
template <class T>
class OptimizedHamiltonian {
public:
    using R = real_t<T>;
    Csr<T> const* original = nullptr;
    bool is_reordered = true;
    bool has_idx = false;
    Indices original_idx, optimized_idx;
    Csr<T> h2;
    SliceMap slice_map;
    std::vector<int32_t> reorder_map;  // original index -> optimized index (empty: identity)
    std::vector<int32_t> locality_order;  // optional site order for whole-system calculations
    bool used_locality = false;
    uint64_t version = 0;
    Chrono timer;

    OptimizedHamiltonian() = default;
    OptimizedHamiltonian(Csr<T> const* h, bool reorder) : original(h), is_reordered(reorder), slice_map(h->rows) {}

    idx_t size() const { return original->rows; }

    // `whole_system`: the calculation never uses the slice map (DOS, moments, conductivity); the
    // GPU then prefers `locality_order` (e.g. sites along a space-filling curve) when one is set,
    // because the moments do not depend on the site order and a compact order keeps the gathered
    // vector rows of a sweep in L2.  Otherwise the reference's BFS reorder is used.
    void optimize_for(Indices const& idx, Scale<> scale, bool whole_system = false) {
        for (auto const* list : {&idx.src, &idx.dest})
            for (auto v : *list)
                if (v < 0 || (idx_t)v >= size())
                    throw std::runtime_error("KPM: site index " + std::to_string(v) + " is outside the system of " +
                                             std::to_string(size()) + " sites");
        const bool loc = whole_system && is_reordered && !locality_order.empty();
        if (has_idx && original_idx == idx && loc == used_locality) return;
        timer.tic();
        if (loc) create_ordered(idx, scale, locality_order);
        else if (is_reordered) create_reordered(idx, scale);
        else create_scaled(idx, scale);
        timer.toc();
        original_idx = idx;
        used_locality = loc;
        has_idx = true;
        ++version;
    }

    template <class V>
    void reorder(std::vector<V>& v) const {
        if (reorder_map.empty()) return;
        std::vector<V> out(v.size());
        for (size_t i = 0; i < v.size(); ++i) out[(size_t)reorder_map[i]] = v[i];
        v.swap(out);
    }

    Csr<T> reorder(Csr<T> const& m) const {
        if (reorder_map.empty()) return m;
        // permute rows and columns, keeping each row's entries sorted by new column
        Csr<T> out;
        out.rows = m.rows;
        out.cols = m.cols;
        out.indptr.assign((size_t)m.rows + 1, 0);
        std::vector<idx_t> inv((size_t)m.rows);
        for (idx_t i = 0; i < m.rows; ++i) inv[(size_t)reorder_map[i]] = i;
        for (idx_t r = 0; r < m.rows; ++r) out.indptr[r + 1] = out.indptr[r] + (m.indptr[inv[r] + 1] - m.indptr[inv[r]]);
        out.indices.resize(m.indices.size());
        out.data.resize(m.data.size());
        std::vector<std::pair<int32_t, T>> tmp;
        for (idx_t r = 0; r < m.rows; ++r) {
            const idx_t o = inv[r];
            tmp.clear();
            for (int32_t k = m.indptr[o]; k < m.indptr[o + 1]; ++k) tmp.push_back({reorder_map[m.indices[k]], m.data[k]});
            std::sort(tmp.begin(), tmp.end(), [](auto const& a, auto const& b) { return a.first < b.first; });
            for (size_t q = 0; q < tmp.size(); ++q) {
                out.indices[out.indptr[r] + q] = tmp[q].first;
                out.data[out.indptr[r] + q] = tmp[q].second;
            }
        }
        return out;
    }

    // Processed matrix entries (OptimizedHamiltonian.cpp:163-192): CSR counts the stored entries of
    // the rows a step covers, ELL counts rows x the ELL width (padding included).
    size_t num_nonzeros(idx_t num_moments, bool optimal_size, bool ell = false) const {
        size_t width = 0;
        if (ell)
            for (idx_t r = 0; r < h2.rows; ++r) width = std::max(width, (size_t)(h2.indptr[(size_t)r + 1] - h2.indptr[(size_t)r]));
        auto entries = [&](idx_t rows) { return ell ? (size_t)rows * width : (size_t)h2.indptr[(size_t)rows]; };
        size_t result = 0;
        if (!optimal_size) {
            result = (size_t)num_moments * entries(h2.rows);
        } else {
            for (idx_t n = 0; n < num_moments; ++n) result += entries(slice_map.optimal_size(n, num_moments));
        }
        if (optimized_idx.is_diagonal()) result /= 2;
        return result;
    }
    size_t num_vec_elements(idx_t num_moments, bool optimal_size) const {
        size_t result = 0;
        if (!optimal_size) result = (size_t)num_moments * (size_t)size();
        else
            for (idx_t n = 0; n < num_moments; ++n) result += (size_t)slice_map.optimal_size(n, num_moments);
        if (optimized_idx.is_diagonal()) result /= 2;
        return result;
    }
    size_t matrix_memory() const { return (size_t)h2.nnz() * (sizeof(T) + 4) + (size_t)(h2.rows + 1) * 4; }
    size_t vector_memory() const { return (size_t)size() * sizeof(T); }

private:
    // Scale and permute by a given order: new index of original site i is order[i].
    void create_ordered(Indices const& idx, Scale<> s, std::vector<int32_t> const& order) {
        create_scaled(idx, s);
        reorder_map = order;
        h2 = reorder(h2);
        optimized_idx = Indices();
        for (auto v : idx.src) optimized_idx.src.push_back(reorder_map[(size_t)v]);
        for (auto v : idx.dest) optimized_idx.dest.push_back(reorder_map[(size_t)v]);
        slice_map = SliceMap(h2.rows);
    }

    void create_scaled(Indices const& idx, Scale<> s) {
        const Scale<R> scale(s);
        auto const& h = *original;
        h2 = Csr<T>();
        h2.rows = h.rows;
        h2.cols = h.cols;
        h2.indptr.assign((size_t)h.rows + 1, 0);
        const R inv = R(2) / scale.a;
        for (idx_t r = 0; r < h.rows; ++r) {
            bool diag = false;
            for (int32_t k = h.indptr[r]; k < h.indptr[r + 1]; ++k) {
                T v = h.data[k];
                if (h.indices[k] == r) {
                    diag = true;
                    if (scale.b != 0) v = v - T(scale.b);
                }
                h2.indices.push_back(h.indices[k]);
                h2.data.push_back(v * inv);
            }
            if (scale.b != 0 && !diag) {
                // insert the missing diagonal element in column order
                const idx_t start = h2.indptr[r];
                auto pos = std::lower_bound(h2.indices.begin() + start, h2.indices.end(), (int32_t)r) - h2.indices.begin();
                h2.indices.insert(h2.indices.begin() + pos, (int32_t)r);
                h2.data.insert(h2.data.begin() + pos, T(-scale.b) * inv);
            }
            h2.indptr[r + 1] = (int32_t)h2.indices.size();
        }
        reorder_map.clear();
        optimized_idx = idx;
        slice_map = SliceMap(h.rows);
    }

    void create_reordered(Indices const& idx, Scale<> s) {
        const Scale<R> scale(s);
        auto const& h = *original;
        const idx_t n = h.rows;
        const R inv = R(2 / scale.a);
        h2 = Csr<T>();
        h2.rows = n;
        h2.cols = n;
        h2.indptr.assign((size_t)n + 1, 0);
        h2.indices.reserve(h.indices.size() + (scale.b != 0 ? (size_t)n : 0));
        h2.data.reserve(h2.indices.capacity());

        std::vector<int32_t> queue;
        queue.reserve((size_t)n);
        queue.push_back(idx.src[0]);
        reorder_map.assign((size_t)n, -1);
        reorder_map[(size_t)idx.src[0]] = 0;
        std::vector<int32_t> borders{1};
        std::vector<std::pair<int32_t, T>> row_buf;
        for (idx_t h2_row = 0; h2_row < n; ++h2_row) {
            if (h2_row >= (idx_t)queue.size()) {
                // disconnected graph: start a new BFS tree at the lowest unvisited index
                // (the reference assumes a connected Hamiltonian)
                for (idx_t q = 0; q < n; ++q)
                    if (reorder_map[(size_t)q] < 0) {
                        reorder_map[(size_t)q] = (int32_t)queue.size();
                        queue.push_back((int32_t)q);
                        break;
                    }
            }
            bool diag = false;
            const int32_t row = queue[(size_t)h2_row];
            row_buf.clear();
            for (int32_t k = h.indptr[row]; k < h.indptr[row + 1]; ++k) {
                const int32_t col = h.indices[k];
                if (reorder_map[(size_t)col] < 0) {
                    reorder_map[(size_t)col] = (int32_t)queue.size();
                    queue.push_back(col);
                }
                T v = h.data[k] * inv;
                if (row == col) {
                    v -= T(scale.b * inv);
                    diag = true;
                }
                row_buf.push_back({reorder_map[(size_t)col], v});
            }
            if (scale.b != 0 && !diag) row_buf.push_back({(int32_t)h2_row, T(-scale.b * inv)});
            std::sort(row_buf.begin(), row_buf.end(), [](auto const& a, auto const& b) { return a.first < b.first; });
            for (auto const& e : row_buf) {
                h2.indices.push_back(e.first);
                h2.data.push_back(e.second);
            }
            h2.indptr[(size_t)h2_row + 1] = (int32_t)h2.indices.size();
            if (h2_row == borders.back() - 1) borders.push_back((int32_t)queue.size());
        }
        borders.pop_back();
        if (borders.empty() || borders.back() != n) borders.push_back((int32_t)n);
        optimized_idx = Indices();
        for (auto v : idx.src) optimized_idx.src.push_back(reorder_map[(size_t)v]);
        for (auto v : idx.dest) optimized_idx.dest.push_back(reorder_map[(size_t)v]);
        slice_map = SliceMap(std::move(borders), optimized_idx);
    }
};

// ------------------------------------------------------------------------------------- Stats
struct Stats {
    idx_t num_moments = 0;
    bool uses_full_system = false;
    size_t nnz = 0, opt_nnz = 0, vec = 0, opt_vec = 0;
    double multiplier = 1;
    size_t matrix_memory = 0, vector_memory = 0;
    Chrono hamiltonian_timer, moments_timer;

    template <class T>
    void reset(idx_t nm, OptimizedHamiltonian<T> const& oh, AlgorithmConfig const& ac, idx_t mult = 1,
               MatrixFormat fmt = MatrixFormat::CSR) {
        num_moments = nm;
        uses_full_system = oh.slice_map.uses_full_system(nm);
        const bool ell = fmt == MatrixFormat::ELL;
        nnz = oh.num_nonzeros(nm, false, ell);
        opt_nnz = oh.num_nonzeros(nm, ac.optimal_size, ell);
        vec = oh.num_vec_elements(nm, false);
        opt_vec = oh.num_vec_elements(nm, ac.optimal_size);
        multiplier = (double)mult;
        matrix_memory = oh.matrix_memory();
        vector_memory = oh.vector_memory();
        hamiltonian_timer = oh.timer;
        moments_timer = Chrono();
    }
    // non-zeros per second (cppcore/src/kpm/Stats.cpp:49-51)
    double eps() const { return multiplier * (double)opt_nnz / moments_timer.elapsed; }
    double ops(bool is_diagonal, bool non_unit_vector) const {
        size_t operations = nnz * 2 + vec;
        if (is_diagonal) operations += vec * 4;
        else if (non_unit_vector) operations += vec * 2;
        return multiplier * (double)operations / moments_timer.elapsed;
    }
};

}  // namespace kpmx_host
