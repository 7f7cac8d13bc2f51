// CoreT<T>: the KPM Core for one scalar type (cppcore/src/kpm/Core.cpp:17-156), computing every
// moment through HipCompute.
#pragma once
#include "kpm_engine.hpp"

namespace kpmx_host {

template <class T>
Csr<T> cast_csr(Csr<cd> const& m) {
    Csr<T> out;
    out.rows = m.rows;
    out.cols = m.cols;
    out.indptr = m.indptr;
    out.indices = m.indices;
    out.data.resize(m.data.size());
    for (size_t i = 0; i < m.data.size(); ++i) out.data[i] = from_cd<T>(m.data[i]);
    return out;
}

// v_ij = (x_i - x_j) H_ij with float positions (cppcore/src/kpm/Moments.cpp:132-156)
template <class T>
Csr<T> velocity(Csr<T> const& h, std::vector<float> const& x) {
    Csr<T> v = h;
    for (idx_t r = 0; r < h.rows; ++r)
        for (int32_t k = h.indptr[r]; k < h.indptr[r + 1]; ++k)
            v.data[k] *= static_cast<T>(x[(size_t)r] - x[(size_t)h.indices[k]]);
    return v;
}

template <class T>
class CoreT final : public CoreBase {
public:
    CoreT(Csr<T> h, Config cfg, ProgressCallback progress, int kubo_mode)
        : h_(std::move(h)), cfg_(std::move(cfg)), oh_(&h_, cfg_.algorithm.reorder()),
          compute_(cfg_.device, std::move(progress), cfg_.autotune), kubo_mode_(kubo_mode) {
        if (cfg_.min_energy > cfg_.max_energy) throw std::invalid_argument("KPM: Invalid energy range specified (min > max).");
        if (cfg_.min_energy != cfg_.max_energy) {
            min_ = cfg_.min_energy;
            max_ = cfg_.max_energy;
            bounds_known_ = true;
        }
    }

    Scale<> scaling_factors() override {
        compute_bounds();
        return Scale<>(min_, max_);
    }
    double min_energy() override { compute_bounds(); return min_; }
    double max_energy() override { compute_bounds(); return max_; }
    int lanczos_loops() const override { return loops_; }
    Stats const& stats() const override { return stats_; }
    Config const& config() const override { return cfg_; }
    idx_t size() const override { return h_.rows; }
    bool is_complex() const override { return is_cplx<T>; }

    std::vector<cd> moments(idx_t num_moments, std::vector<cd> const& alpha, std::vector<cd> const& beta,
                            Csr<cd> const& op) override {
        auto alg = cfg_.algorithm;
        alg.optimal_size = false;
        oh_.optimize_for({0, 0}, scaling_factors(), true);
        stats_.reset(num_moments, oh_, alg, 1, cfg_.matrix_format);
        auto starter = constant_starter(oh_, alpha);
        const idx_t M = round_num_moments(num_moments);
        std::vector<T> data;
        if (beta.empty() && op.empty()) {
            DiagonalMoments<T> m{M, {}};
            timed_compute(&m, starter, alg);
            data = std::move(m.data);
        } else {
            Csr<T> opt;
            if (!op.empty()) opt = cast_csr<T>(op);
            GenericMoments<T> m{M, &alpha, &beta, op.empty() ? nullptr : &opt, {}};
            timed_compute(&m, starter, alg);
            data = std::move(m.data);
        }
        apply_damping(data, cfg_.kernel, M);
        std::vector<cd> out((size_t)num_moments);
        for (idx_t i = 0; i < num_moments; ++i) out[(size_t)i] = cd(data[(size_t)i]);
        return out;
    }

    std::vector<double> ldos(std::vector<idx_t> const& idx, std::vector<double> const& energy, double broadening,
                             idx_t* cols_out) override {
        const auto scale = scaling_factors();
        const idx_t M = cfg_.kernel.required_num_moments(broadening / scale.a);
        const idx_t K = (idx_t)idx.size();
        oh_.optimize_for(Indices(idx, idx), scale);
        stats_.reset(M, oh_, cfg_.algorithm, K, cfg_.matrix_format);
        auto starter = unit_starter(oh_);
        BatchDiagonalMoments<T> m{M, K, false, {}};
        timed_compute(&m, starter, cfg_.algorithm);
        apply_damping(m.data, cfg_.kernel, M, K);
        *cols_out = K;
        return spectral_density(compute_.ctx(), m.data, M, K, energy, scale);
    }

    std::vector<double> dos(std::vector<double> const& energy, double broadening, idx_t num_random) override {
        const auto scale = scaling_factors();
        const idx_t M = cfg_.kernel.required_num_moments(broadening / scale.a);
        auto alg = cfg_.algorithm;
        alg.optimal_size = false;
        oh_.optimize_for({0, 0}, scale, true);
        stats_.reset(M, oh_, alg, num_random, cfg_.matrix_format);
        auto starter = random_starter(oh_);
        BatchDiagonalMoments<T> m{M, num_random, true, {}};
        timed_compute(&m, starter, alg);
        apply_damping(m.data, cfg_.kernel, M);
        return spectral_density(compute_.ctx(), m.data, M, 1, energy, scale);
    }

    // Undamped mean diagonal moments of `num_random` reference random starters (the DOS moments).
    std::vector<cd> dos_moments(idx_t num_moments, idx_t num_random) override {
        auto alg = cfg_.algorithm;
        alg.optimal_size = false;
        oh_.optimize_for({0, 0}, scaling_factors(), true);
        stats_.reset(num_moments, oh_, alg, num_random, cfg_.matrix_format);
        auto starter = random_starter(oh_);
        BatchDiagonalMoments<T> m{num_moments, num_random, true, {}};
        timed_compute(&m, starter, alg);
        std::vector<cd> out(m.data.size());
        for (size_t i = 0; i < out.size(); ++i) out[i] = cd(m.data[i]);
        return out;
    }

    std::vector<std::vector<cd>> greens_vector(idx_t row, std::vector<idx_t> const& cols,
                                               std::vector<double> const& energy, double broadening) override {
        if (cols.empty()) throw std::invalid_argument("KPM::calc_greens: no column indices given");
        const auto scale = scaling_factors();
        const idx_t M = cfg_.kernel.required_num_moments(broadening / scale.a);
        oh_.optimize_for(Indices(row, cols), scale);
        stats_.reset(M, oh_, cfg_.algorithm, 1, cfg_.matrix_format);
        if (oh_.optimized_idx.is_diagonal()) {
            DiagonalMoments<T> m{M, {}};
            timed_compute(&m, unit_starter(oh_), cfg_.algorithm);
            apply_damping(m.data, cfg_.kernel, M);
            return {greens_function(compute_.ctx(), m.data, energy, scale)};
        }
        MultiUnitMoments<T> m{M, &oh_.optimized_idx, {}};
        timed_compute(&m, unit_starter(oh_), cfg_.algorithm);
        std::vector<std::vector<cd>> out;
        for (auto& d : m.data) {
            apply_damping(d, cfg_.kernel, M);
            out.push_back(greens_function(compute_.ctx(), d, energy, scale));
        }
        return out;
    }

    std::vector<cd> conductivity(std::vector<float> const& left, std::vector<float> const& right,
                                 std::vector<double> const& chemical_potential, double broadening, double temperature,
                                 idx_t num_random, idx_t num_points) override {
        const auto scale = scaling_factors();
        const idx_t M = cfg_.kernel.required_num_moments(broadening / scale.a);
        auto alg = cfg_.algorithm;
        alg.optimal_size = false;
        oh_.optimize_for({0, 0}, scale, true);
        stats_.reset(M, oh_, alg, num_random, cfg_.matrix_format);
        const Csr<T> vl = velocity(h_, left), vr = velocity(h_, right);
        stats_.moments_timer.tic();
        auto mu = compute_.kubo(oh_, format(), vl, vr, M, num_random, kubo_mode_);
        stats_.moments_timer.toc_accumulate();
        // 2-D damping g_n g_m (cppcore/include/kpm/Kernel.hpp:47-56)
        const auto g = cfg_.kernel.damping_coefficients(M);
        for (idx_t n = 0; n < M; ++n)
            for (idx_t m = 0; m < M; ++m) mu[(size_t)(n * M + m)] *= (real_t<T>)(g[(size_t)n]) * (real_t<T>)(g[(size_t)m]);
        // Bounds::linspaced(num_points) (cppcore/include/kpm/Bounds.hpp:50)
        std::vector<double> samples((size_t)num_points);
        for (idx_t i = 0; i < num_points; ++i)
            samples[(size_t)i] = (num_points == 1) ? max_ : min_ + (max_ - min_) * (double)i / (double)(num_points - 1);
        return kubo_bastin(compute_.ctx(), mu, M, chemical_potential, samples, temperature, scale);
    }

    void set_locality_order(std::vector<int32_t> order) override {
        if (!order.empty() && (idx_t)order.size() != h_.rows) throw std::invalid_argument("site order has the wrong size");
        oh_.locality_order = std::move(order);
    }

    // Raw access for the engine's own tools
    OptimizedHamiltonian<T>& optimized() { return oh_; }
    HipCompute<T>& compute() { return compute_; }
    Csr<T> const& hamiltonian() const { return h_; }
    int format() const { return cfg_.matrix_format == MatrixFormat::ELL ? KPMX_FMT_SELL : KPMX_FMT_CSR; }

private:
    void compute_bounds() {
        if (bounds_known_) return;
        auto m = upload(compute_.ctx(), h_, KPMX_FMT_CSR);
        double mn = 0, mx = 0;
        int loops = 0;
        check(kpmx_lanczos_bounds(m->m, (double)cfg_.lanczos_precision, &mn, &mx, &loops));
        min_ = mn;
        max_ = mx;
        loops_ = loops;
        bounds_known_ = true;
    }

    template <class M>
    void timed_compute(M* m, Starter<T> const& starter, AlgorithmConfig const& ac) {
        stats_.moments_timer.tic();
        compute_.moments(MomentsRef<T>(m), starter, ac, oh_, format());
        stats_.moments_timer.toc_accumulate();
    }

    Csr<T> h_;
    Config cfg_;
    OptimizedHamiltonian<T> oh_;
    HipCompute<T> compute_;
    int kubo_mode_ = KPMX_KUBO_FP64;
    Stats stats_;
    double min_ = 0, max_ = 0;
    bool bounds_known_ = false;
    int loops_ = 0;
};

}  // namespace kpmx_host
