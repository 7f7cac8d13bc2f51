// Starters, moment containers, the compute interface with its HIP backend, the KPM Core and the
// reconstruction functions (host side of the engine).
#pragma once
#include <thread>
#include <mutex>
#include <variant>

#include "kpm_core.hpp"

namespace kpmx_host {

// ------------------------------------------------------------------------------------- Starter
// constant / unit / random starters (cppcore/include/kpm/Starter.hpp:12-34,
// cppcore/src/kpm/Starter.cpp:9-97).  Each make() returns one vector in optimized order.
template <class T>
struct Starter {
    std::function<std::vector<T>()> make;
    idx_t vector_size = 0;
    mutable idx_t count = 0;
};

template <class T>
Starter<T> constant_starter(OptimizedHamiltonian<T> const& oh, std::vector<cd> const& alpha) {
    return {[&oh, alpha]() {
                std::vector<T> r0(alpha.size());
                for (size_t i = 0; i < alpha.size(); ++i) r0[i] = from_cd<T>(alpha[i]);
                oh.reorder(r0);
                return r0;
            },
            oh.size()};
}

template <class T>
Starter<T> unit_starter(OptimizedHamiltonian<T> const& oh) {
    auto sources = std::make_shared<std::vector<int32_t>>(oh.optimized_idx.src);
    auto i = std::make_shared<size_t>(0);
    const idx_t size = oh.size();
    return {[sources, i, size]() {
                std::vector<T> r0((size_t)size, T(0));
                if (*i < sources->size()) {
                    r0[(size_t)(*sources)[*i]] = T(1);
                    ++*i;
                }
                return r0;
            },
            size};
}

// Random starter: std::mt19937 default seed, std::uniform_real_distribution<real_t>; real:
// x < 0.5f ? -1 : 1, complex: exp(i 2 pi_f x); optional op applied before the reorder.
template <class T>
Starter<T> random_starter(OptimizedHamiltonian<T> const& oh, Csr<T> const* op = nullptr) {
    using R = real_t<T>;
    auto gen = std::make_shared<std::mt19937>();
    std::shared_ptr<Csr<T>> opc = op ? std::make_shared<Csr<T>>(*op) : nullptr;
    const idx_t size = oh.size();
    return {[&oh, gen, opc, size]() {
                std::vector<T> r0((size_t)size);
                std::uniform_real_distribution<R> dist;
                for (idx_t i = 0; i < size; ++i) {
                    const R x = dist(*gen);
                    if constexpr (is_cplx<T>) {
                        const auto k = T{2 * PI_F * std::complex<float>(0, 1)};
                        r0[(size_t)i] = std::exp(k * x);
                    } else {
                        r0[(size_t)i] = (x < 0.5f) ? R(-1.f) : R(1.f);
                    }
                }
                if (opc) {
                    std::vector<T> tmp((size_t)size, T(0));
                    for (idx_t r = 0; r < size; ++r) {
                        T acc = T(0);
                        for (int32_t k = opc->indptr[r]; k < opc->indptr[r + 1]; ++k) acc += opc->data[k] * r0[opc->indices[k]];
                        tmp[(size_t)r] = acc;
                    }
                    r0.swap(tmp);
                }
                oh.reorder(r0);
                return r0;
            },
            size};
}

// ------------------------------------------------------------------------------------- moments
template <class T>
struct DiagonalMoments {
    idx_t num_moments;
    std::vector<T> data;
};

// collect(result, new_moments (M x cols, column-major), first_index, num_vectors)
template <class T>
struct BatchDiagonalMoments {
    idx_t num_moments, num_vectors;
    bool accumulate;  // true: BatchAccumulator (mean), false: BatchConcatenator (columns)
    std::vector<T> data;  // accumulate: M; concatenate: M x num_vectors column-major
};

template <class T>
struct GenericMoments {
    idx_t num_moments;
    std::vector<cd> const* alpha;
    std::vector<cd> const* beta;
    Csr<T> const* op;  // null: identity
    std::vector<T> data;
};

template <class T>
struct MultiUnitMoments {
    idx_t num_moments;
    Indices const* idx;
    std::vector<std::vector<T>> data;
};

template <class T>
using MomentsRef = std::variant<DiagonalMoments<T>*, BatchDiagonalMoments<T>*, GenericMoments<T>*, MultiUnitMoments<T>*>;

template <class T>
void apply_damping(std::vector<T>& m, Kernel const& k, idx_t rows, idx_t cols = 1) {
    const auto g = k.damping_coefficients(rows);
    for (idx_t c = 0; c < cols; ++c)
        for (idx_t r = 0; r < rows; ++r) m[(size_t)(c * rows + r)] *= (real_t<T>)g[(size_t)r];
}

// ------------------------------------------------------------------------------------- compute
using ProgressCallback = std::function<void(idx_t, idx_t)>;

// One GPU context per device, shared by every Core on that device.
class Device {
public:
    explicit Device(int device) {
        check(kpmx_ctx_create(device, &ctx_));
    }
    ~Device() { kpmx_ctx_destroy(ctx_); }
    kpmx_ctx* get() const { return ctx_; }
    static std::shared_ptr<Device> instance(int device) {
        static std::mutex mu;
        static std::vector<std::weak_ptr<Device>> devices(64);
        std::lock_guard<std::mutex> lk(mu);
        if (device < 0 || device >= 64) throw std::invalid_argument("invalid device index");
        auto d = devices[(size_t)device].lock();
        if (!d) {
            d = std::make_shared<Device>(device);
            devices[(size_t)device] = d;
        }
        return d;
    }

private:
    kpmx_ctx* ctx_ = nullptr;
};

struct MatHandle {
    kpmx_mat* m = nullptr;
    MatHandle() = default;
    MatHandle(MatHandle const&) = delete;
    MatHandle& operator=(MatHandle const&) = delete;
    ~MatHandle() { if (m) kpmx_mat_destroy(m); }
};
struct BlockHandle {
    kpmx_block* b = nullptr;
    BlockHandle() = default;
    BlockHandle(BlockHandle const&) = delete;
    BlockHandle& operator=(BlockHandle const&) = delete;
    ~BlockHandle() { if (b) kpmx_block_destroy(b); }
};

template <class T>
std::unique_ptr<MatHandle> upload(kpmx_ctx* ctx, Csr<T> const& m, int format) {
    auto h = std::make_unique<MatHandle>();
    check(kpmx_mat_upload(ctx, scalar_code<T>(), format, m.rows, m.nnz(), m.indptr.data(), m.indices.data(),
                          m.data.data(), &h->m));
    return h;
}

template <class T>
std::unique_ptr<BlockHandle> upload_block(kpmx_ctx* ctx, idx_t n, idx_t nvec, std::vector<T> const& r0) {
    auto h = std::make_unique<BlockHandle>();
    check(kpmx_block_upload(ctx, scalar_code<T>(), n, nvec, r0.data(), &h->b));
    return h;
}

// HipCompute: the compute plugin (Compute::Interface::moments, cppcore/include/kpm/Core.hpp:21-38)
// implemented with the engine's C-ABI.  The scaled H2 is uploaded once per optimization and
// reused by every later call.
template <class T>
class HipCompute {
public:
    HipCompute(int device, ProgressCallback progress, bool autotune = false)
        : dev_(Device::instance(device)), progress_(std::move(progress)), autotune_(autotune) {}

    kpmx_ctx* ctx() const { return dev_->get(); }

    kpmx_mat* matrix(OptimizedHamiltonian<T> const& oh, int format) {
        if (!h2_ || version_ != oh.version || format_ != format || oh_ != &oh) {
            h2_.reset();
            h2_ = upload(ctx(), oh.h2, format);
            version_ = oh.version;
            format_ = format;
            oh_ = &oh;
        }
        return h2_->m;
    }

    void moments(MomentsRef<T> m, Starter<T> const& s, AlgorithmConfig const&, OptimizedHamiltonian<T> const& oh,
                 int format) {
        kpmx_mat* h2 = matrix(oh, format);
        const idx_t n = oh.size();
        std::visit(
            [&](auto* mm) {
                using M = std::decay_t<decltype(*mm)>;
                if constexpr (std::is_same<M, DiagonalMoments<T>>::value) {
                    auto r0 = s.make();
                    ++s.count;
                    auto blk = upload_block(ctx(), n, 1, r0);
                    mm->data.assign((size_t)mm->num_moments, T(0));
                    check(kpmx_moments_diagonal(h2, blk->b, mm->num_moments, mm->data.data()));
                } else if constexpr (std::is_same<M, BatchDiagonalMoments<T>>::value) {
                    const idx_t M_ = mm->num_moments, R = mm->num_vectors;
                    report(-1, R);
                    if (mm->accumulate) mm->data.assign((size_t)M_, T(0));
                    else mm->data.assign((size_t)(M_ * R), T(0));
                    // vectors in groups of up to 64 (one device block per group)
                    std::vector<T> acc((size_t)M_, T(0));
                    for (idx_t v0 = 0; v0 < R; v0 += 64) {
                        const idx_t nv = std::min<idx_t>(64, R - v0);
                        std::vector<T> r0((size_t)(n * nv));
                        for (idx_t v = 0; v < nv; ++v) {
                            auto col = s.make();
                            ++s.count;
                            std::copy(col.begin(), col.end(), r0.begin() + v * n);
                        }
                        auto blk = upload_block(ctx(), n, nv, r0);
                        if (autotune_ && nv >= 4) {
                            int lw = 0;
                            while ((idx_t(1) << lw) < nv) ++lw;   // the block width: next power of two
                            if (!tuned_[lw]) {
                                check(kpmx_autotune_diagonal(h2, blk->b, nullptr));
                                tuned_[lw] = true;
                            }
                        }
                        std::vector<T> mu((size_t)(nv * M_));
                        check(kpmx_moments_diagonal(h2, blk->b, M_, mu.data()));
                        for (idx_t v = 0; v < nv; ++v) {
                            if (mm->accumulate) {
                                for (idx_t k = 0; k < M_; ++k) acc[(size_t)k] += mu[(size_t)(v * M_ + k)];
                            } else {
                                std::copy(mu.begin() + v * M_, mu.begin() + (v + 1) * M_, mm->data.begin() + (v0 + v) * M_);
                            }
                        }
                        report(nv, R);
                    }
                    if (mm->accumulate) {
                        const real_t<T> div = (R != 1) ? (real_t<T>)R : real_t<T>(1);
                        for (idx_t k = 0; k < M_; ++k) mm->data[(size_t)k] = acc[(size_t)k] / div;
                    }
                    report(R, R);
                } else if constexpr (std::is_same<M, GenericMoments<T>>::value) {
                    auto r0 = s.make();
                    ++s.count;
                    auto blk = upload_block(ctx(), n, 1, r0);
                    // beta defaults to alpha; both reordered like the starter (collectors.cpp:36-47)
                    std::vector<cd> const& b = (mm->beta && !mm->beta->empty()) ? *mm->beta : *mm->alpha;
                    std::vector<T> beta(b.size());
                    for (size_t i = 0; i < b.size(); ++i) beta[i] = from_cd<T>(b[i]);
                    oh.reorder(beta);
                    mm->data.assign((size_t)mm->num_moments, T(0));
                    if (mm->op && mm->op->nnz() > 0) {
                        Csr<T> op = oh.reorder(*mm->op);
                        check(kpmx_moments_generic(h2, blk->b, mm->num_moments, beta.data(), op.nnz(), op.indptr.data(),
                                                   op.indices.data(), op.data.data(), mm->data.data()));
                    } else {
                        check(kpmx_moments_generic(h2, blk->b, mm->num_moments, beta.data(), 0, nullptr, nullptr, nullptr,
                                                   mm->data.data()));
                    }
                } else {
                    auto r0 = s.make();
                    ++s.count;
                    auto blk = upload_block(ctx(), n, 1, r0);
                    const auto& dest = mm->idx->dest;
                    std::vector<T> out(dest.size() * (size_t)mm->num_moments);
                    check(kpmx_moments_multi_unit(h2, blk->b, mm->num_moments, (idx_t)dest.size(), dest.data(), out.data()));
                    mm->data.assign(dest.size(), std::vector<T>());
                    for (size_t i = 0; i < dest.size(); ++i)
                        mm->data[i].assign(out.begin() + i * mm->num_moments, out.begin() + (i + 1) * mm->num_moments);
                }
            },
            m);
    }

    // Kubo-Bastin mu_nm on the device (L/R panels + contraction never leave the GPU).
    std::vector<T> kubo(OptimizedHamiltonian<T> const& oh, int format, Csr<T> const& vl, Csr<T> const& vr,
                        idx_t num_moments, idx_t num_random, int mode) {
        kpmx_mat* h2 = matrix(oh, format);
        const idx_t n = oh.size();
        auto vlr = upload(ctx(), oh.reorder(vl), format);
        auto vrr = upload(ctx(), oh.reorder(vr), format);
        // Both starters are default-seeded random starters, so they draw the same vectors
        // (cppcore/src/kpm/Core.cpp:132-137); vl is applied on the device.
        auto st = random_starter(oh);
        std::vector<T> r0((size_t)(n * num_random));
        for (idx_t v = 0; v < num_random; ++v) {
            auto col = st.make();
            std::copy(col.begin(), col.end(), r0.begin() + v * n);
        }
        auto blk = upload_block(ctx(), n, num_random, r0);
        std::vector<T> mu((size_t)(num_moments * num_moments));
        check(kpmx_kubo_moments(h2, vlr->m, vrr->m, blk->b, num_moments, mode, mu.data()));
        return mu;
    }

    void report(idx_t delta, idx_t total) {
        if (progress_) progress_(delta, total);
    }

private:
    std::shared_ptr<Device> dev_;
    ProgressCallback progress_;
    bool autotune_ = false;
    bool tuned_[7] = {};  // block widths 1 .. 64 already tuned on this matrix
    std::unique_ptr<MatHandle> h2_;
    uint64_t version_ = 0;
    int format_ = -1;
    OptimizedHamiltonian<T> const* oh_ = nullptr;
};

// ------------------------------------------------------------------------------------- reconstruct
// SpectralDensity (cppcore/include/kpm/reconstruct.hpp:16-54): out[e * cols + c].  The scaled
// energies, their arccosines and the prefactors are formed here in the moments' real type, as the
// reference does; the E x M x cols cosine sums run on the device (kpmx_chebyshev_cos_series).
template <class R> constexpr int real_code() { return std::is_same<R, float>::value ? KPMX_F32 : KPMX_F64; }

template <class T>
std::vector<double> spectral_density(kpmx_ctx* ctx, std::vector<T> const& moments, idx_t M, idx_t cols,
                                     std::vector<double> const& energy, Scale<> s) {
    using R = real_t<T>;
    const Scale<R> scale(s);
    const R k = R(2 / PI_F) / scale.a;
    std::vector<R> mu(moments.size()), th(energy.size()), pref(energy.size());
    for (size_t i = 0; i < mu.size(); ++i) mu[i] = (R)re(moments[i]);
    for (size_t i = 0; i < energy.size(); ++i) {
        const R E = scale((R)energy[i]);
        th[i] = std::acos(E);
        pref[i] = k / std::sqrt(1 - E * E);
    }
    std::vector<double> out(energy.size() * (size_t)cols);
    check(kpmx_chebyshev_cos_series(ctx, real_code<R>(), M, cols, mu.data(), (int64_t)energy.size(), th.data(),
                                    pref.data(), out.data()));
    return out;
}

// GreensFunction (reconstruct.hpp:59-85); the exponential sums run on the device
// (kpmx_chebyshev_exp_series).
template <class T>
std::vector<cd> greens_function(kpmx_ctx* ctx, std::vector<T> const& moments, std::vector<double> const& energy,
                                Scale<> s) {
    using R = real_t<T>;
    using C = std::complex<R>;
    const Scale<R> scale(s);
    const C i1(0, 1);
    const C k = -R(2) * i1 / scale.a;
    const idx_t M = (idx_t)moments.size();
    std::vector<C> mu((size_t)M), pref(energy.size());
    std::vector<R> th(energy.size());
    for (idx_t n = 0; n < M; ++n) mu[(size_t)n] = C(moments[(size_t)n]);
    for (size_t i = 0; i < energy.size(); ++i) {
        const R E = scale((R)energy[i]);
        th[i] = std::acos(E);
        pref[i] = k / std::sqrt(1 - E * E);
    }
    std::vector<cd> out(energy.size());
    check(kpmx_chebyshev_exp_series(ctx, real_code<R>(), M, 1, mu.data(), (int64_t)energy.size(), th.data(),
                                    pref.data(), reinterpret_cast<double*>(out.data())));
    return out;
}

// KuboBastin (reconstruct.hpp:90-143); mu is M x M row-major.
// The reference builds the M x M matrix Gamma_nm = g(n,m) + conj(g(m,n)) with
// g(n,m) = (E - i m sqrt(1-E^2)) e^{i m acos E} T_n(E) for every energy sample, i.e. M^2 complex
// exponentials per sample.  g factorises into T_n(E) a_m(E) with a_m = (E - i m sqrt(1-E^2))
// e^{i m acos E}, so
//     sum_nm mu_nm Gamma_nm = sum_n T_n (mu a)_n + sum_n conj(a_n) (mu t)_n,
// two matrix-vector products per sample, formed on the device for all samples at once
// (kpmx_kubo_bastin_sums) in double precision (the reference uses the model's real type).  The
// Fermi-Dirac trapezoid over the samples stays here.
template <class T>
std::vector<cd> kubo_bastin(kpmx_ctx* ctx, std::vector<T> const& mu, idx_t M,
                            std::vector<double> const& chemical_potential, std::vector<double> const& energy_samples,
                            double temperature, Scale<> s) {
    using R = real_t<T>;
    const Scale<R> scale(s);
    const double inv_kbt = (double)static_cast<R>(scale.a / (KB_F * temperature));
    std::vector<double> es(energy_samples.size());
    for (size_t i = 0; i < es.size(); ++i) es[i] = (double)scale((R)energy_samples[i]);
    std::vector<cd> mud((size_t)(M * M));
    for (size_t i = 0; i < mud.size(); ++i) mud[i] = cd(mu[i]);
    std::vector<cd> sum_nm(es.size());
    check(kpmx_kubo_bastin_sums(ctx, M, reinterpret_cast<const double*>(mud.data()), (int64_t)es.size(), es.data(),
                                reinterpret_cast<double*>(sum_nm.data())));
    const size_t P = es.size();
    double emin = *std::min_element(es.begin(), es.end()), emax = *std::max_element(es.begin(), es.end());
    const double coeff = (emax - emin) / static_cast<double>(2 * es.size());
    const double prefix = 4.0 / ((double)scale.a * (double)scale.a);
    std::vector<cd> out(chemical_potential.size());
    for (size_t j = 0; j < chemical_potential.size(); ++j) {
        const double mu_ = (double)scale((R)chemical_potential[j]);
        auto fd = [&](double e) { return 1.0 / (1.0 + std::exp((e - mu_) * inv_kbt)); };
        cd total(0, 0);
        for (size_t q = 0; q < P; ++q) total += fd(es[q]) * sum_nm[q];
        const cd integ = coeff * (2.0 * total - fd(es.front()) * sum_nm.front() - fd(es.back()) * sum_nm.back());
        out[j] = prefix * integ;
    }
    return out;
}

// ------------------------------------------------------------------------------------- Core
// kpm::Core (cppcore/include/kpm/Core.hpp:40-89, cppcore/src/kpm/Core.cpp:17-156)
class CoreBase {
public:
    virtual ~CoreBase() = default;
    virtual Scale<> scaling_factors() = 0;
    virtual std::vector<cd> moments(idx_t num_moments, std::vector<cd> const& alpha, std::vector<cd> const& beta,
                                    Csr<cd> const& op) = 0;
    virtual std::vector<double> ldos(std::vector<idx_t> const& idx, std::vector<double> const& energy, double broadening,
                                     idx_t* cols_out) = 0;
    virtual std::vector<double> dos(std::vector<double> const& energy, double broadening, idx_t num_random) = 0;
    virtual std::vector<std::vector<cd>> greens_vector(idx_t row, std::vector<idx_t> const& cols,
                                                       std::vector<double> const& energy, double broadening) = 0;
    virtual std::vector<cd> conductivity(std::vector<float> const& left, std::vector<float> const& right,
                                         std::vector<double> const& chemical_potential, double broadening,
                                         double temperature, idx_t num_random, idx_t num_points) = 0;
    virtual std::vector<cd> dos_moments(idx_t num_moments, idx_t num_random) = 0;
    virtual Stats const& stats() const = 0;
    virtual Config const& config() const = 0;
    virtual idx_t size() const = 0;
    virtual bool is_complex() const = 0;
    virtual double min_energy() = 0;
    virtual double max_energy() = 0;
    virtual int lanczos_loops() const = 0;
    virtual void set_locality_order(std::vector<int32_t> order) = 0;
};

}  // namespace kpmx_host
