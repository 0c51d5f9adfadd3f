// effm_hamiltonian.hpp — C++ mirror of the reference's H_psi module API over
// the C-ABI of libeffm_hip.so (include/effm_hpsi.h).
//
// Reference: src/H_psi/hamiltonian.cppm:115-363.  The class and method names,
// argument meaning and error behaviour are kept so that reference callers
// (test/H_psi/test_hpsi_eigen.cpp:88, test_hgrad_eigen_fd.cpp:191) switch by
// changing the include:
//
//   effm::Hamiltonian h(cell);            // finalize()d cell data
//   h.loadKPoint(ikpt, kpoint);           // GKK::loadKPoint
//   auto op = h.at_k(ikpt);
//   op(psi, hpsi);                        // hpsi += H psi   (one band)
//   op(n_bands, psi, hpsi);               // batched form, band-major
//   auto g = h.gradient().at_k(ikpt);
//   g(psi, effm::GradHPsi{dx, dy, dz});   // empty span = skip direction
//   auto hs = h.hessian().at_k(ikpt);
//   hs(psi, effm::HessHPsi{xx, yy, zz, xy, xz, yz});   // d2H/dk dk |psi>
//
// Spans are host memory, complex<double> (the reference's FP64 path) or
// complex<float> (the mixed path: FP32 FFT / GEMM, FP64 D contraction); the
// facade stages them through the device.  Callers whose bands already live
// in HBM use the device() forms (device pointers, band-major with leading
// dimensions, either precision, ordered on the stream given to
// Hamiltonian::set_stream).  Status codes are rethrown as the reference's
// exception classes: std::runtime_error, std::out_of_range,
// std::domain_error, std::invalid_argument, std::bad_alloc-like
// std::runtime_error for NOMEM.
#pragma once

#include <algorithm>
#include <complex>
#include <cstdint>
#include <map>
#include <span>
#include <stdexcept>
#include <string>
#include <vector>

#include "effm_hpsi.h"

namespace effm {

struct GradHPsi {
    std::span<std::complex<double>> x;
    std::span<std::complex<double>> y;
    std::span<std::complex<double>> z;
};

// The same output container for the mixed path (complex64 bands).
struct GradHPsi32 {
    std::span<std::complex<float>> x;
    std::span<std::complex<float>> y;
    std::span<std::complex<float>> z;
};

// d2H/dk_a dk_b |psi> outputs (reference: hamiltonian.cppm:84-91); an empty
// span skips that component
struct HessHPsi {
    std::span<std::complex<double>> xx;
    std::span<std::complex<double>> yy;
    std::span<std::complex<double>> zz;
    std::span<std::complex<double>> xy;
    std::span<std::complex<double>> xz;
    std::span<std::complex<double>> yz;
};

enum class Precision : int { FP64 = EFFM_PREC_FP64, Mixed = EFFM_PREC_MIXED };

struct SpeciesData {
    int lmax = -1;
    std::vector<int32_t> beta_l;
    std::vector<double> lambda;       // diagonalised l-block D (Ha)
    int nq = 0;
    double dq = 0.01;
    double qmax = 20.0;
    std::vector<double> beta_tab;     // [nbeta*nq]
    std::vector<double> dbeta_tab;    // [nbeta*nq] or empty
};

struct CellData {
    int n1 = 0, n2 = 0, n3 = 0;
    double lattice[9] = {0};          // rows a1, a2, a3 (bohr)
    std::vector<double> vloc;         // [n1*n2*n3] Ha
    std::vector<int32_t> species;     // [natom]
    std::vector<double> tau;          // [natom*3] fractional
    std::vector<SpeciesData> sp;
};

struct KPointData {
    std::vector<double> K;            // [npw*3] k+G Cartesian
    std::vector<double> kinetic;      // [npw]
    std::vector<int32_t> gidx;        // [npw*3]
    double kfrac[3] = {0, 0, 0};
    int npw() const { return static_cast<int>(kinetic.size()); }
};

inline void throw_status(int rc, const effm_ctx* ctx, const char* what) {
    if (rc >= 0) return;
    std::string msg = std::string(what) + ": " + (ctx ? effm_last_error(ctx) : "");
    switch (rc) {
        case EFFM_ERR_OUT_OF_RANGE: throw std::out_of_range(msg);
        case EFFM_ERR_DOMAIN: throw std::domain_error(msg);
        case EFFM_ERR_INVALID: throw std::invalid_argument(msg);
        default: throw std::runtime_error(msg);
    }
}

class Hamiltonian {
public:
    explicit Hamiltonian(const CellData& cell, int device = 0, uint64_t hbm_budget = 0) {
        throw_status(effm_ctx_create(&ctx_, device, hbm_budget), nullptr, "effm_ctx_create");
        std::vector<effm_species> sps(cell.sp.size());
        for (size_t i = 0; i < cell.sp.size(); ++i) {
            const auto& s = cell.sp[i];
            sps[i] = effm_species{s.lmax, static_cast<int32_t>(s.beta_l.size()), s.beta_l.data(),
                                  s.lambda.data(), s.nq, s.dq, s.qmax, s.beta_tab.data(),
                                  s.dbeta_tab.empty() ? nullptr : s.dbeta_tab.data()};
        }
        for (int i = 0; i < 9; ++i) lattice_[i] = cell.lattice[i];
        effm_cell c{};
        c.n1 = cell.n1; c.n2 = cell.n2; c.n3 = cell.n3;
        for (int i = 0; i < 9; ++i) c.lattice[i] = cell.lattice[i];
        c.vloc = cell.vloc.data();
        c.natom = static_cast<int32_t>(cell.species.size());
        c.species = cell.species.data();
        c.tau = cell.tau.data();
        c.nspecies = static_cast<int32_t>(sps.size());
        c.sp = sps.data();
        int rc = effm_cell_upload(ctx_, &c);
        if (rc) {
            std::string m = effm_last_error(ctx_);
            effm_ctx_destroy(ctx_);
            ctx_ = nullptr;
            throw_status(rc, nullptr, ("effm_cell_upload: " + m).c_str());
        }
    }
    Hamiltonian(const Hamiltonian&) = delete;
    Hamiltonian& operator=(const Hamiltonian&) = delete;
    ~Hamiltonian() { if (ctx_) effm_ctx_destroy(ctx_); }

    void loadKPoint(int ikpt, const KPointData& k) {
        throw_status(effm_kpt_upload(ctx_, ikpt, k.npw(), k.K.data(), k.kinetic.data(), k.gidx.data(), k.kfrac),
                     ctx_, "loadKPoint");
        kpts_[ikpt] = k;          // host copy: the Hessian's shifted k-points are built from it
    }

    effm_ctx* ctx() const { return ctx_; }
    // order every later device() call on this HIP stream (hipStream_t)
    void set_stream(void* stream) { throw_status(effm_set_stream(ctx_, stream), ctx_, "set_stream"); }

    // -- 1. Callable: H|psi> at fixed k (hamiltonian.cppm:173-205) ---------
    class Callable {
    public:
        Callable() = default;
        void operator()(std::span<const std::complex<double>> psi,
                        std::span<std::complex<double>> hpsi) const {
            (*this)(1, psi, hpsi);
        }
        void operator()(int n_bands, std::span<const std::complex<double>> psi,
                        std::span<std::complex<double>> hpsi) const {
            host(n_bands, psi.data(), psi.size(), hpsi.data(), hpsi.size(), EFFM_PREC_FP64);
        }
        // mixed path on complex64 host spans
        void operator()(std::span<const std::complex<float>> psi, std::span<std::complex<float>> hpsi) const {
            (*this)(1, psi, hpsi);
        }
        void operator()(int n_bands, std::span<const std::complex<float>> psi,
                        std::span<std::complex<float>> hpsi) const {
            host(n_bands, psi.data(), psi.size(), hpsi.data(), hpsi.size(), EFFM_PREC_MIXED);
        }
        // device pointers (band b at psi + b * ld complex elements of the
        // precision's type); hpsi += H psi unless overwrite
        void device(int n_bands, const void* psi, int64_t ld, void* hpsi, int64_t ld_out, Precision p,
                    bool overwrite = false) const {
            throw_status(effm_hpsi(ctx_, ikpt_, n_bands, psi, ld, hpsi, ld_out, static_cast<int>(p),
                                   overwrite ? EFFM_F_OVERWRITE : 0),
                         ctx_, "H|psi>");
        }
        int dim() const { int n = effm_kpt_dim(ctx_, ikpt_); throw_status(n, ctx_, "Callable::dim"); return n; }
        void set_ikpt(int ikpt) { ikpt_ = ikpt; (void)dim(); }

    private:
        friend class Hamiltonian;
        Callable(effm_ctx* c, int ikpt) : ctx_(c), ikpt_(ikpt) { (void)dim(); }
        void host(int n_bands, const void* psi, size_t npsi, void* hpsi, size_t nh, int prec) const {
            const int ng = dim();
            if (n_bands < 0 || npsi < static_cast<size_t>(n_bands) * ng || nh < static_cast<size_t>(n_bands) * ng)
                throw std::invalid_argument("Callable: spans shorter than n_bands * dim()");
            throw_status(effm_hpsi(ctx_, ikpt_, n_bands, psi, ng, hpsi, ng, prec, EFFM_F_HOST), ctx_, "H|psi>");
        }
        effm_ctx* ctx_ = nullptr;
        int ikpt_ = 0;
    };
    Callable at_k(int ikpt) const { return Callable(ctx_, ikpt); }

    // -- 2. Gradient: dH/dk_alpha |psi> (hamiltonian.cppm:218-263) ----------
    class Gradient {
    public:
        class Callable {
        public:
            Callable() = default;
            void operator()(std::span<const std::complex<double>> psi, GradHPsi out) const {
                host(1, psi.data(), psi.size(), out.x, out.y, out.z, EFFM_PREC_FP64);
            }
            // batched bands (band-major, dim() apart) and the mixed path
            void operator()(int n_bands, std::span<const std::complex<double>> psi, GradHPsi out) const {
                host(n_bands, psi.data(), psi.size(), out.x, out.y, out.z, EFFM_PREC_FP64);
            }
            void operator()(std::span<const std::complex<float>> psi, GradHPsi32 out) const {
                host(1, psi.data(), psi.size(), out.x, out.y, out.z, EFFM_PREC_MIXED);
            }
            void operator()(int n_bands, std::span<const std::complex<float>> psi, GradHPsi32 out) const {
                host(n_bands, psi.data(), psi.size(), out.x, out.y, out.z, EFFM_PREC_MIXED);
            }
            // device pointers; a NULL output skips that direction
            void device(int n_bands, const void* psi, int64_t ld, void* x, void* y, void* z, int64_t ld_out,
                        Precision p, bool overwrite = false) const {
                throw_status(effm_dhdk(ctx_, ikpt_, n_bands, psi, ld, x, y, z, ld_out, static_cast<int>(p),
                                       overwrite ? EFFM_F_OVERWRITE : 0),
                             ctx_, "dH/dk|psi>");
            }
            int dim() const { int n = effm_kpt_dim(ctx_, ikpt_); throw_status(n, ctx_, "Gradient::dim"); return n; }

        private:
            friend class Gradient;
            Callable(effm_ctx* c, int ikpt) : ctx_(c), ikpt_(ikpt) { (void)dim(); }
            template <typename S>
            void host(int n_bands, const void* psi, size_t npsi, S x, S y, S z, int prec) const {
                const int ng = dim();
                const size_t need = static_cast<size_t>(n_bands) * ng;
                if (n_bands < 0 || npsi < need)
                    throw std::invalid_argument("Gradient::Callable: psi span shorter than n_bands * dim()");
                auto ptr = [&](S s) -> void* {
                    if (s.empty()) return nullptr;
                    if (s.size() < need) throw std::invalid_argument("GradHPsi span shorter than n_bands * dim()");
                    return s.data();
                };
                throw_status(effm_dhdk(ctx_, ikpt_, n_bands, psi, ng, ptr(x), ptr(y), ptr(z), ng, prec, EFFM_F_HOST),
                             ctx_, "dH/dk|psi>");
            }
            effm_ctx* ctx_ = nullptr;
            int ikpt_ = 0;
        };
        Callable at_k(int ikpt) const { return Callable(ctx_, ikpt); }

    private:
        friend class Hamiltonian;
        explicit Gradient(effm_ctx* c) : ctx_(c) {}
        effm_ctx* ctx_ = nullptr;
    };
    Gradient gradient() const { return Gradient(ctx_); }

    // -- 3. Hessian: d2H/dk_a dk_b |psi> (hamiltonian.cppm:267-306) ---------
    // The reference declares this operator and leaves its body empty
    // (hamiltonian_hessian.cpp:48-50).  Here d2H/dk_a dk_b = delta_ab +
    // d2V_NL/dk_a dk_b is the central difference over k_b of the analytic
    // dH/dk_a at k +- delta e_b (Cartesian, bohr^-1) on the same G set: exact
    // for the kinetic part (dH/dk_a = K_a there), O(delta^2) for the nonlocal
    // part; xy, xz, yz difference dH/dk_x, dH/dk_x, dH/dk_y along y, z, z.
    // The shifted k-points use one scratch k-point index past the highest
    // one loaded, released afterwards.  Results accumulate into the spans.
    class Hessian {
    public:
        class Callable {
        public:
            Callable() = default;
            void operator()(std::span<const std::complex<double>> psi, HessHPsi out) const {
                apply(1, psi, out);
            }
            // batched bands (band-major, dim() apart)
            void operator()(int n_bands, std::span<const std::complex<double>> psi, HessHPsi out) const {
                apply(n_bands, psi, out);
            }
            int dim() const {
                int n = effm_kpt_dim(h_->ctx_, ikpt_);
                throw_status(n, h_->ctx_, "Hessian::dim");
                return n;
            }
            void set_ikpt(int ikpt) { ikpt_ = ikpt; (void)dim(); }

        private:
            friend class Hessian;
            Callable(const Hamiltonian* h, int ikpt, double delta) : h_(h), ikpt_(ikpt), delta_(delta) { (void)dim(); }
            void apply(int n_bands, std::span<const std::complex<double>> psi, HessHPsi out) const {
                using cd = std::complex<double>;
                const int ng = dim();
                const size_t need = static_cast<size_t>(n_bands) * ng;
                if (n_bands < 0 || psi.size() < need)
                    throw std::invalid_argument("Hessian::Callable: psi span shorter than n_bands * dim()");
                std::span<cd>* comp[6] = {&out.xx, &out.yy, &out.zz, &out.xy, &out.xz, &out.yz};
                for (auto* c : comp)
                    if (!c->empty() && c->size() < need)
                        throw std::invalid_argument("HessHPsi span shorter than n_bands * dim()");
                auto it = h_->kpts_.find(ikpt_);
                if (it == h_->kpts_.end())
                    throw std::out_of_range("Hessian: k-point " + std::to_string(ikpt_) +
                                            " was not loaded through loadKPoint");
                const KPointData& kp = it->second;
                const int slot = h_->kpts_.rbegin()->first + 1;
                // shift direction b -> (component, a): xx; xy, yy; xz, yz, zz
                struct Part { int comp, a; };
                const std::vector<Part> plan[3] = {{{0, 0}}, {{3, 0}, {1, 1}}, {{4, 0}, {5, 1}, {2, 2}}};
                std::vector<cd> d[2][3];
                struct Release {
                    effm_ctx* c; int s; bool on = false;
                    ~Release() { if (on) (void)effm_kpt_release(c, s); }
                } rel{h_->ctx_, slot};
                for (int b = 0; b < 3; ++b) {
                    bool any = false;
                    bool dir[3] = {false, false, false};
                    for (const Part& q : plan[b])
                        if (!comp[q.comp]->empty()) any = dir[q.a] = true;
                    if (!any) continue;
                    for (int sg = 0; sg < 2; ++sg) {
                        const double dk = (sg == 0 ? 1.0 : -1.0) * delta_;
                        KPointData sk = kp;
                        for (int g = 0; g < sk.npw(); ++g) {
                            double* K = &sk.K[3 * static_cast<size_t>(g)];
                            K[b] += dk;
                            sk.kinetic[g] = 0.5 * (K[0] * K[0] + K[1] * K[1] + K[2] * K[2]);
                        }
                        // k_frac + dk . A^T / 2 pi (rows of A are the lattice vectors)
                        for (int i = 0; i < 3; ++i) sk.kfrac[i] += dk * h_->lattice_[3 * i + b] / (2.0 * kPi);
                        throw_status(effm_kpt_upload(h_->ctx_, slot, sk.npw(), sk.K.data(), sk.kinetic.data(),
                                                     sk.gidx.data(), sk.kfrac),
                                     h_->ctx_, "Hessian: shifted k-point");
                        rel.on = true;
                        void* ptr[3] = {nullptr, nullptr, nullptr};
                        for (int a = 0; a < 3; ++a)
                            if (dir[a]) {
                                d[sg][a].assign(need, cd(0.0));
                                ptr[a] = d[sg][a].data();
                            }
                        throw_status(effm_dhdk(h_->ctx_, slot, n_bands, psi.data(), ng, ptr[0], ptr[1], ptr[2], ng,
                                               EFFM_PREC_FP64, EFFM_F_HOST | EFFM_F_OVERWRITE),
                                     h_->ctx_, "Hessian: dH/dk at k +- delta");
                    }
                    const double s = 1.0 / (2.0 * delta_);
                    for (const Part& q : plan[b]) {
                        std::span<cd>& o = *comp[q.comp];
                        if (o.empty()) continue;
                        for (size_t i = 0; i < need; ++i) o[i] += (d[0][q.a][i] - d[1][q.a][i]) * s;
                    }
                }
            }
            static constexpr double kPi = 3.14159265358979323846;
            const Hamiltonian* h_ = nullptr;
            int ikpt_ = 0;
            double delta_ = 1e-3;
        };
        Callable at_k(int ikpt) const { return Callable(h_, ikpt, delta_); }

    private:
        friend class Hamiltonian;
        Hessian(const Hamiltonian* h, double delta) : h_(h), delta_(delta) {}
        const Hamiltonian* h_ = nullptr;
        double delta_ = 1e-3;
    };
    Hessian hessian(double delta = 1e-3) const { return Hessian(this, delta); }

private:
    effm_ctx* ctx_ = nullptr;
    double lattice_[9] = {0};
    std::map<int, KPointData> kpts_;
};

}  // namespace effm
