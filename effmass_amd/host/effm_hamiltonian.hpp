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

#include <complex>
#include <cstdint>
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

private:
    effm_ctx* ctx_ = nullptr;
};

}  // namespace effm
