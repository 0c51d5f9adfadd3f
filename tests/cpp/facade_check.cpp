// Exercises the C++ facade (effmass_amd/host/effm_hamiltonian.hpp) the way the
// reference's tests drive H_psi::Hamiltonian: at_k(ikpt)(psi, hpsi) with
// accumulate semantics, gradient().at_k(ikpt)(psi, GradHPsi{...}) with a
// skipped direction, and the exception classes of the error paths; then the
// mixed-precision host spans, the device-pointer forms and
// hessian().at_k(ikpt)(psi, HessHPsi{...}).
// Closed-form case: simple cubic cell, constant V, one s projector
// beta(q) = exp(-q^2/4) on one atom, so H psi is known analytically.
// Built twice: with #include of the facade (C++20), and with
// -DEFFM_USE_MODULE against the C++23 module unit `import H_psi.hip;`.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <span>
#include <stdexcept>
#include <vector>

#ifdef EFFM_USE_MODULE
import H_psi.hip;
#else
#include "effm_hamiltonian.hpp"
#endif

using cd = std::complex<double>;
using cf = std::complex<float>;

int main() {
    const double a = 10.0, V0 = 0.1, lam = 0.5, PI = 3.14159265358979323846;
    const int n = 24;
    effm::CellData cell;
    cell.n1 = cell.n2 = cell.n3 = n;
    cell.lattice[0] = cell.lattice[4] = cell.lattice[8] = a;
    cell.vloc.assign((size_t)n * n * n, V0);
    cell.species = {0};
    cell.tau = {0.1, 0.2, 0.3};
    effm::SpeciesData sp;
    sp.lmax = 0;
    sp.beta_l = {0};
    sp.lambda = {lam};
    sp.nq = 2004;
    for (int i = 0; i < sp.nq; ++i) {
        double q = i * sp.dq;
        sp.beta_tab.push_back(std::exp(-q * q / 4));
        sp.dbeta_tab.push_back(-0.5 * q * std::exp(-q * q / 4));
    }
    cell.sp = {sp};
    effm::KPointData k;
    const double b = 2 * PI / a, ecut_ry = 10.0;
    for (int i = -n / 2; i < n / 2; ++i)
        for (int j = -n / 2; j < n / 2; ++j)
            for (int l = -n / 2; l < n / 2; ++l) {
                double K[3] = {i * b, j * b, l * b};
                double k2 = K[0] * K[0] + K[1] * K[1] + K[2] * K[2];
                if (k2 > ecut_ry) continue;
                k.K.insert(k.K.end(), K, K + 3);
                k.kinetic.push_back(0.5 * k2);
                k.gidx.insert(k.gidx.end(), {i, j, l});
            }
    const int npw = k.npw();
    effm::Hamiltonian h(cell);
    h.loadKPoint(0, k);
    std::vector<cd> psi(npw), hpsi(npw, cd(1.0, -1.0)), p(npw), dp[3];
    for (auto& v : dp) v.assign(npw, 0.0);
    const double omega = a * a * a, y00 = 0.5 / std::sqrt(PI);
    cd inner = 0, dinner[3] = {0, 0, 0};
    for (int g = 0; g < npw; ++g) {
        psi[g] = cd(std::cos(0.3 * g), std::sin(0.17 * g)) / (1.0 + k.kinetic[g]);
        const double* K = &k.K[3 * g];
        double q = std::sqrt(K[0] * K[0] + K[1] * K[1] + K[2] * K[2]);
        double ph = -2 * PI * (k.gidx[3 * g] * 0.1 + k.gidx[3 * g + 1] * 0.2 + k.gidx[3 * g + 2] * 0.3);
        cd S(std::cos(ph), std::sin(ph));
        p[g] = std::exp(-q * q / 4) / std::sqrt(omega) * y00 * S;
        inner += std::conj(p[g]) * psi[g];
        for (int d = 0; d < 3; ++d) {
            dp[d][g] = (-0.5 * q * std::exp(-q * q / 4)) * (q > 0 ? K[d] / q : 0.0) / std::sqrt(omega) * y00 * S;
            dinner[d] += std::conj(dp[d][g]) * psi[g];
        }
    }
    auto op = h.at_k(0);
    op(psi, hpsi);                                     // accumulate onto (1 - i)
    double err = 0, ref = 0;
    for (int g = 0; g < npw; ++g) {
        cd e = cd(1.0, -1.0) + (k.kinetic[g] + V0) * psi[g] + lam * inner * p[g];
        err = std::max(err, std::abs(hpsi[g] - e));
        ref = std::max(ref, std::abs(e));
    }
    std::printf("facade H|psi> max err %.3e (npw %d)\n", err, npw);
    if (err > 1e-9 * ref) return 1;
    std::vector<cd> gx(npw), gz(npw);
    h.gradient().at_k(0)(psi, effm::GradHPsi{gx, {}, gz});
    double gerr = 0;
    for (int g = 0; g < npw; ++g) {
        for (int d : {0, 2}) {
            cd e = k.K[3 * g + d] * psi[g] + lam * (dinner[d] * p[g] + inner * dp[d][g]);
            gerr = std::max(gerr, std::abs((d == 0 ? gx : gz)[g] - e));
        }
    }
    std::printf("facade dH/dk max err %.3e\n", gerr);
    if (gerr > 1e-8) return 2;
    try {
        h.at_k(7);
        return 3;
    } catch (const std::out_of_range&) {
    }
    try {
        effm::KPointData far = k;
        for (auto& v : far.K) v *= 8.0;
        for (int g = 0; g < npw; ++g) far.kinetic[g] *= 64.0;
        h.loadKPoint(1, far);
        return 4;
    } catch (const std::domain_error&) {
    }
    // mixed path on complex64 host spans: closed form within FP32 rounding
    std::vector<cf> psi32(npw), h32(npw), gx32(npw), gz32(npw);
    for (int g = 0; g < npw; ++g) psi32[g] = cf((float)psi[g].real(), (float)psi[g].imag());
    op(std::span<const cf>(psi32), std::span<cf>(h32));
    h.gradient().at_k(0)(std::span<const cf>(psi32), effm::GradHPsi32{gx32, {}, gz32});
    double merr = 0, mgerr = 0, gref = 0;
    for (int g = 0; g < npw; ++g) {
        cd e = (k.kinetic[g] + V0) * psi[g] + lam * inner * p[g];
        merr = std::max(merr, std::abs(cd(h32[g]) - e));
        for (int d : {0, 2}) {
            cd ge = k.K[3 * g + d] * psi[g] + lam * (dinner[d] * p[g] + inner * dp[d][g]);
            mgerr = std::max(mgerr, std::abs(cd((d == 0 ? gx32 : gz32)[g]) - ge));
            gref = std::max(gref, std::abs(ge));
        }
    }
    std::printf("facade mixed H|psi> max err %.3e, dH/dk %.3e\n", merr, mgerr);
    if (merr > 1e-5 * ref || mgerr > 1e-5 * gref) return 5;
    // device pointers, FP64, overwrite; two bands (the second is 2 psi)
    void *dpsi = nullptr, *dh = nullptr, *dgx = nullptr;
    const size_t bytes = sizeof(cd) * (size_t)npw * 2;
    if (hipMalloc(&dpsi, bytes) != hipSuccess || hipMalloc(&dh, bytes) != hipSuccess ||
        hipMalloc(&dgx, bytes) != hipSuccess)
        return 6;
    std::vector<cd> two(2 * (size_t)npw), back(2 * (size_t)npw), gback(2 * (size_t)npw);
    for (int g = 0; g < npw; ++g) two[g] = psi[g], two[npw + g] = 2.0 * psi[g];
    if (hipMemcpy(dpsi, two.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) return 6;
    op.device(2, dpsi, npw, dh, npw, effm::Precision::FP64, true);
    h.gradient().at_k(0).device(2, dpsi, npw, dgx, nullptr, nullptr, npw, effm::Precision::FP64, true);
    if (hipDeviceSynchronize() != hipSuccess) return 6;
    if (hipMemcpy(back.data(), dh, bytes, hipMemcpyDeviceToHost) != hipSuccess) return 6;
    if (hipMemcpy(gback.data(), dgx, bytes, hipMemcpyDeviceToHost) != hipSuccess) return 6;
    double derr = 0, dgerr = 0;
    for (int bnd = 0; bnd < 2; ++bnd)
        for (int g = 0; g < npw; ++g) {
            const double f = bnd + 1.0;
            cd e = f * ((k.kinetic[g] + V0) * psi[g] + lam * inner * p[g]);
            cd ge = f * (k.K[3 * g] * psi[g] + lam * (dinner[0] * p[g] + inner * dp[0][g]));
            derr = std::max(derr, std::abs(back[(size_t)bnd * npw + g] - e));
            dgerr = std::max(dgerr, std::abs(gback[(size_t)bnd * npw + g] - ge));
        }
    (void)hipFree(dpsi);
    (void)hipFree(dh);
    (void)hipFree(dgx);
    std::printf("facade device H|psi> max err %.3e, dH/dk %.3e\n", derr, dgerr);
    if (derr > 2e-9 * ref || dgerr > 2e-8) return 7;
    // Hessian, closed form: d2H/dk_a dk_b psi = delta_ab psi + lam (<d_ab p|psi> p
    // + <d_a p|psi> d_b p + <d_b p|psi> d_a p + <p|psi> d_ab p), with p ~ f(|K|),
    // f = exp(-|K|^2/4): d_a f = -(K_a/2) f, d_ab f = (K_a K_b/4 - delta_ab/2) f.
    // Outputs start at (0.25 + 0.5i) (accumulate); xz is skipped.
    {
        const int ca[6] = {0, 1, 2, 0, 0, 1}, cb[6] = {0, 1, 2, 1, 2, 2};
        const cd base(0.25, 0.5);
        std::vector<cd> hs[6];
        for (auto& v : hs) v.assign(npw, base);
        h.hessian().at_k(0)(psi, effm::HessHPsi{hs[0], hs[1], hs[2], hs[3], {}, hs[5]});
        double herr = 0, href = 0;
        std::vector<cd> d2p(npw);
        for (int c = 0; c < 6; ++c) {
            const int a = ca[c], b = cb[c];
            cd in2 = 0;
            for (int g = 0; g < npw; ++g) {
                const double* K = &k.K[3 * g];
                d2p[g] = (K[a] * K[b] / 4.0 - (a == b ? 0.5 : 0.0)) * p[g];
                in2 += std::conj(d2p[g]) * psi[g];
            }
            for (int g = 0; g < npw; ++g) {
                cd e = base;
                if (c != 4)
                    e += (a == b ? psi[g] : cd(0.0)) +
                         lam * (in2 * p[g] + dinner[a] * dp[b][g] + dinner[b] * dp[a][g] + inner * d2p[g]);
                herr = std::max(herr, std::abs(hs[c][g] - e));
                href = std::max(href, std::abs(e - base));
            }
        }
        std::printf("facade Hessian max err %.3e (scale %.3e)\n", herr, href);
        if (herr > 1e-5 * href) return 9;
    }
#ifdef EFFM_USE_MODULE
    {
        // the reference's spelling, exported at global scope by the module
        Hamiltonian h2(cell);
        h2.loadKPoint(0, k);
        std::vector<cd> hp(npw, 0.0);
        h2.at_k(0)(psi, hp);
        double e2 = 0;
        for (int g = 0; g < npw; ++g) e2 = std::max(e2, std::abs(hp[g] - (hpsi[g] - cd(1.0, -1.0))));
        std::printf("module Hamiltonian vs facade %.3e\n", e2);
        if (e2 > 1e-12 * ref) return 8;
    }
#endif
    std::printf("facade checks passed\n");
    return 0;
}
