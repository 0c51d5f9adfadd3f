// facade_bench.cpp — times the C++ facade (effm_hamiltonian.hpp) the way the
// reference's callers drive H_psi::Hamiltonian: one band per call on host
// spans (test/H_psi/test_hpsi_eigen.cpp:88, src/H_psi/hamiltonian.cppm:
// 173-205), against the batched host form and the device-pointer forms
// (FP64, mixed) on the same cell.  Input: a binary dump of one cell,
// k-point and band block (written by bench.py); output: one JSON line.
//
//   facade_bench <dump> [repeats]
//
// Dump layout (little-endian): "EFFMDUMP", int32 version = 1;
// int32 n1 n2 n3; f64 lattice[9]; f64 vloc[n1 n2 n3];
// int32 natom; int32 species[natom]; f64 tau[3 natom];
// int32 nspecies; per species: int32 lmax, nbeta; int32 beta_l[nbeta];
//   f64 lambda[nbeta]; int32 nq; f64 dq, qmax; f64 beta_tab[nbeta nq];
//   int32 has_dbeta; f64 dbeta_tab[nbeta nq] if has_dbeta;
// int32 npw; f64 K[3 npw]; f64 kinetic[npw]; int32 gidx[3 npw]; f64 kfrac[3];
// int32 nb; complex128 psi[nb npw].
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <complex>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "effm_hamiltonian.hpp"

using cd = std::complex<double>;
using cf = std::complex<float>;

namespace {

struct Reader {
    std::ifstream f;
    explicit Reader(const char* path) : f(path, std::ios::binary) {
        if (!f) throw std::runtime_error(std::string("cannot open ") + path);
    }
    template <typename T>
    T get() {
        T v;
        if (!f.read(reinterpret_cast<char*>(&v), sizeof(T))) throw std::runtime_error("dump truncated");
        return v;
    }
    template <typename T>
    std::vector<T> vec(size_t n) {
        std::vector<T> v(n);
        if (n && !f.read(reinterpret_cast<char*>(v.data()), sizeof(T) * n)) throw std::runtime_error("dump truncated");
        return v;
    }
};

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <dump> [repeats]\n", argv[0]);
        return 2;
    }
    const int reps = argc > 2 ? std::max(1, std::atoi(argv[2])) : 3;
    try {
        Reader r(argv[1]);
        char magic[8];
        r.f.read(magic, 8);
        if (std::memcmp(magic, "EFFMDUMP", 8) != 0 || r.get<int32_t>() != 1) throw std::runtime_error("not a dump");
        effm::CellData cell;
        cell.n1 = r.get<int32_t>();
        cell.n2 = r.get<int32_t>();
        cell.n3 = r.get<int32_t>();
        for (double& v : cell.lattice) v = r.get<double>();
        cell.vloc = r.vec<double>((size_t)cell.n1 * cell.n2 * cell.n3);
        const int natom = r.get<int32_t>();
        cell.species = r.vec<int32_t>(natom);
        cell.tau = r.vec<double>(3 * (size_t)natom);
        const int nsp = r.get<int32_t>();
        for (int s = 0; s < nsp; ++s) {
            effm::SpeciesData sp;
            sp.lmax = r.get<int32_t>();
            const int nbeta = r.get<int32_t>();
            sp.beta_l = r.vec<int32_t>(nbeta);
            sp.lambda = r.vec<double>(nbeta);
            sp.nq = r.get<int32_t>();
            sp.dq = r.get<double>();
            sp.qmax = r.get<double>();
            sp.beta_tab = r.vec<double>((size_t)nbeta * sp.nq);
            if (r.get<int32_t>()) sp.dbeta_tab = r.vec<double>((size_t)nbeta * sp.nq);
            cell.sp.push_back(std::move(sp));
        }
        effm::KPointData k;
        const int npw = r.get<int32_t>();
        k.K = r.vec<double>(3 * (size_t)npw);
        k.kinetic = r.vec<double>(npw);
        k.gidx = r.vec<int32_t>(3 * (size_t)npw);
        for (double& v : k.kfrac) v = r.get<double>();
        const int nb = r.get<int32_t>();
        const std::vector<cd> psi = r.vec<cd>((size_t)nb * npw);

        effm::Hamiltonian h(cell);
        h.loadKPoint(0, k);
        auto op = h.at_k(0);
        std::vector<cd> hpsi((size_t)nb * npw);
        std::vector<cf> psi32((size_t)nb * npw), h32((size_t)nb * npw);
        for (size_t i = 0; i < psi.size(); ++i) psi32[i] = cf((float)psi[i].real(), (float)psi[i].imag());
        auto one_band = [&] {
            for (int b = 0; b < nb; ++b)
                op(std::span<const cd>(psi.data() + (size_t)b * npw, npw),
                   std::span<cd>(hpsi.data() + (size_t)b * npw, npw));
        };
        auto batched = [&] { op(nb, psi, hpsi); };
        auto batched32 = [&] { op(nb, std::span<const cf>(psi32), std::span<cf>(h32)); };
        // device pointers, bands resident in HBM
        void *dpsi = nullptr, *dh = nullptr, *dpsi32 = nullptr, *dh32 = nullptr;
        const size_t b64 = sizeof(cd) * psi.size(), b32 = sizeof(cf) * psi.size();
        if (hipMalloc(&dpsi, b64) != hipSuccess || hipMalloc(&dh, b64) != hipSuccess ||
            hipMalloc(&dpsi32, b32) != hipSuccess || hipMalloc(&dh32, b32) != hipSuccess)
            throw std::runtime_error("hipMalloc failed");
        if (hipMemcpy(dpsi, psi.data(), b64, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(dpsi32, psi32.data(), b32, hipMemcpyHostToDevice) != hipSuccess)
            throw std::runtime_error("hipMemcpy failed");
        auto dev64 = [&] {
            op.device(nb, dpsi, npw, dh, npw, effm::Precision::FP64, true);
            (void)hipDeviceSynchronize();
        };
        auto dev32 = [&] {
            op.device(nb, dpsi32, npw, dh32, npw, effm::Precision::Mixed, true);
            (void)hipDeviceSynchronize();
        };
        auto dev_one_band = [&] {
            for (int b = 0; b < nb; ++b)
                op.device(1, static_cast<const cd*>(dpsi) + (size_t)b * npw, npw, static_cast<cd*>(dh) + (size_t)b * npw,
                          npw, effm::Precision::FP64, true);
            (void)hipDeviceSynchronize();
        };
        auto timeit = [&](auto&& fn) {
            fn();                                     // warm-up (buffers, projectors)
            double best = 1e30;
            for (int i = 0; i < reps; ++i) {
                const double t0 = now_s();
                fn();
                best = std::min(best, now_s() - t0);
            }
            return best * 1e3 / nb;                   // ms per band
        };
        const double t_one = timeit(one_band), t_batch = timeit(batched), t_batch32 = timeit(batched32);
        const double t_dev1 = timeit(dev_one_band), t_dev = timeit(dev64), t_dev32 = timeit(dev32);
        (void)hipFree(dpsi);
        (void)hipFree(dh);
        (void)hipFree(dpsi32);
        (void)hipFree(dh32);
        std::printf("{\"npw\": %d, \"bands\": %d, \"repeats\": %d, \"ms_per_band\": {\"host_one_band_per_call_fp64\": %.6f, "
                    "\"host_batched_fp64\": %.6f, \"host_batched_mixed\": %.6f, \"device_one_band_per_call_fp64\": %.6f, "
                    "\"device_batched_fp64\": %.6f, \"device_batched_mixed\": %.6f}}\n",
                    npw, nb, reps, t_one, t_batch, t_batch32, t_dev1, t_dev, t_dev32);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "facade_bench: %s\n", e.what());
        return 1;
    }
    return 0;
}
