// k_nonlocal.hip — nonlocal pseudopotential term and dH/dk on gfx950.
//
// Reference: src/H_psi/hamiltonian_callable.cpp:200-304 (V_NL psi) and
// src/H_psi/hamiltonian_gradient.cpp:116-294 (dH/dk psi).  The reference
// loops atom x (l, ib) x m with two passes over the G-sphere per channel and
// band; across bands those passes are the complex GEMMs
//   C[c][b]  = sum_g conj(P[c][g]) psi[b][g]        (<p|psi>)
//   hpsi[b][g] += sum_c P[c][g] lambda_c C[c][b]      (sum lambda <p|psi> p)
// which run here on the matrix cores (v_mfma_f64_16x16x4f64 for FP64,
// v_mfma_f32_16x16x4f32 for the mixed path).  The split-K partial sums of
// the first GEMM are reduced and scaled by lambda in FP64 (the D
// contraction), in both precisions.
//
// P is generated on the device once per k-point:
//   P[c][g] = beta_{l,ib}(q_g)/sqrt(Omega) * Y_lm(theta_g, phi_g) * S_a(g)
// with the reference's cubic-Lagrange table rule
// (src/pseudo/ncpp_advance.cppm:644-674), real Y_lm recurrences
// (src/math/sph_harmonics.cppm:120-335, 702-875) and separable structure
// factor (src/H_psi/structure_factor.cppm:92-159).
#include <algorithm>
#include <cmath>
#include <type_traits>

#include "effm_internal.hpp"

namespace effm {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename T> struct Acc;
template <> struct Acc<double> { using type = v4d; };
template <> struct Acc<float>  { using type = v4f; };

__device__ __forceinline__ v4d mfma16(double a, double b, v4d c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4f mfma16(float a, float b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Row of the 16x16 output block held in accumulator register v of a lane in
// lane group kk = lane / 16 (the column is lane % 16).  A and B operands
// share one layout (row or column = lane % 16, k = lane / 16), but the
// outputs differ: f32 16x16x4 keeps 4 consecutive rows per lane group,
// f64 16x16x4 interleaves them with stride 4.
template <typename T>
__device__ __forceinline__ int out_row(int kk, int v) {
    return sizeof(T) == 8 ? v * 4 + kk : kk * 4 + v;
}

// ---------------------------------------------------------------------------
//  Projector generation
// ---------------------------------------------------------------------------
__device__ __forceinline__ double lagrange_cubic(double f0, double f1, double f2, double f3,
                                                 double px) {
    double ux = 1.0 - px;
    double vx = 2.0 - px;
    double wx = 3.0 - px;
    return f0 * ux * vx * wx / 6.0 + f1 * px * vx * wx / 2.0 - f2 * px * ux * wx / 2.0 +
           f3 * px * ux * vx / 6.0;
}

__device__ __forceinline__ double interp_tab(const double* t, double dq, double q) {
    if (q == 0.0) return t[0];
    double s = q / dq;
    int i0 = (int)floor(s) - 1;
    double px = s - (double)i0;
    if (i0 < 0) { i0 = 0; px = s; }
    return lagrange_cubic(t[i0], t[i0 + 1], t[i0 + 2], t[i0 + 3], px);
}

__device__ __forceinline__ double seed_coeff(int m) {
    double c = 0.5 / sqrt(M_PI);
    if (m & 1) c = -c;
    for (int k = 1; k <= m; ++k) c *= sqrt((double)(2 * k + 1) / (2 * k));
    return c;
}

// Y_lm, dY/dtheta and (1/sin theta) dY/dphi for one (l, m) by the column
// recurrences of the reference (QRecurrence, AngGradThetaRecurrence,
// AngGradPhiRecurrence).
__device__ void ylm_eval(int l, int m, double ct, double st, double cp, double sp, bool grad,
                         double* Y, double* dT, double* dP) {
    const int ma = m < 0 ? -m : m;
    const double c0 = seed_coeff(ma);
    double Qc = (ma == 0) ? 0.5 / sqrt(M_PI) : c0 * pow(st, (double)ma);
    double Qp = 0.0;
    double dQc = 0.0, dQp = 0.0, Rc = 0.0, Rp = 0.0;
    if (grad) {
        if (ma == 0) dQc = 0.0;
        else if (ma == 1) dQc = seed_coeff(1) * ct;
        else dQc = (double)ma * c0 * pow(st, (double)(ma - 1)) * ct;
        if (ma == 1) Rc = c0;
        else if (ma >= 2) Rc = c0 * pow(st, (double)(ma - 1));
    }
    if (l > ma) {
        double f = sqrt(2.0 * ma + 3.0);
        double Qmm = Qc;
        Qp = Qc; Qc = f * ct * Qc;
        if (grad) {
            dQp = dQc; dQc = f * (-st * Qmm + ct * dQc);
            Rp = Rc; Rc = f * ct * Rc;
        }
        for (int ll = ma + 2; ll <= l; ++ll) {
            double r1 = (2.0 * ll - 1.0) *
                        sqrt(((2.0 * ll + 1.0) * (ll - ma)) / ((2.0 * ll - 1.0) * (ll + ma)));
            double r2 = (ll + ma - 1.0) *
                        sqrt(((2.0 * ll + 1.0) * (ll - ma) * (ll - ma - 1)) /
                             ((2.0 * ll - 3.0) * (ll + ma) * (ll + ma - 1)));
            double den = 1.0 / (ll - ma);
            double Ql1 = Qc;
            double nq = (ct * r1 * Qc - r2 * Qp) * den;
            Qp = Qc; Qc = nq;
            if (grad) {
                double nd = (ct * r1 * dQc - st * r1 * Ql1 - r2 * dQp) * den;
                dQp = dQc; dQc = nd;
                double nr = (ct * r1 * Rc - r2 * Rp) * den;
                Rp = Rc; Rc = nr;
            }
        }
    }
    double cm = 1.0, sm = 0.0;
    for (int k = 0; k < ma; ++k) {
        double oc = cm;
        cm = cm * cp - sm * sp;
        sm = sm * cp + oc * sp;
    }
    const double r2 = sqrt(2.0);
    if (m == 0) *Y = Qc;
    else *Y = (m > 0) ? (r2 * Qc) * cm : (r2 * Qc) * sm;
    if (grad) {
        if (m == 0) *dT = dQc;
        else *dT = (m > 0) ? r2 * dQc * cm : r2 * dQc * sm;
        if (m == 0) *dP = 0.0;
        else if (m > 0) *dP = (-(double)m * r2) * Rc * sm;
        else *dP = ((double)ma * r2) * Rc * cm;
    }
}

__device__ __forceinline__ double2 cmul_d(double2 a, double2 b) {
    return double2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

template <bool GRAD>
__global__ void __launch_bounds__(256)
k_build_proj(const double* __restrict__ K, const int* __restrict__ gidx, int npw, double kf0,
             double kf1, double kf2, const ChanDev* __restrict__ chans,
             const TableDev* __restrict__ tinfo, const double* __restrict__ btab,
             const double* __restrict__ dbtab, int nq_stride, const double* __restrict__ tau,
             double2* __restrict__ P64, float2* __restrict__ P32, double2* __restrict__ dP64,
             float2* __restrict__ dP32, int nch) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (g >= npw) return;
    const ChanDev ch = chans[c];
    const double kx = K[g], ky = K[npw + g], kz = K[2 * npw + g];
    // GKK::computeSpherical (src/io/GKK.cppm:431-449)
    const double q = sqrt(kx * kx + ky * ky + kz * kz);
    const double th = (q > 0.0) ? acos(kz / q) : 0.0;
    const double ph = atan2(ky, kx);
    const double ct = cos(th), st = sin(th), cp = cos(ph), sp = sin(ph);
    double Y, dT = 0.0, dPh = 0.0;
    ylm_eval(ch.l, ch.m, ct, st, cp, sp, GRAD, &Y, &dT, &dPh);
    const TableDev t = tinfo[ch.table];
    const double beta = interp_tab(btab + (size_t)ch.table * nq_stride, t.dq, q) * t.norm;
    // S = exp(-2 pi i k.tau) p_x[g_x] p_y[g_y] p_z[g_z]
    const double* ta = tau + 3 * ch.atom;
    double s, co;
    sincospi(-2.0 * (kf0 * ta[0] + kf1 * ta[1] + kf2 * ta[2]), &s, &co);
    double2 S = double2{co, s};
    for (int d = 0; d < 3; ++d) {
        sincospi(-2.0 * (double)gidx[3 * g + d] * ta[d], &s, &co);
        S = cmul_d(S, double2{co, s});
    }
    const double by = beta * Y;
    const size_t o = (size_t)c * npw + g;
    const double2 p = double2{by * S.x, by * S.y};
    if (P64) P64[o] = p;
    if (P32) P32[o] = float2{(float)p.x, (float)p.y};
    if (GRAD) {
        const double dbeta = interp_tab(dbtab + (size_t)ch.table * nq_stride, t.dq, q) * t.norm;
        const double bo = (q > 1e-12) ? beta / q : t.dbeta0;
        const double gq = dbeta * Y;
        const double gt = bo * dT;
        const double gp = bo * dPh;
        const double r[3] = {st * cp * gq + ct * cp * gt - sp * gp,
                             st * sp * gq + ct * sp * gt + cp * gp,
                             ct * gq - st * gt};
        for (int a = 0; a < 3; ++a) {
            const size_t oa = ((size_t)a * nch + c) * npw + g;
            const double2 v = double2{r[a] * S.x, r[a] * S.y};
            if (dP64) dP64[oa] = v;
            if (dP32) dP32[oa] = float2{(float)v.x, (float)v.y};
        }
    }
}

template <typename T>
hipError_t launch_build_proj(Ctx& c, KptDev& k, bool grad) {
    const int nch = c.nch();
    if (nch == 0) return hipSuccess;
    const bool f64 = sizeof(T) == 8;
    dim3 grid((k.npw + 255) / 256, nch);
    timer_begin(c);
    if (grad)
        hipLaunchKernelGGL((k_build_proj<true>), grid, dim3(256), 0, c.stream,
                           static_cast<const double*>(k.K.ptr), static_cast<const int*>(k.gidx.ptr),
                           k.npw, k.kfrac[0], k.kfrac[1], k.kfrac[2],
                           static_cast<const ChanDev*>(c.chan_dev.ptr),
                           static_cast<const TableDev*>(c.tab_info.ptr),
                           static_cast<const double*>(c.btab.ptr), static_cast<const double*>(c.dbtab.ptr),
                           c.nq_max, static_cast<const double*>(c.tau.ptr),
                           f64 ? static_cast<double2*>(k.P64.ptr) : nullptr,
                           f64 ? nullptr : static_cast<float2*>(k.P32.ptr),
                           f64 ? static_cast<double2*>(k.dP64.ptr) : nullptr,
                           f64 ? nullptr : static_cast<float2*>(k.dP32.ptr), nch);
    else
        hipLaunchKernelGGL((k_build_proj<false>), grid, dim3(256), 0, c.stream,
                           static_cast<const double*>(k.K.ptr), static_cast<const int*>(k.gidx.ptr),
                           k.npw, k.kfrac[0], k.kfrac[1], k.kfrac[2],
                           static_cast<const ChanDev*>(c.chan_dev.ptr),
                           static_cast<const TableDev*>(c.tab_info.ptr),
                           static_cast<const double*>(c.btab.ptr), nullptr, c.nq_max,
                           static_cast<const double*>(c.tau.ptr),
                           f64 ? static_cast<double2*>(k.P64.ptr) : nullptr,
                           f64 ? nullptr : static_cast<float2*>(k.P32.ptr), nullptr, nullptr, nch);
    timer_end(c, "proj_build");
    return hipGetLastError();
}

template hipError_t launch_build_proj<double>(Ctx&, KptDev&, bool);
template hipError_t launch_build_proj<float>(Ctx&, KptDev&, bool);

// ---------------------------------------------------------------------------
//  GEMM 1: part[split][row0 + m][n] = sum_{k in split} conj(A[m][k]) B[n][k]
//  A: [M][lda] (projectors, G contiguous), B: [N][ldb] (bands, G contiguous).
//  Workgroup = 4 waves, 64 x 64 tile; wave = 2 x 2 blocks of 16 x 16.
//  Each lane owns 4 consecutive k of its row per 16-wide k step, so both
//  operands are read as contiguous 4-element runs.  1D grid of
//  tiles x splits; the tiles of one split are consecutive (one XCD), so its
//  slabs of A and B are fetched into that XCD's L2 once.
// ---------------------------------------------------------------------------
template <typename T>
struct ProjSeg {
    const cx_t<T>* A;      // rows [0, M1): [M1][lda]
    const cx_t<T>* B;      // [N][ldb]
    cx_t<T>* part;         // [nsplit][Mtot][N]
    int64_t lda, ldb;
    int K, kchunk;
    // rows [M1, M) from A2 ([M - M1][lda], the gradient's dP after P); by
    // default every row comes from A
    const cx_t<T>* A2 = nullptr;
    int M1 = 1 << 30;
};

template <typename T>
struct ProjBatch {
    int nseg;
    ProjSeg<T> s[kMaxSeg];
};

// Workgroup tile 64 x TN: TN = 64 (waves 2 x 2, each 2 x 2 blocks of
// 16 x 16); for N <= 32 (the shrinking row blocks of a Chebyshev filter,
// small eigensolver blocks) TN = 32 (waves 2 x 2, each 2 x 1 blocks); for
// N <= 16 (one band per call) TN = 16 (waves 4 x 1, one block each): a
// 64-wide tile there spends 64x the MFMA work of the product and made the
// call MFMA-bound.
template <int TN> struct GemmWaves {
    static constexpr int WGC = TN == 16 ? 1 : 2;    // waves along N
    static constexpr int WGR = 4 / WGC;             // waves along M
    static constexpr int WR = 64 / (16 * WGR);      // 16-row blocks per wave
    static constexpr int WC = TN / (16 * WGC);      // 16-column blocks per wave
    static_assert(TN == 64 || TN == 32 || TN == 16, "tile width");
};

// Tile width for an N-column product.
inline int gemm_tile_width(int n) { return n <= 16 ? 16 : n <= 32 ? 32 : 64; }

// Workgroups per CU the projection's register allocation must allow: FP32
// keeps four (its LDS allows four) with the second register set of slabs in
// flight; FP64 is held to three by its accumulators.
template <typename T> struct ProjTune { static constexpr int MINB = sizeof(T) == 4 ? 4 : 2; };

template <typename T, int TN>
__global__ void __launch_bounds__(256, ProjTune<T>::MINB)
k_gemm_proj(const ProjBatch<T> bt, int M, int N, int nsplit, int Mtot, int row0) {
    using C = cx_t<T>;
    using V = typename Acc<T>::type;
    using W = GemmWaves<TN>;
    constexpr int WR = W::WR, WC = W::WC;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = lane & 15, kk = lane >> 4;
    const int nx = (M + 63) / 64, nt = nx * ((N + TN - 1) / TN);
    const int L0 = xcd_order(blockIdx.x, gridDim.x);
    // segment-major, then split, then tile: one segment's splits stay together
    const int seg = L0 / (nt * nsplit);
    const int L = L0 - seg * nt * nsplit;
    const int split = L / nt, t = L - split * nt;
    const ProjSeg<T>& sg = bt.s[seg];
    const C* __restrict__ A = sg.A;
    const C* __restrict__ B = sg.B;
    C* __restrict__ part = sg.part;
    const int64_t lda = sg.lda, ldb = sg.ldb;
    const int K = sg.K, kchunk = sg.kchunk;
    const int tm0 = (t % nx) * 64, tn0 = (t / nx) * TN;      // workgroup tile
    const int m0 = tm0 + (wave % W::WGR) * (16 * WR);         // wave tile
    const int n0 = tn0 + (wave / W::WGR) * (16 * WC);
    const int ks = split * kchunk;
    const int ke = min(K, ks + kchunk);
    // Gauss's three-product complex MAC: for conj(P) psi with a = Re P,
    // b = -Im P, c = Re psi, d = Im psi:
    //   k1 += c (a + b),  k2 += a (d - c),  k3 += b (c + d)
    //   Re = k1 - k3,     Im = k1 + k2
    V k1[WR][WC], k2[WR][WC], k3[WR][WC];
    for (int r = 0; r < WR; ++r)
        for (int q = 0; q < WC; ++q) {
            k1[r][q] = V{0, 0, 0, 0};
            k2[r][q] = V{0, 0, 0, 0};
            k3[r][q] = V{0, 0, 0, 0};
        }
    // Slabs of BK k (128 bytes of a row) of the tile's 64 rows of A and 64 of
    // B are staged through LDS, double-buffered: each global load instruction
    // covers whole cache lines (consecutive lanes, consecutive k), and every
    // element is fetched once per workgroup instead of once per wave.
    constexpr int BK = 128 / (int)sizeof(C);          // FP32 16, FP64 8
    // row stride of 144 bytes (36 banks): the 16 lanes of one 16-byte LDS
    // read cycle hit 16 distinct 4-bank groups
    constexpr int LDK = BK + 16 / (int)sizeof(C);
    constexpr int KPL = BK / 4;                       // consecutive k per lane and slab
    // each global load is 16 bytes: one complex128, or two consecutive
    // complex64 of one row (kchunk is a multiple of 16, so a pair never
    // straddles a split; only k = K - 1 at the end of the last split can
    // stand alone)
    constexpr int VEC = 16 / (int)sizeof(C);
    constexpr int LPR = BK / VEC;                     // loading lanes per slab row
    constexpr int RPI = 256 / LPR;                    // rows per load instruction
    constexpr int NLR = 64 / RPI;                     // loads per lane, matrix and slab
    // B has TN rows: with TN < RPI only the first TN / (RPI / 4) waves load
    // it (a wave-uniform predicate)
    constexpr bool BALL = TN >= RPI;
    using LV = typename std::conditional<sizeof(T) == 8, double2, float4>::type;
    __shared__ __align__(16) C sA[2][64 * LDK];
    __shared__ __align__(16) C sB[2][TN * LDK];
    const bool bload = BALL || (int)threadIdx.x / LPR < TN;
    // Rows / columns past M / N are clamped onto the last one: they only feed
    // outputs that are never stored, so the loads need no branch.
    const C* ga[NLR];
    const C* gb[NLR];
    int lofs[NLR];
    const int kofs = (threadIdx.x % LPR) * VEC;
#pragma unroll
    for (int q = 0; q < NLR; ++q) {
        const int row = (int)threadIdx.x / LPR + q * RPI;
        const int gr = min(tm0 + row, M - 1);
        ga[q] = (sg.A2 != nullptr && gr >= sg.M1) ? sg.A2 + (int64_t)(gr - sg.M1) * lda : A + (int64_t)gr * lda;
        gb[q] = B + (int64_t)min(tn0 + (BALL ? row : min(row, TN - 1)), N - 1) * ldb;
        lofs[q] = row * LDK + kofs;
    }
    // FP32 loads each slab two slabs ahead (a second register set at the
    // same four workgroups per CU; measured on the 64-atom projection: 271
    // -> 259 us), FP64 one ahead
    constexpr int DEPTH = sizeof(T) == 4 ? 2 : 1;
    LV pa[DEPTH][NLR], pb[DEPTH][NLR];
    // k past the split end loads as zero
    auto gload = [&](int kb, int d) {
        const int k = kb + kofs;
        if constexpr (VEC == 1) {
            // clamped address, zeroed operands: no branch (and no scratch
            // array, which an unselected B operand made the compiler use)
            const bool kv = k < ke;
            const int kc = kv ? k : ke - 1;
#pragma unroll
            for (int q = 0; q < NLR; ++q) {
                const C x = ga[q][kc];
                pa[d][q] = kv ? x : C{0, 0};
                if (q * RPI < TN && bload) {
                    const C y = gb[q][kc];
                    pb[d][q] = kv ? y : C{0, 0};
                }
            }
        } else {
#pragma unroll
            for (int q = 0; q < NLR; ++q) {
                LV x{}, y{};
                const bool qb = q * RPI < TN && bload;
                if (k + VEC <= ke) {
                    x = *reinterpret_cast<const LV*>(ga[q] + k);
                    if (qb) y = *reinterpret_cast<const LV*>(gb[q] + k);
                } else if (k < ke) {
                    const C u = ga[q][k];
                    x.x = u.x; x.y = u.y;
                    if (qb) {
                        const C v = gb[q][k];
                        y.x = v.x; y.y = v.y;
                    }
                }
                pa[d][q] = x;
                pb[d][q] = y;
            }
        }
    };
    auto lstore = [&](int buf, int d) {
#pragma unroll
        for (int q = 0; q < NLR; ++q) {
            *reinterpret_cast<LV*>(&sA[buf][lofs[q]]) = pa[d][q];
            if (q * RPI < TN && bload) *reinterpret_cast<LV*>(&sB[buf][lofs[q]]) = pb[d][q];
        }
    };
    // lane (i, kk) takes k = kk*KPL + s of rows / columns i of its blocks
    auto mma = [&](int buf) {
        constexpr int WM = WR > WC ? WR : WC;
        C a[WR][KPL], b[WC][KPL];
#pragma unroll
        for (int r = 0; r < WM; ++r)
#pragma unroll
            for (int s = 0; s < KPL; ++s) {
                if (r < WR) a[r][s] = sA[buf][(m0 - tm0 + r * 16 + i) * LDK + kk * KPL + s];
                if (r < WC) b[r][s] = sB[buf][(n0 - tn0 + r * 16 + i) * LDK + kk * KPL + s];
            }
#pragma unroll
        for (int s = 0; s < KPL; ++s) {
            T sa[WR], sc[WC], dc[WC];
#pragma unroll
            for (int r = 0; r < WM; ++r) {
                if (r < WR) sa[r] = a[r][s].x - a[r][s].y;             // a + b
                if (r < WC) {
                    sc[r] = b[r][s].x + b[r][s].y;             // c + d
                    dc[r] = b[r][s].y - b[r][s].x;             // d - c
                }
            }
#pragma unroll
            for (int r = 0; r < WR; ++r)
#pragma unroll
                for (int q = 0; q < WC; ++q) k1[r][q] = mfma16(sa[r], b[q][s].x, k1[r][q]);
#pragma unroll
            for (int r = 0; r < WR; ++r)
#pragma unroll
                for (int q = 0; q < WC; ++q) k2[r][q] = mfma16(a[r][s].x, dc[q], k2[r][q]);
#pragma unroll
            for (int r = 0; r < WR; ++r)
#pragma unroll
                for (int q = 0; q < WC; ++q) k3[r][q] = mfma16(-a[r][s].y, sc[q], k3[r][q]);
        }
    };
    if (ks < ke) {
        if constexpr (DEPTH == 1) {
            gload(ks, 0);
            lstore(0, 0);
            __syncthreads();
            int buf = 0;
            for (int kb = ks; kb < ke; kb += BK) {
                const bool more = kb + BK < ke;
                if (more) gload(kb + BK, 0);   // next slab in flight during the MFMAs
                mma(buf);
                if (more) lstore(buf ^ 1, 0);
                __syncthreads();
                buf ^= 1;
            }
        } else {
            // the slab at kb sits in LDS buffer and register set
            // ((kb - ks) / BK) % 2; two slabs in flight during the MFMAs
            gload(ks, 0);
            if (ks + BK < ke) gload(ks + BK, 1);
            lstore(0, 0);
            __syncthreads();
            for (int kb = ks; kb < ke; kb += 2 * BK) {
                if (kb + 2 * BK < ke) gload(kb + 2 * BK, 0);
                mma(0);
                if (kb + BK < ke) lstore(1, 1);
                __syncthreads();
                if (kb + BK < ke) {
                    if (kb + 3 * BK < ke) gload(kb + 3 * BK, 1);
                    mma(1);
                    if (kb + 2 * BK < ke) lstore(0, 0);
                    __syncthreads();
                }
            }
        }
    }
    C* po = part + (int64_t)split * Mtot * N;
    for (int r = 0; r < WR; ++r)
        for (int q = 0; q < WC; ++q)
            for (int v = 0; v < 4; ++v) {
                int row = m0 + r * 16 + out_row<T>(kk, v);
                int col = n0 + q * 16 + i;
                if (row < M && col < N)
                    po[(int64_t)(row0 + row) * N + col] =
                        C{k1[r][q][v] - k3[r][q][v], k1[r][q][v] + k2[r][q][v]};
            }
}

// out[m][n] = lambda[m % nch] * sum_s part[s][m][n], summed in FP64.
template <typename T>
struct ReduceBatch {
    int nseg;
    const cx_t<T>* part[kMaxSeg];
    cx_t<T>* out[kMaxSeg];
};

// grid: nseg x per_seg blocks of 256
template <typename T>
__global__ void k_reduce_scale(const ReduceBatch<T> bt, int per_seg, int nsplit, int Mtot, int N,
                               const double* __restrict__ lam, int nch, int64_t ldo) {
    const int seg = blockIdx.x / per_seg;
    const int64_t idx = (int64_t)(blockIdx.x - seg * per_seg) * blockDim.x + threadIdx.x;
    const int64_t tot = (int64_t)Mtot * N;
    if (idx >= tot) return;
    const cx_t<T>* __restrict__ part = bt.part[seg] + idx;
    cx_t<T>* __restrict__ out = bt.out[seg];
    double re = 0.0, im = 0.0;
    // eight splits' loads in flight at a time, added in split order
    int s = 0;
    for (; s + 8 <= nsplit; s += 8) {
        cx_t<T> v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = part[(int64_t)(s + u) * tot];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            re += (double)v[u].x;
            im += (double)v[u].y;
        }
    }
    for (; s < nsplit; ++s) {
        const cx_t<T> v = part[(int64_t)s * tot];
        re += (double)v.x;
        im += (double)v.y;
    }
    const int m = (int)(idx / N);
    const double l = lam ? lam[m % nch] : 1.0;          // no lambda: a plain overlap
    out[(int64_t)m * ldo + (idx - (int64_t)m * N)] = cx_t<T>{(T)(re * l), (T)(im * l)};
}

// The same sum for few outputs and many splits (one band per call: 512
// outputs over ~150 splits left k_reduce_scale two workgroups of long
// serial loops): 32 lanes per output, lane l summing splits l, l + 32, ...
// in FP64, then a fixed-order shuffle tree, so the result is deterministic.
// grid: nseg x per_seg blocks of 256 (8 outputs each)
template <typename T>
__global__ void __launch_bounds__(256) k_reduce_scale_wide(const ReduceBatch<T> bt, int per_seg, int nsplit,
                                                           int Mtot, int N, const double* __restrict__ lam,
                                                           int nch, int64_t ldo) {
    constexpr int LPO = 32;
    const int seg = blockIdx.x / per_seg;
    const int l = threadIdx.x % LPO;
    const int64_t idx = (int64_t)(blockIdx.x - seg * per_seg) * (256 / LPO) + threadIdx.x / LPO;
    const int64_t tot = (int64_t)Mtot * N;
    double re = 0.0, im = 0.0;
    if (idx < tot) {
        const cx_t<T>* __restrict__ part = bt.part[seg];
        for (int s = l; s < nsplit; s += LPO) {
            const cx_t<T> v = part[(int64_t)s * tot + idx];
            re += (double)v.x;
            im += (double)v.y;
        }
    }
#pragma unroll
    for (int off = LPO / 2; off > 0; off >>= 1) {       // every lane takes part: no divergence
        re += __shfl_down(re, off, LPO);
        im += __shfl_down(im, off, LPO);
    }
    if (idx < tot && l == 0) {
        const int m = (int)(idx / N);
        const double lm = lam ? lam[m % nch] : 1.0;
        bt.out[seg][(int64_t)m * ldo + (idx - (int64_t)m * N)] = cx_t<T>{(T)(re * lm), (T)(im * lm)};
    }
}

// out = lambda * sum over the splits, with the kernel that fills the device
// for the shape (the choice depends on the shape only, so equal shapes give
// equal bits)
template <typename T>
static hipError_t launch_reduce(Ctx& c, const ReduceBatch<T>& bt, int nseg, int nsplit, int Mtot, int N,
                                const double* lam, int nch, int64_t ldo) {
    const int64_t tot = (int64_t)Mtot * N;
    // many splits and too few outputs for one 256-output workgroup per CU:
    // 32 lanes per output
    if (nsplit >= 16 && (tot < 32768 || tot < (int64_t)c.ncu * 256)) {
        const int per = (int)((tot + 7) / 8);
        hipLaunchKernelGGL((k_reduce_scale_wide<T>), dim3((unsigned)(per * nseg)), dim3(256), 0, c.stream, bt, per,
                           nsplit, Mtot, N, lam, nch, ldo);
    } else {
        const int per = (int)((tot + 255) / 256);
        hipLaunchKernelGGL((k_reduce_scale<T>), dim3((unsigned)(per * nseg)), dim3(256), 0, c.stream, bt, per,
                           nsplit, Mtot, N, lam, nch, ldo);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
//  GEMM 2: out[n][m] (+)= diag[m] * src[n][m] + sum_k A[k][m] B[k][n]
//  A: [K][lda] (projectors, G contiguous), B: [K][ldb] (scaled <p|psi>).
//  1D grid; the column tiles of one row tile are consecutive (one XCD), so
//  each projector slab is fetched into one L2.
// ---------------------------------------------------------------------------
template <typename T>
struct ApplySegK {
    const cx_t<T>* A;      // [K][lda]
    const cx_t<T>* B;      // [K][ldb]
    cx_t<T>* out;          // [N][ldo]
    const cx_t<T>* A2;     // optional second product A2 B2 (same K, lda, ldb), or NULL
    const cx_t<T>* B2;
    const double* diag;    // [M] or NULL
    const cx_t<T>* src;    // [N][lds] (with diag)
    int64_t lda, ldb, ldo, lds;
    int M;
    int tile0;             // first tile of the segment
};

template <typename T>
struct ApplyBatch {
    int nseg;
    // tiles [0, full) are whole 64 x TN tiles, one workgroup each; with TN =
    // 64 the tiles [full, total) run as four 64 x 16 tiles each, after the
    // whole ones (the per-CU remainder of the tile count, see gemm_apply)
    int full;
    ApplySegK<T> s[kMaxSeg];
};

// Slab depth and LDS row pitch of the apply GEMM.  FP64: 12 k-rows per slab
// and unpadded rows, 48 KB, so three workgroups share a CU (measured on the
// 64-atom apply: 8 rows 680 us, 16 rows 640 us at two workgroups per CU, 12
// rows 590 us); FP32: 16 rows (wider slabs halve the workgroups per CU,
// narrower add barriers).  Row pitch: a 16-byte read cycle (8 lanes) covers
// 128 contiguous bytes of one k-row, so FP64 needs no padding; FP32 rows are
// 64 elements + 128 bytes, consecutive k-rows start 32 banks apart and the
// two 16-lane halves of an 8-byte read cycle do not collide.
template <typename T> struct ApplyShape {
    static constexpr int BK = sizeof(T) == 8 ? 12 : 16;
    static constexpr int LDM = sizeof(T) == 8 ? 64 : 64 + 128 / (int)sizeof(cx_t<T>);
};

// One 64 x TN tile of hpsi (+)= A B (+ A2 B2): rows tm0.., columns tn0.. of
// the segment's output; sA / sB are the workgroup's double-buffered LDS
// slabs (room for TN = 64).
template <typename T, bool TWO, int TN>
__device__ __forceinline__ void apply_tile(const ApplySegK<T>& sg, int tm0, int tn0, int N, int K, int accumulate,
                                           cx_t<T>* sA, cx_t<T>* sB) {
    using C = cx_t<T>;
    using V = typename Acc<T>::type;
    using W = GemmWaves<TN>;
    constexpr int WR = W::WR, WC = W::WC;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = lane & 15, kk = lane >> 4;
    const C* __restrict__ A = sg.A;
    const C* __restrict__ B = sg.B;
    const C* __restrict__ A2 = sg.A2;
    const C* __restrict__ B2 = sg.B2;
    C* __restrict__ out = sg.out;
    const double* __restrict__ diag = sg.diag;
    const C* __restrict__ src = sg.src;
    const int64_t lda = sg.lda, ldb = sg.ldb, ldo = sg.ldo, lds = sg.lds;
    const int M = sg.M;
    const int m0 = tm0 + (wave % W::WGR) * (16 * WR);         // wave tile
    const int n0 = tn0 + (wave / W::WGR) * (16 * WC);
    // Gauss's three-product complex MAC with a = Re P, b = Im P, c = Re C,
    // d = Im C: k1 += c (a + b), k2 += a (d - c), k3 += b (c + d).
    V k1[WR][WC], k2[WR][WC], k3[WR][WC];
    for (int r = 0; r < WR; ++r)
        for (int q = 0; q < WC; ++q) {
            k1[r][q] = V{0, 0, 0, 0};
            k2[r][q] = V{0, 0, 0, 0};
            k3[r][q] = V{0, 0, 0, 0};
        }
    // Slabs of BK k-rows of A (64 consecutive m) and of B (TN consecutive n)
    // are staged through LDS, double-buffered.  Every global load is 16
    // bytes (one complex128, or two consecutive complex64), consecutive
    // lanes on consecutive addresses, so one instruction covers whole cache
    // lines; the row pointers advance by BK rows per slab (no per-element
    // 64-bit index arithmetic).
    constexpr int BK = ApplyShape<T>::BK, LDM = ApplyShape<T>::LDM;
    constexpr int VEC = 16 / (int)sizeof(C);          // complex per 16-byte load
    constexpr int CPR = 64 / VEC;                     // loading lanes per slab row
    constexpr int RSTEP = 256 / CPR;                  // slab rows between a lane's loads
    constexpr int NL = BK / RSTEP;                    // loads per lane, matrix and slab
    static_assert(NL * RSTEP == BK, "slab rows must split evenly over the loads");
    // B slabs are BK x TN: TN = 64 as A; TN = 16 unpadded (a 16-lane read
    // cycle covers one whole 128- or 256-byte row); TN = 32 rows of 48 (the
    // two 16-lane halves of an FP32 8-byte read cycle, k-rows kk and kk + 1,
    // start 32 banks apart); loaded by the first BK * TN / VEC lanes
    constexpr int LDB = TN == 64 ? LDM : TN == 32 ? 48 : TN;
    constexpr int CPRB = TN / VEC;
    constexpr int RSTEPB = 256 / CPRB;
    constexpr int NLB = (BK + RSTEPB - 1) / RSTEPB;
    // FP32 loads each slab two slabs ahead (a second register set, 16 more
    // VGPRs; measured on the 64-atom apply: 2-3% faster); FP64 one ahead (a
    // second set would cost a workgroup per CU)
    constexpr int DEPTH = sizeof(T) == 4 ? 2 : 1;
    using LV = typename std::conditional<sizeof(T) == 8, double2, float4>::type;
    const int lrow = threadIdx.x / CPR, lcol = (threadIdx.x % CPR) * VEC;
    const int lrowb = threadIdx.x / CPRB, lcolb = (threadIdx.x % CPRB) * VEC;
    const int ca = tm0 + lcol, cb = tn0 + lcolb;      // first column this lane loads
    // the k loop runs over nsl slabs of (A, B) and, with TWO, nsl more of
    // (A2, B2); rows past K of a slab load as zero, columns past M / N load
    // as zero (they only feed outputs that are never stored)
    const int nsl = (K + BK - 1) / BK;
    const int ntot = TWO ? 2 * nsl : nsl;
    const C* pa = K > 0 ? A + (int64_t)lrow * lda + ca : nullptr;
    const C* pb = K > 0 ? B + (int64_t)lrowb * ldb + cb : nullptr;
    int krow = 0, lt = 0;                             // next slab: its first row, its index
    LV ra[DEPTH][NL], rb[DEPTH][NLB];
    auto load16 = [](const C* p, int col, int ncol) -> LV {
        LV v{};
        if (col + VEC <= ncol) {
            v = *reinterpret_cast<const LV*>(p);
        } else if constexpr (VEC == 2) {
            if (col < ncol) {                         // the last odd column: one complex
                const C x = *p;
                v.x = x.x;
                v.y = x.y;
            }
        }
        return v;
    };
    // the next slab into register set d
    auto gload = [&](int d) {
        if (TWO && lt == nsl) {
            pa = A2 + (int64_t)lrow * lda + ca;
            pb = B2 + (int64_t)lrowb * ldb + cb;
            krow = 0;
        }
#pragma unroll
        for (int q = 0; q < NL; ++q) {
            const bool kv = krow + lrow + q * RSTEP < K;
            ra[d][q] = kv ? load16(pa + (int64_t)(q * RSTEP) * lda, ca, M) : LV{};
        }
#pragma unroll
        for (int q = 0; q < NLB; ++q) {
            const int rr = lrowb + q * RSTEPB;
            const bool kv = (TN == 64 || rr < BK) && krow + rr < K;   // TN = 64: rr < BK always
            rb[d][q] = kv ? load16(pb + (int64_t)(q * RSTEPB) * ldb, cb, N) : LV{};
        }
        pa += (int64_t)BK * lda;
        pb += (int64_t)BK * ldb;
        krow += BK;
        ++lt;
    };
    auto lstore = [&](int buf, int d) {
        C* a = sA + buf * (BK * LDM);
        C* b = sB + buf * (BK * LDB);
#pragma unroll
        for (int q = 0; q < NL; ++q) *reinterpret_cast<LV*>(&a[(lrow + q * RSTEP) * LDM + lcol]) = ra[d][q];
#pragma unroll
        for (int q = 0; q < NLB; ++q) {
            const int rr = lrowb + q * RSTEPB;
            if (TN == 64 || rr < BK) *reinterpret_cast<LV*>(&b[rr * LDB + lcolb]) = rb[d][q];
        }
    };
    // MFMA k-step s of the slab: lane (i, kk) takes k-row 4 s + kk
    auto mma = [&](int buf) {
        const C* a_s = sA + buf * (BK * LDM);
        const C* b_s = sB + buf * (BK * LDB);
#pragma unroll
        for (int s = 0; s < BK / 4; ++s) {
            const int kr = 4 * s + kk;
            constexpr int WM = WR > WC ? WR : WC;
            C a[WR], b[WC];
#pragma unroll
            for (int r = 0; r < WM; ++r) {
                if (r < WR) a[r] = a_s[kr * LDM + (m0 - tm0) + r * 16 + i];
                if (r < WC) b[r] = b_s[kr * LDB + (n0 - tn0) + r * 16 + i];
            }
            T sa[WR], sc[WC], dc[WC];
#pragma unroll
            for (int r = 0; r < WM; ++r) {
                if (r < WR) sa[r] = a[r].x + a[r].y;
                if (r < WC) {
                    sc[r] = b[r].x + b[r].y;
                    dc[r] = b[r].y - b[r].x;
                }
            }
#pragma unroll
            for (int r = 0; r < WR; ++r)
#pragma unroll
                for (int q = 0; q < WC; ++q) k1[r][q] = mfma16(sa[r], b[q].x, k1[r][q]);
#pragma unroll
            for (int r = 0; r < WR; ++r)
#pragma unroll
                for (int q = 0; q < WC; ++q) k2[r][q] = mfma16(a[r].x, dc[q], k2[r][q]);
#pragma unroll
            for (int r = 0; r < WR; ++r)
#pragma unroll
                for (int q = 0; q < WC; ++q) k3[r][q] = mfma16(a[r].y, sc[q], k3[r][q]);
        }
    };
    if (ntot > 0) {
        if constexpr (DEPTH == 1) {
            gload(0);
            lstore(0, 0);
            __syncthreads();
            int buf = 0;
            for (int t = 0; t < ntot; ++t) {
                const bool more = t + 1 < ntot;
                if (more) gload(0);            // next slab in flight during the MFMAs
                mma(buf);
                if (more) lstore(buf ^ 1, 0);
                __syncthreads();
                buf ^= 1;
            }
        } else {
            // slab t sits in LDS buffer t % 2 and, before that, in register
            // set t % 2; two slabs are in flight during the MFMAs (the loop
            // is unrolled by two so that the register sets are fixed)
            gload(0);
            if (ntot > 1) gload(1);
            lstore(0, 0);
            __syncthreads();
            for (int t = 0; t < ntot; t += 2) {
                if (t + 2 < ntot) gload(0);
                mma(0);
                if (t + 1 < ntot) lstore(1, 1);
                __syncthreads();
                if (t + 1 < ntot) {
                    if (t + 3 < ntot) gload(1);
                    mma(1);
                    if (t + 2 < ntot) lstore(0, 0);
                    __syncthreads();
                }
            }
        }
    }
    for (int r = 0; r < WR; ++r)
        for (int q = 0; q < WC; ++q) {
            const int n = n0 + q * 16 + i;
            if (n >= N) continue;
            if constexpr (sizeof(T) == 4) {
                // FP32 lanes hold four consecutive m: two 16-byte accesses
                // per operand instead of four 8-byte ones (same arithmetic)
                const int mb = m0 + r * 16 + kk * 4;
                if (mb + 3 < M) {
                    float4 o4[2], s4[2] = {};
                    double dg[4] = {};
                    float4* op = reinterpret_cast<float4*>(out + (int64_t)n * ldo + mb);
                    if (accumulate) o4[0] = op[0], o4[1] = op[1];
                    if (diag) {
                        const float4* sp = reinterpret_cast<const float4*>(src + (int64_t)n * lds + mb);
                        s4[0] = sp[0];
                        s4[1] = sp[1];
#pragma unroll
                        for (int v = 0; v < 4; ++v) dg[v] = diag[mb + v];
                    }
                    const float* sv = reinterpret_cast<const float*>(s4);
                    float* ov = reinterpret_cast<float*>(o4);
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        double re = (double)(k1[r][q][v] - k3[r][q][v]);
                        double im = (double)(k1[r][q][v] + k2[r][q][v]);
                        if (diag) {
                            re += dg[v] * (double)sv[2 * v];
                            im += dg[v] * (double)sv[2 * v + 1];
                        }
                        if (accumulate) {
                            re += (double)ov[2 * v];
                            im += (double)ov[2 * v + 1];
                        }
                        ov[2 * v] = (float)re;
                        ov[2 * v + 1] = (float)im;
                    }
                    op[0] = o4[0];
                    op[1] = o4[1];
                    continue;
                }
            }
            for (int v = 0; v < 4; ++v) {
                const int m = m0 + r * 16 + out_row<T>(kk, v);
                if (m >= M) continue;
                double re = (double)(k1[r][q][v] - k3[r][q][v]);
                double im = (double)(k1[r][q][v] + k2[r][q][v]);
                if (diag) {
                    C s = src[(int64_t)n * lds + m];
                    re += diag[m] * (double)s.x;
                    im += diag[m] * (double)s.y;
                }
                C* o = out + (int64_t)n * ldo + m;
                if (accumulate) {
                    C h = *o;
                    re += (double)h.x;
                    im += (double)h.y;
                }
                *o = C{(T)re, (T)im};
            }
        }
}

// GEMM 2: out[n][m] (+)= sum_k A[k][m] B[k][n] (+ A2 B2), per segment, in
// 64 x TN tiles.  Workgroups [0, full) take whole tiles (1D grid renumbered
// so that the column tiles of one row tile run on one XCD); with TN = 64 the
// workgroups past `full` each take a quarter (64 x 16) of one of the
// remaining tiles, so those tiles spread over four times as many CUs.
template <typename T, bool TWO, int TN>
__global__ void __launch_bounds__(256, 2)
k_gemm_apply(const ApplyBatch<T> bt, int N, int K, int accumulate) {
    using C = cx_t<T>;
    constexpr int BK = ApplyShape<T>::BK, LDM = ApplyShape<T>::LDM;
    constexpr int LDB = TN == 64 ? LDM : TN == 32 ? 48 : TN;
    __shared__ __align__(16) C sA[2 * BK * LDM];
    __shared__ __align__(16) C sB[2 * BK * LDB];
    const int ny = (N + TN - 1) / TN;
    int L0, sub = -1;
    if (TN == 64 && (int)blockIdx.x >= bt.full) {
        const int b = (int)blockIdx.x - bt.full;
        L0 = bt.full + b / 4;
        sub = b % 4;
    } else {
        L0 = xcd_order(blockIdx.x, bt.full);
    }
    int seg = 0;
    while (seg + 1 < bt.nseg && L0 >= bt.s[seg + 1].tile0) ++seg;
    const ApplySegK<T>& sg = bt.s[seg];
    const int L = L0 - sg.tile0;
    const int tm0 = (L / ny) * 64, tn0 = (L % ny) * TN;
    if constexpr (TN == 64) {
        if (sub >= 0) {
            if (tn0 + 16 * sub < N) apply_tile<T, TWO, 16>(sg, tm0, tn0 + 16 * sub, N, K, accumulate, sA, sB);
            return;
        }
    }
    apply_tile<T, TWO, TN>(sg, tm0, tn0, N, K, accumulate, sA, sB);
}

// Resident workgroups of the projection GEMM on the whole device.
template <typename T>
static int gemm_proj_slots(Ctx& c) {
    const void* fn = reinterpret_cast<const void*>(&k_gemm_proj<T, 64>);
    auto it = c.occupancy.find(fn);
    if (it == c.occupancy.end()) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, 0) != hipSuccess || per_cu < 1)
            per_cu = 1;
        it = c.occupancy.emplace(fn, per_cu).first;
    }
    return c.ncu * it->second;
}

// Split-K count: at least `wgs` workgroups over the batch (the tuned
// target), then, among up to three times that many splits, the one whose
// last wave of workgroups is fullest (slots = resident workgroups); chunks
// of >= 256 G.
static int choose_split(int wgs, int slots, int M, int N, int K, int nseg) {
    const int tiles = ((M + 63) / 64) * ((N + 63) / 64) * std::max(1, nseg);
    const int maxs = std::max(1, (K + 255) / 256);
    const int smin = std::min(maxs, std::max(1, (wgs + tiles - 1) / tiles));
    const int smax = std::min(maxs, std::max(3, 3 * smin));
    int best = smin;
    double beff = -1.0;
    for (int sp = smin; sp <= smax; ++sp) {
        const double w = (double)tiles * sp;
        const double eff = w / (std::ceil(w / slots) * slots);   // busy fraction of the waves
        if (eff > beff + 1e-9) best = sp, beff = eff;
    }
    return best;
}

// Partial-sum bytes for the largest split any candidate target can choose,
// so that re-tuning never needs a reallocation.
template <typename T>
size_t vnl_partial_bytes(Ctx& c, int nrows, int nb, int min_npw, int nseg) {
    const int slots = gemm_proj_slots<T>(c);
    int ns = choose_split(c.tuning<T>().gemm_wgs, slots, nrows, nb, min_npw, nseg);
    for (int w : kGemmWgChoices) ns = std::max(ns, choose_split(w, slots, nrows, nb, min_npw, nseg));
    return (size_t)ns * nrows * nb * sizeof(cx_t<T>) * nseg;
}
template size_t vnl_partial_bytes<double>(Ctx&, int, int, int, int);
template size_t vnl_partial_bytes<float>(Ctx&, int, int, int, int);

static int min_npw(const ApplySeg* segs, int nseg) {
    int m = segs[0].k->npw;
    for (int i = 1; i < nseg; ++i) m = std::min(m, segs[i].k->npw);
    return m;
}

// <A_s|psi_s> for every segment: rows [row0, row0 + M) of each segment's
// partial buffer part_s = partial + s * nsplit * Mtot * nb.  With A2, rows
// [M1, M) of the product come from A2 (the projector k-gradients after the
// projectors: one launch for all 4 nch rows, so the split fills the device
// for the whole projection rather than for each part).
template <typename T>
static hipError_t gemm_proj(Ctx& c, const ApplySeg* segs, int nseg, const cx_t<T>* const* A, int M,
                            int nb, cx_t<T>* partial, int Mtot, int row0, int nsplit,
                            const cx_t<T>* const* A2 = nullptr, int M1 = 0) {
    ProjBatch<T> bt{};
    bt.nseg = nseg;
    for (int i = 0; i < nseg; ++i) {
        const int npw = segs[i].k->npw;
        int kchunk = (npw + nsplit - 1) / nsplit;
        kchunk = (kchunk + 15) / 16 * 16;
        bt.s[i] = ProjSeg<T>{A[i], static_cast<const cx_t<T>*>(segs[i].psi),
                             partial + (size_t)i * nsplit * Mtot * nb, (int64_t)npw, segs[i].ld, npw, kchunk,
                             A2 ? A2[i] : nullptr, A2 ? M1 : M};
    }
    // nb <= 32: one 64 x 16 or 64 x 32 tile column (same grid as one 64 x 64 one)
    dim3 grid(((M + 63) / 64) * ((nb + 63) / 64) * nsplit * nseg);
    switch (gemm_tile_width(nb)) {
        case 16:
            hipLaunchKernelGGL((k_gemm_proj<T, 16>), grid, dim3(256), 0, c.stream, bt, M, nb, nsplit, Mtot, row0);
            break;
        case 32:
            hipLaunchKernelGGL((k_gemm_proj<T, 32>), grid, dim3(256), 0, c.stream, bt, M, nb, nsplit, Mtot, row0);
            break;
        default:
            hipLaunchKernelGGL((k_gemm_proj<T, 64>), grid, dim3(256), 0, c.stream, bt, M, nb, nsplit, Mtot, row0);
    }
    return hipGetLastError();
}

template <typename T>
static hipError_t reduce_scale(Ctx& c, int nseg, cx_t<T>* partial, int nsplit, int Mtot, int nb,
                               cx_t<T>* cbuf) {
    ReduceBatch<T> bt{};
    bt.nseg = nseg;
    for (int i = 0; i < nseg; ++i) {
        bt.part[i] = partial + (size_t)i * nsplit * Mtot * nb;
        bt.out[i] = cbuf + (size_t)i * Mtot * nb;
    }
    return launch_reduce<T>(c, bt, nseg, nsplit, Mtot, nb, static_cast<const double*>(c.lam64.ptr), c.nch(),
                            (int64_t)nb);
}

// out_s (+)= diag_s * src_s + A_s B_s for the segments with out_s != NULL
template <typename T>
static hipError_t gemm_apply(Ctx& c, int nseg, const ApplySegK<T>* in, int nb, int K, bool acc) {
    ApplyBatch<T> bt{};
    int tiles = 0;
    // above 32 bands: 64 x 64 tiles, or 64 x 32 ones (the tuned width)
    const bool narrow = nb > 32 && c.tuning<T>().apply_tile == 32;
    const int ny = narrow ? (nb + 31) / 32 : (nb + 63) / 64;
    for (int i = 0; i < nseg; ++i) {
        if (!in[i].out) continue;
        ApplySegK<T>& s = bt.s[bt.nseg++];
        s = in[i];
        s.tile0 = tiles;
        tiles += ((s.M + 63) / 64) * ny;
    }
    if (bt.nseg == 0 || tiles == 0) return hipSuccess;
    // all segments of a call either have the second product or none; nb <=
    // 16 takes the 64 x 16 tile (the same tile count as 64 x 64 there)
    const int a = acc ? 1 : 0;
    bt.full = tiles;
    if (nb <= 32) {
        // one 64 x 16 or 64 x 32 tile column: the same tile count as 64 x 64
        if (nb <= 16) {
            if (bt.s[0].A2)
                hipLaunchKernelGGL((k_gemm_apply<T, true, 16>), dim3(tiles), dim3(256), 0, c.stream, bt, nb, K, a);
            else
                hipLaunchKernelGGL((k_gemm_apply<T, false, 16>), dim3(tiles), dim3(256), 0, c.stream, bt, nb, K, a);
        } else {
            if (bt.s[0].A2)
                hipLaunchKernelGGL((k_gemm_apply<T, true, 32>), dim3(tiles), dim3(256), 0, c.stream, bt, nb, K, a);
            else
                hipLaunchKernelGGL((k_gemm_apply<T, false, 32>), dim3(tiles), dim3(256), 0, c.stream, bt, nb, K, a);
        }
        return hipGetLastError();
    }
    if (narrow) {
        if (bt.s[0].A2)
            hipLaunchKernelGGL((k_gemm_apply<T, true, 32>), dim3(tiles), dim3(256), 0, c.stream, bt, nb, K, a);
        else
            hipLaunchKernelGGL((k_gemm_apply<T, false, 32>), dim3(tiles), dim3(256), 0, c.stream, bt, nb, K, a);
        return hipGetLastError();
    }
    // Workgroups on one CU share its MFMA pipes, so the kernel lasts as long
    // as the CU with the most tiles: with tiles = q * CUs + r, r small, r CUs
    // take a whole tile more than the rest.  Those r tiles run as 4 r
    // quarter tiles after the others instead (measured on the 64-atom apply,
    // 2312 tiles = 9 * 256 + 8: the 8 extra tiles cost 6.6% of the kernel).
    const int r = tiles % c.ncu;
    if (r > 0 && 4 * r <= c.ncu) bt.full = tiles - r;
    const int grid = bt.full + 4 * (tiles - bt.full);
    if (bt.s[0].A2)
        hipLaunchKernelGGL((k_gemm_apply<T, true, 64>), dim3(grid), dim3(256), 0, c.stream, bt, nb, K, a);
    else
        hipLaunchKernelGGL((k_gemm_apply<T, false, 64>), dim3(grid), dim3(256), 0, c.stream, bt, nb, K, a);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_vnl(Ctx& c, const ApplySeg* segs, int nseg, int nb, void* partial, size_t partial_bytes,
                      void* cbuf) {
    using C = cx_t<T>;
    const int nch = c.nch();
    if (nch == 0) return hipSuccess;
    if (nseg < 1 || nseg > kMaxSeg) return hipErrorInvalidValue;
    const bool f64 = sizeof(T) == 8;
    const C* P[kMaxSeg];
    for (int i = 0; i < nseg; ++i) P[i] = static_cast<const C*>(f64 ? segs[i].k->P64.ptr : segs[i].k->P32.ptr);
    const int nsplit = choose_split(c.tuning<T>().gemm_wgs, gemm_proj_slots<T>(c), nch, nb, min_npw(segs, nseg), nseg);
    if ((size_t)nsplit * nch * nb * sizeof(C) * nseg > partial_bytes) return hipErrorInvalidValue;
    C* part = static_cast<C*>(partial);
    C* cb = static_cast<C*>(cbuf);
    timer_begin(c);
    hipError_t e = gemm_proj<T>(c, segs, nseg, P, nch, nb, part, nch, 0, nsplit);
    timer_end(c, "proj_gemm");
    if (e != hipSuccess) return e;
    timer_begin(c);
    e = reduce_scale<T>(c, nseg, part, nsplit, nch, nb, cb);
    timer_end(c, "proj_reduce");
    if (e != hipSuccess) return e;
    ApplySegK<T> ap[kMaxSeg];
    for (int i = 0; i < nseg; ++i)
        ap[i] = ApplySegK<T>{P[i], cb + (size_t)i * nch * nb, static_cast<C*>(segs[i].out[0]), nullptr, nullptr,
                             nullptr, nullptr, (int64_t)segs[i].k->npw, (int64_t)nb, segs[i].ldo, 0,
                             segs[i].k->npw, 0};
    timer_begin(c);
    e = gemm_apply<T>(c, nseg, ap, nb, nch, true);
    timer_end(c, "proj_apply");
    return e;
}

template <typename T>
hipError_t launch_dhdk(Ctx& c, const ApplySeg* segs, int nseg, int nb, bool accumulate, void* partial,
                       size_t partial_bytes, void* cbuf, bool with_h) {
    using C = cx_t<T>;
    const int nch = c.nch();
    const bool f64 = sizeof(T) == 8;
    if (nseg < 1 || nseg > kMaxSeg) return hipErrorInvalidValue;
    hipError_t e = hipSuccess;
    ApplySegK<T> ap[kMaxSeg];
    if (nch == 0) {
        // kinetic part only: out_a (+)= K_a psi  (gradient.cpp:139-141)
        for (int a = 0; a < 3 && e == hipSuccess; ++a) {
            for (int i = 0; i < nseg; ++i) {
                const KptDev& k = *segs[i].k;
                ap[i] = ApplySegK<T>{nullptr, nullptr, static_cast<C*>(segs[i].out[a]), nullptr, nullptr,
                                     static_cast<const double*>(k.K.ptr) + (size_t)a * k.npw,
                                     static_cast<const C*>(segs[i].psi), 0, 0, segs[i].ldo, segs[i].ld, k.npw, 0};
            }
            e = gemm_apply<T>(c, nseg, ap, nb, 0, accumulate);
        }
        return e;
    }
    const C* P[kMaxSeg];
    const C* dP[kMaxSeg];
    for (int i = 0; i < nseg; ++i) {
        P[i] = static_cast<const C*>(f64 ? segs[i].k->P64.ptr : segs[i].k->P32.ptr);
        dP[i] = static_cast<const C*>(f64 ? segs[i].k->dP64.ptr : segs[i].k->dP32.ptr);
    }
    const int Mtot = 4 * nch;
    const int nsplit = choose_split(c.tuning<T>().gemm_wgs, gemm_proj_slots<T>(c), Mtot, nb, min_npw(segs, nseg),
                                    nseg);
    if ((size_t)nsplit * Mtot * nb * sizeof(C) * nseg > partial_bytes) return hipErrorInvalidValue;
    C* part = static_cast<C*>(partial);
    C* cb = static_cast<C*>(cbuf);
    // <p|psi> and <dp_a|psi> for a = x, y, z in one launch into one partial buffer
    timer_begin(c);
    e = gemm_proj<T>(c, segs, nseg, P, Mtot, nb, part, Mtot, 0, nsplit, dP, nch);
    timer_end(c, "proj_gemm");
    if (e != hipSuccess) return e;
    timer_begin(c);
    e = reduce_scale<T>(c, nseg, part, nsplit, Mtot, nb, cb);
    timer_end(c, "proj_reduce");
    if (e != hipSuccess) return e;
    timer_begin(c);
    if (with_h) {
        // hpsi += sum_c p_c lambda_c <p_c|psi>, the first block of the projection
        for (int i = 0; i < nseg; ++i)
            ap[i] = ApplySegK<T>{P[i], cb + (size_t)i * Mtot * nb, static_cast<C*>(segs[i].hout), nullptr,
                                 nullptr, nullptr, nullptr, (int64_t)segs[i].k->npw, (int64_t)nb, segs[i].ldo, 0,
                                 segs[i].k->npw, 0};
        e = gemm_apply<T>(c, nseg, ap, nb, nch, true);
    }
    for (int a = 0; a < 3 && e == hipSuccess; ++a) {
        // out_a (+)= K_a psi + sum_c p_c lambda_c <dp_a,c|psi> + sum_c dp_a,c lambda_c <p_c|psi>:
        // one launch, the two projector products concatenated along k
        for (int i = 0; i < nseg; ++i) {
            const KptDev& k = *segs[i].k;
            const C* cbs = cb + (size_t)i * Mtot * nb;
            ap[i] = ApplySegK<T>{P[i], cbs + (size_t)(1 + a) * nch * nb, static_cast<C*>(segs[i].out[a]),
                                 dP[i] + (size_t)a * nch * k.npw, cbs,
                                 static_cast<const double*>(k.K.ptr) + (size_t)a * k.npw,
                                 static_cast<const C*>(segs[i].psi), (int64_t)k.npw, (int64_t)nb, segs[i].ldo,
                                 segs[i].ld, k.npw, 0};
        }
        e = gemm_apply<T>(c, nseg, ap, nb, nch, accumulate);
    }
    timer_end(c, "proj_apply");
    return e;
}

// C[m][n] = sum_g conj(A[m][g]) B[n][g] (no lambda), split-K partial sums
// reduced in FP64: the overlap matrices of an eigensolver's Rayleigh-Ritz
// step, whose K = npw is long and M, N are band counts.
// Split count of an overlap: about 512 workgroups, partial sums capped at
// 32 MiB (an eigensolver's context runs under a small HBM budget).
template <typename T>
static int overlap_split(Ctx& c, int m, int n, int k) {
    const int ns = choose_split(512, gemm_proj_slots<T>(c), m, n, k, 1);
    const size_t per = (size_t)m * n * sizeof(cx_t<T>);
    const int cap = (int)std::max<size_t>(1, kOverlapPartialCap / std::max<size_t>(per, 1));
    return std::max(1, std::min(ns, cap));
}

template <typename T>
size_t overlap_partial_bytes(Ctx& c, int m, int n, int k) {
    return (size_t)overlap_split<T>(c, m, n, k) * m * n * sizeof(cx_t<T>);
}

template <typename T>
hipError_t launch_overlap(Ctx& c, int m, int n, int k, const void* A, int64_t lda, const void* B, int64_t ldb,
                          void* C, int64_t ldc, void* partial, size_t partial_bytes, bool timed) {
    using Cx = cx_t<T>;
    const int nsplit = overlap_split<T>(c, m, n, k);
    if ((size_t)nsplit * m * n * sizeof(Cx) > partial_bytes) return hipErrorInvalidValue;
    int kchunk = (k + nsplit - 1) / nsplit;
    kchunk = (kchunk + 15) / 16 * 16;
    ProjBatch<T> bt{};
    bt.nseg = 1;
    bt.s[0] = ProjSeg<T>{static_cast<const Cx*>(A), static_cast<const Cx*>(B), static_cast<Cx*>(partial), lda, ldb,
                         k, kchunk, nullptr, m};
    if (timed) timer_begin(c);
    dim3 grid(((m + 63) / 64) * ((n + 63) / 64) * nsplit);
    switch (gemm_tile_width(n)) {
        case 16: hipLaunchKernelGGL((k_gemm_proj<T, 16>), grid, dim3(256), 0, c.stream, bt, m, n, nsplit, m, 0); break;
        case 32: hipLaunchKernelGGL((k_gemm_proj<T, 32>), grid, dim3(256), 0, c.stream, bt, m, n, nsplit, m, 0); break;
        default: hipLaunchKernelGGL((k_gemm_proj<T, 64>), grid, dim3(256), 0, c.stream, bt, m, n, nsplit, m, 0);
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        ReduceBatch<T> rb{};
        rb.nseg = 1;
        rb.part[0] = static_cast<const Cx*>(partial);
        rb.out[0] = static_cast<Cx*>(C);
        e = launch_reduce<T>(c, rb, 1, nsplit, m, n, nullptr, 1, ldc);
    }
    if (timed) timer_end(c, "overlap");
    return e;
}

// Resident-workgroup counts of the GEMM kernels, cached at context creation
// so that later launches (effm_overlap_noalloc among them) only read the cache.
void prime_occupancy(Ctx& c) {
    (void)gemm_proj_slots<double>(c);
    (void)gemm_proj_slots<float>(c);
}

template size_t overlap_partial_bytes<double>(Ctx&, int, int, int);
template size_t overlap_partial_bytes<float>(Ctx&, int, int, int);
template hipError_t launch_overlap<double>(Ctx&, int, int, int, const void*, int64_t, const void*, int64_t, void*,
                                           int64_t, void*, size_t, bool);
template hipError_t launch_overlap<float>(Ctx&, int, int, int, const void*, int64_t, const void*, int64_t, void*,
                                          int64_t, void*, size_t, bool);

template hipError_t launch_vnl<double>(Ctx&, const ApplySeg*, int, int, void*, size_t, void*);
template hipError_t launch_vnl<float>(Ctx&, const ApplySeg*, int, int, void*, size_t, void*);
template hipError_t launch_dhdk<double>(Ctx&, const ApplySeg*, int, int, bool, void*, size_t, void*, bool);
template hipError_t launch_dhdk<float>(Ctx&, const ApplySeg*, int, int, bool, void*, size_t, void*, bool);

// ---------------------------------------------------------------------------
//  FP64 Rayleigh quotients
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256)
k_rayleigh(const cx_t<T>* __restrict__ psi, int64_t ld, const cx_t<T>* __restrict__ h,
           int64_t ldh, int npw, double* __restrict__ E) {
    __shared__ double sn[256], sd[256];
    const int b = blockIdx.x;
    double num = 0.0, den = 0.0;
    for (int g = threadIdx.x; g < npw; g += blockDim.x) {
        cx_t<T> p = psi[(int64_t)b * ld + g];
        cx_t<T> q = h[(int64_t)b * ldh + g];
        num += (double)p.x * (double)q.x + (double)p.y * (double)q.y;
        den += (double)p.x * (double)p.x + (double)p.y * (double)p.y;
    }
    sn[threadIdx.x] = num;
    sd[threadIdx.x] = den;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sn[threadIdx.x] += sn[threadIdx.x + s];
            sd[threadIdx.x] += sd[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) E[b] = sn[0] / sd[0];
}

template <typename T>
hipError_t launch_rayleigh(hipStream_t s, int npw, int nb, const void* psi, int64_t ld,
                           const void* hpsi, int64_t ldh, double* E) {
    hipLaunchKernelGGL((k_rayleigh<T>), dim3(nb), dim3(256), 0, s,
                       static_cast<const cx_t<T>*>(psi), ld, static_cast<const cx_t<T>*>(hpsi), ldh,
                       npw, E);
    return hipGetLastError();
}

template hipError_t launch_rayleigh<double>(hipStream_t, int, int, const void*, int64_t,
                                            const void*, int64_t, double*);
template hipError_t launch_rayleigh<float>(hipStream_t, int, int, const void*, int64_t,
                                           const void*, int64_t, double*);

// ---------------------------------------------------------------------------
//  k.p sum over states in FP64: one workgroup per (selected band n, a, b)
//    M_ab = (d2h_ab + d2h_ba) / 2
//         + sum_{m: |E_n - E_m| > tol, m != n} Re(V_a[n][m] V_b[m][n] + V_b[n][m] V_a[m][n]) / (E_n - E_m)
//  (the symmetrised 2 Re sum of the k.p expression, denominators in FP64)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_kp_sum(const KpBands bs, int nb, const double* __restrict__ E, const double2* __restrict__ V0,
         const double2* __restrict__ V1, const double2* __restrict__ V2, int64_t ldv,
         const double* __restrict__ d2h, double tol, double* __restrict__ out) {
    __shared__ double red[256];
    const int sel = blockIdx.x, a = blockIdx.y / 3, b = blockIdx.y % 3;
    const int n = bs.n[sel];
    const double2* Va = a == 0 ? V0 : (a == 1 ? V1 : V2);
    const double2* Vb = b == 0 ? V0 : (b == 1 ? V1 : V2);
    const double en = E[n];
    double acc = 0.0;
    for (int m = threadIdx.x; m < nb; m += blockDim.x) {
        const double d = en - E[m];
        if (m == n || !(fabs(d) > tol)) continue;
        const double2 x = Va[(int64_t)n * ldv + m], y = Vb[(int64_t)m * ldv + n];
        const double2 u = Vb[(int64_t)n * ldv + m], v = Va[(int64_t)m * ldv + n];
        acc += ((x.x * y.x - x.y * y.y) + (u.x * v.x - u.y * v.y)) / d;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double* dd = d2h + (size_t)sel * 9;
        out[(size_t)sel * 9 + a * 3 + b] = 0.5 * (dd[a * 3 + b] + dd[b * 3 + a]) + red[0];
    }
}

hipError_t launch_kp_sum(hipStream_t s, const KpBands& bs, int nsel, int nb, const double* E,
                         const void* const* V, int64_t ldv, const double* d2h, double tol, double* out) {
    hipLaunchKernelGGL(k_kp_sum, dim3(nsel, 9), dim3(256), 0, s, bs, nb, E, static_cast<const double2*>(V[0]),
                       static_cast<const double2*>(V[1]), static_cast<const double2*>(V[2]), ldv, d2h, tol, out);
    return hipGetLastError();
}

// z <- a z + b y + c x elementwise over n complex values, in FP64 for either
// storage precision: one pass where three scale / add passes would read z
// three times (the eigensolver's Chebyshev three-term recurrence).
template <typename T>
__global__ void __launch_bounds__(256)
k_lincomb(int64_t n, double a, cx_t<T>* __restrict__ z, double b, const cx_t<T>* __restrict__ y, double c,
          const cx_t<T>* __restrict__ x) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const cx_t<T> zv = z[i], yv = y[i], xv = x[i];
        z[i] = cx_t<T>{(T)(a * (double)zv.x + b * (double)yv.x + c * (double)xv.x),
                       (T)(a * (double)zv.y + b * (double)yv.y + c * (double)xv.y)};
    }
}

template <typename T>
hipError_t launch_lincomb(Ctx& c, int64_t n, double a, void* z, double b, const void* y, double cc,
                          const void* x) {
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, (int64_t)c.ncu * 16);
    hipLaunchKernelGGL((k_lincomb<T>), dim3((unsigned)std::max<int64_t>(blocks, 1)), dim3(256), 0, c.stream, n, a,
                       static_cast<cx_t<T>*>(z), b, static_cast<const cx_t<T>*>(y), cc,
                       static_cast<const cx_t<T>*>(x));
    return hipGetLastError();
}
template hipError_t launch_lincomb<double>(Ctx&, int64_t, double, void*, double, const void*, double, const void*);
template hipError_t launch_lincomb<float>(Ctx&, int64_t, double, void*, double, const void*, double, const void*);

}  // namespace effm
