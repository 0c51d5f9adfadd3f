// fft_ct.hpp — compile-time two-stage plans for the common grid axes.
//
// Same in-place algorithm and digit-reversed order as fft_lds.hpp (so the
// host-side permutations are shared), but the length N = P1 * P2, the line
// count and the workgroup size are template parameters: every index
// division folds to a multiply/shift, the radix dispatch disappears and the
// per-thread loop is fully unrolled.  A two-stage plan is one LDS round trip
// per radix, i.e. two per axis.
#pragma once
#include "fft_lds.hpp"

namespace effm {

// Radix pair (first DIT stage P1, second P2) of the specialised axes.  Must
// match factor_axis() on the host for these sizes.
template <int N> struct Plan2 { static constexpr bool ok = false; static constexpr int P1 = 1, P2 = 1; };
#define EFFM_PLAN2(N, A, B) \
    template <> struct Plan2<N> { static constexpr bool ok = true; static constexpr int P1 = A, P2 = B; };
EFFM_PLAN2(15, 5, 3)
EFFM_PLAN2(16, 4, 4)
EFFM_PLAN2(18, 6, 3)
EFFM_PLAN2(20, 5, 4)
EFFM_PLAN2(24, 6, 4)
EFFM_PLAN2(25, 5, 5)
EFFM_PLAN2(36, 6, 6)
EFFM_PLAN2(40, 8, 5)
EFFM_PLAN2(45, 9, 5)
EFFM_PLAN2(48, 8, 6)
EFFM_PLAN2(90, 10, 9)
#undef EFFM_PLAN2

// The axis lengths above (the 64-atom cell's 90, the 8-atom 48, and the
// grids of 2-64 atom cells at 30-50 Ry); X-macro for dispatch.
#define EFFM_CT_SIZES(X) X(15) X(16) X(18) X(20) X(24) X(25) X(36) X(40) X(45) X(48) X(90)

// One radix-P stage (combining blocks of L) over NL lines of an N-point axis;
// element e of line l at l*LS + e*ES.  Lines vary fastest across threads when
// the axis is strided (ES > 1), butterflies otherwise.
template <int N, int P, int L, int S, bool DIT, int ES, int LS, int NL, int BLOCK, typename T>
__device__ __forceinline__ void ct_stage(cx_t<T>* buf, const cx_t<T>* tw) {
    constexpr int NBF = N / P;
    constexpr int TOTAL = NL * NBF;
    constexpr int LT = L * P;
    constexpr int TSTEP = N / LT;
    constexpr int ITER = (TOTAL + BLOCK - 1) / BLOCK;
#pragma unroll
    for (int it = 0; it < ITER; ++it) {
        const int t = threadIdx.x + it * BLOCK;
        if (ITER * BLOCK == TOTAL || t < TOTAL) {
            int line, r;
            if constexpr (ES != 1) {
                r = t / NL;
                line = t - r * NL;
            } else {
                line = t / NBF;
                r = t - line * NBF;
            }
            const int b = r / L;
            const int k1 = r - b * L;
            const int base = line * LS + (b * LT + k1) * ES;
            butterfly<P, S, DIT, T>(buf, base, L * ES, k1, TSTEP, tw);
        }
    }
}

// Full N-point transform along one axis; ends with a barrier.
template <int N, int S, bool DIT, int ES, int LS, int NL, int BLOCK, typename T>
__device__ __forceinline__ void ct_fft(cx_t<T>* buf, const cx_t<T>* tw) {
    constexpr int P1 = Plan2<N>::P1, P2 = Plan2<N>::P2;
    if constexpr (DIT) {
        ct_stage<N, P1, 1, S, true, ES, LS, NL, BLOCK, T>(buf, tw);
        __syncthreads();
        ct_stage<N, P2, P1, S, true, ES, LS, NL, BLOCK, T>(buf, tw);
        __syncthreads();
    } else {
        ct_stage<N, P2, P1, S, false, ES, LS, NL, BLOCK, T>(buf, tw);
        __syncthreads();
        ct_stage<N, P1, 1, S, false, ES, LS, NL, BLOCK, T>(buf, tw);
        __syncthreads();
    }
}

// The contiguous-axis transform of only the listed rows of an N x N plane
// (rows[0 .. nrows) are row indices, e.g. in LDS): same stages as ct_fft
// with ES = 1, LS = N.  Rows that hold no data (G2R) or whose results are
// not gathered (R2G) are skipped.  Ends with a barrier.
template <int N, int P, int L, int S, bool DIT, int BLOCK, typename T>
__device__ __forceinline__ void ct_stage_rows(cx_t<T>* buf, const cx_t<T>* tw, const int* rows, int nrows) {
    constexpr int NBF = N / P;
    constexpr int LT = L * P;
    constexpr int TSTEP = N / LT;
    constexpr int ITER = (N * NBF + BLOCK - 1) / BLOCK;
    const int total = nrows * NBF;
#pragma unroll
    for (int it = 0; it < ITER; ++it) {
        const int t = threadIdx.x + it * BLOCK;
        if (t < total) {
            const int li = t / NBF;
            const int r = t - li * NBF;
            const int b = r / L;
            const int k1 = r - b * L;
            butterfly<P, S, DIT, T>(buf, rows[li] * N + b * LT + k1, L, k1, TSTEP, tw);
        }
    }
}

template <int N, int S, bool DIT, int BLOCK, typename T>
__device__ __forceinline__ void ct_fft_rows(cx_t<T>* buf, const cx_t<T>* tw, const int* rows, int nrows) {
    constexpr int P1 = Plan2<N>::P1, P2 = Plan2<N>::P2;
    if constexpr (DIT) {
        ct_stage_rows<N, P1, 1, S, true, BLOCK, T>(buf, tw, rows, nrows);
        __syncthreads();
        ct_stage_rows<N, P2, P1, S, true, BLOCK, T>(buf, tw, rows, nrows);
        __syncthreads();
    } else {
        ct_stage_rows<N, P2, P1, S, false, BLOCK, T>(buf, tw, rows, nrows);
        __syncthreads();
        ct_stage_rows<N, P1, 1, S, false, BLOCK, T>(buf, tw, rows, nrows);
        __syncthreads();
    }
}

constexpr int round_up(int a, int b) { return (a + b - 1) / b * b; }
constexpr int cmax(int a, int b) { return a > b ? a : b; }
constexpr int cmin(int a, int b) { return a < b ? a : b; }

// Workgroup size for a square N x N plane: one butterfly per thread and
// stage, rounded to whole waves, at most 1024.
template <int N>
constexpr int plane_block() {
    return cmin(1024, round_up(N * cmax(N / Plan2<N>::P1, N / Plan2<N>::P2), 64));
}

}  // namespace effm
