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
EFFM_PLAN2(27, 9, 3)
EFFM_PLAN2(32, 8, 4)
EFFM_PLAN2(36, 6, 6)
EFFM_PLAN2(40, 8, 5)
EFFM_PLAN2(45, 9, 5)
EFFM_PLAN2(48, 8, 6)
EFFM_PLAN2(80, 10, 8)
EFFM_PLAN2(90, 10, 9)
EFFM_PLAN2(96, 12, 8)
#undef EFFM_PLAN2

// Prime-factor axes: with gcd(P1, P2) = 1 the two stages need no twiddle
// factors if the inputs sit in the Good-Thomas order and the outputs are
// read in the Chinese-remainder order (see dft_pfa).  In place, input
// natural index (P2 u + P1 b) mod N sits at b P1 + u (the host builds the
// plane scatter / gather positions that way), and after both DIT stages
// position k1 + P1 q holds natural index nat(k1, q) = (P2 a k1 + P1 c q)
// mod N (a = P2^-1 mod P1, c = P1^-1 mod P2), i.e. natural index x sits
// at pos(x) = x mod P1 + P1 (x mod P2).  The DIF stages run the same maps
// backwards.
template <int N> struct Pfa {
    static constexpr int P1 = Plan2<N>::P1, P2 = Plan2<N>::P2;
    static constexpr bool on = Plan2<N>::ok && gcd_c(P1, P2) == 1;
    static constexpr int A = inv_mod(P2 % P1, P1), C = inv_mod(P1 % P2, P2);
    static __device__ __forceinline__ int nat(int k1, int q) { return (P2 * A * k1 + P1 * C * q) % N; }
    static __device__ __forceinline__ int pos(int x) { return x % P1 + P1 * (x % P2); }
};

// The axis lengths above (the 64-atom cell's 90, the 8-atom 48, and the
// density-sphere grids of 2-64 atom cells at 30-50 Ry: 27, 32, 36, 40, 45,
// 48, 80, 90, 96); X-macro for dispatch.
#define EFFM_CT_SIZES(X) \
    X(15) X(16) X(18) X(20) X(24) X(25) X(27) X(32) X(36) X(40) X(45) X(48) X(80) X(90) X(96)

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
            butterfly<P, S, DIT, T>(buf, base, L * ES, Pfa<N>::on ? 0 : k1, TSTEP, tw);
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
            // a 24-bit multiply: as a 32-bit one it became a 64-bit mad whose
            // unused high addend half was allocated to the destination of an
            // in-flight load, and the wait for that load stalled the row pass
            butterfly<P, S, DIT, T>(buf, (int)__umul24((unsigned)rows[li], (unsigned)N) + b * LT + k1, L,
                                    Pfa<N>::on ? 0 : k1, TSTEP, tw);
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

// The radix-P2 stage of the column (strided) transform of an N x N plane
// with global memory on one side: the last DIT stage writing its results
// straight to the rows of a global plane (row pitch WP >= N), or the first
// DIF stage taking its inputs from one.  With N = P1 * P2 that stage has P1
// butterflies per column; work item t takes column t % W (consecutive lanes
// on consecutive columns, so the global accesses of one instruction are row
// segments) and butterfly r = t / W, whose elements are rows r + u * P1 in
// LDS (in global memory rows Pfa<N>::nat(r, u) on prime-factor axes, whose
// LDS columns are then permuted too: natural column x at Pfa<N>::pos(x)).
// Each skips an LDS round trip and a barrier against ct_fft + a copy.
template <int N, int W, int BLOCK>
struct ColStage {
    static constexpr int P1 = Plan2<N>::P1, P2 = Plan2<N>::P2;
    static constexpr int TOTAL = W * P1;
    static constexpr int ITER = (TOTAL + BLOCK - 1) / BLOCK;
};

// Streaming (nontemporal) store of one complex value: the plane workspace
// is written once and read by the next kernel, so it need not stay in the
// caches.  (Measured on the 90^3 workspace, 1.36 GB FP64: the G2R pass 1-3%
// faster and the x-pencil pass that reads it 2-3% faster than with plain
// stores; nontemporal stores in the x-pencil pass itself were slower.)
template <typename C>
__device__ __forceinline__ void store_stream(C v, C* p) {
    using E = decltype(v.x);
    typedef E v2 __attribute__((ext_vector_type(2)));
    const v2 w = {v.x, v.y};
    __builtin_nontemporal_store(w, reinterpret_cast<v2*>(p));
}

// Streaming (nontemporal) load of one complex value: a plane is read once,
// so it should not push the arrays every plane reuses (sphere indices,
// kinetic energies, psi) out of L2.  (Measured on the 64-atom R2G pass:
// 249 -> 213 us mixed, 412 -> 403 us FP64, fetch 1003 -> 969 MB mixed.  In
// the x-pencil pass, whose neighbouring z-chunks share lines, they were
// slower: FP64 570 -> 726 us, fetch 1.55 -> 1.93 GB.)
template <typename C>
__device__ __forceinline__ C load_stream(const C* p) {
    using E = decltype(p->x);
    typedef E v2 __attribute__((ext_vector_type(2)));
    const v2 w = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
    return C{w.x, w.y};
}

// G2R: last DIT stage of the columns, from LDS to out[row * rs + col]
// (rs: the workspace's row stride, >= WP).  The pad columns N <= col < WP
// repeat column N - 1 (whole rows are written).
template <int N, int WP, int S, int BLOCK, typename T>
__device__ __forceinline__ void ct_cols_last_to_global(const cx_t<T>* buf, const cx_t<T>* tw, cx_t<T>* out,
                                                       int64_t rs) {
    using Cs = ColStage<N, WP, BLOCK>;
#pragma unroll
    for (int it = 0; it < Cs::ITER; ++it) {
        const int t = threadIdx.x + it * BLOCK;
        if (Cs::ITER * BLOCK == Cs::TOTAL || t < Cs::TOTAL) {
            const int r = t / WP, col = t - r * WP;
            const int c = min(col, N - 1);
            const int cp = Pfa<N>::on ? Pfa<N>::pos(c) : c;
            cx_t<T> y[Cs::P2];
#pragma unroll
            for (int u = 0; u < Cs::P2; ++u) y[u] = buf[(r + u * Cs::P1) * N + cp];
            if (!Pfa<N>::on && r != 0) {
#pragma unroll
                for (int u = 1; u < Cs::P2; ++u) y[u] = cmul(y[u], twid<S>(tw, u * r));
            }
            dft<Cs::P2, S, T>(y);
#pragma unroll
            for (int u = 0; u < Cs::P2; ++u) {
                const int row = Pfa<N>::on ? Pfa<N>::nat(r, u) : r + u * Cs::P1;
                store_stream(y[u], &out[(int64_t)row * rs + col]);
            }
        }
    }
}

// R2G: the inputs u in [U0, U1) of this thread's first DIF column
// butterflies, from in[row * rs + col], col < N.  Spare work items load a
// valid duplicate.
template <int N, int WP, int BLOCK, typename T, int U0 = 0, int U1 = Plan2<N>::P2>
__device__ __forceinline__ void ct_cols_first_load(const cx_t<T>* in,
                                                   cx_t<T> (&y)[ColStage<N, N, BLOCK>::ITER][Plan2<N>::P2],
                                                   int64_t rs) {
    using Cs = ColStage<N, N, BLOCK>;
#pragma unroll
    for (int it = 0; it < Cs::ITER; ++it) {
        const int t = min((int)threadIdx.x + it * BLOCK, Cs::TOTAL - 1);
        const int r = t / N, col = t - r * N;
#pragma unroll
        for (int u = U0; u < U1; ++u) {
            const int row = Pfa<N>::on ? Pfa<N>::nat(r, u) : r + u * Cs::P1;
            y[it][u] = load_stream(&in[(int64_t)row * rs + col]);
        }
    }
}

// R2G: first DIF stage of the columns on the loaded inputs, results to the
// LDS plane (no barrier).
template <int N, int S, int BLOCK, typename T>
__device__ __forceinline__ void ct_cols_first_to_lds(cx_t<T>* buf, const cx_t<T>* tw,
                                                     cx_t<T> (&y)[ColStage<N, N, BLOCK>::ITER][Plan2<N>::P2]) {
    using Cs = ColStage<N, N, BLOCK>;
#pragma unroll
    for (int it = 0; it < Cs::ITER; ++it) {
        const int t = threadIdx.x + it * BLOCK;
        if (Cs::ITER * BLOCK == Cs::TOTAL || t < Cs::TOTAL) {
            const int r = t / N, col = t - r * N;
            dft<Cs::P2, S, T>(y[it]);
            if (!Pfa<N>::on && r != 0) {
#pragma unroll
                for (int u = 1; u < Cs::P2; ++u) y[it][u] = cmul(y[it][u], twid<S>(tw, u * r));
            }
            const int cp = Pfa<N>::on ? Pfa<N>::pos(col) : col;
#pragma unroll
            for (int u = 0; u < Cs::P2; ++u) buf[(r + u * Cs::P1) * N + cp] = y[it][u];
        }
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
