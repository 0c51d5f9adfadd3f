// fft_lds.hpp — in-place mixed-radix complex FFT on data resident in LDS.
//
// A length-n transform with n = p_1 p_2 ... p_s runs as s radix stages over
// all threads of the workgroup.  Stage t combines blocks of L = p_1..p_{t-1}
// into blocks of L*p_t in place, so no ping-pong buffer is needed and a whole
// 90x90 complex128 plane (129.6 KB) fits in the 160 KB LDS of a CU.
//
//   DIT: load x[j] at position perm(j), run stages t = 1..s  -> natural X
//   DIF: natural x,                 run stages t = s..1  -> X[k] at perm(k)
//
// perm(j) reverses the mixed-radix digits of j (host side: fft_perm()).
// Twiddles W_n^k = exp(-2 pi i k / n) sit in LDS; the backward transform
// uses their conjugates.  Forward = exp(-i), backward = exp(+i), both
// unnormalised here (the callers apply 1/N where the reference's R2G does).
#pragma once
#include <hip/hip_runtime.h>

namespace effm {

constexpr int kMaxStages = 12;

// One FFT axis: its length and radix sequence p_1..p_s (kernel argument).
struct FftAxis {
    int n;
    int ns;
    int p[kMaxStages];
};

template <typename T> struct Cx;
template <> struct Cx<double> { using type = double2; };
template <> struct Cx<float>  { using type = float2; };
template <typename T> using cx_t = typename Cx<T>::type;

template <typename C>
__device__ __forceinline__ C cadd(C a, C b) { return C{a.x + b.x, a.y + b.y}; }
template <typename C>
__device__ __forceinline__ C csub(C a, C b) { return C{a.x - b.x, a.y - b.y}; }
template <typename C>
__device__ __forceinline__ C cmul(C a, C b) {
    return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// multiply by conj(b)
template <typename C>
__device__ __forceinline__ C cmulc(C a, C b) {
    return C{a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
// multiply by (S * i), S = +-1
template <int S, typename C>
__device__ __forceinline__ C cmul_si(C a) {
    return S > 0 ? C{-a.y, a.x} : C{a.y, -a.x};
}

// Twiddle W^k for direction S (S=-1 forward exp(-i), S=+1 backward).
template <int S, typename C>
__device__ __forceinline__ C twid(const C* tw, int k) {
    C w = tw[k];
    if (S > 0) w.y = -w.y;
    return w;
}

// ---- small DFTs: y_q <- sum_u y_u exp(S 2 pi i u q / p) -------------------
template <int S, typename T>
__device__ __forceinline__ void dft2(cx_t<T>* y) {
    auto a = y[0], b = y[1];
    y[0] = cadd(a, b);
    y[1] = csub(a, b);
}

template <int S, typename T>
__device__ __forceinline__ void dft3(cx_t<T>* y) {
    using C = cx_t<T>;
    const T c1 = T(-0.5);
    const T s1 = T(0.86602540378443864676372317075294);   // sin(2 pi / 3)
    C t = cadd(y[1], y[2]);
    C d = csub(y[1], y[2]);
    C m = C{y[0].x + c1 * t.x, y[0].y + c1 * t.y};
    C r = cmul_si<S>(C{s1 * d.x, s1 * d.y});
    y[0] = cadd(y[0], t);
    y[1] = cadd(m, r);
    y[2] = csub(m, r);
}

template <int S, typename T>
__device__ __forceinline__ void dft4(cx_t<T>* y) {
    using C = cx_t<T>;
    C a = cadd(y[0], y[2]);
    C b = csub(y[0], y[2]);
    C c = cadd(y[1], y[3]);
    C d = cmul_si<S>(csub(y[1], y[3]));
    y[0] = cadd(a, c);
    y[2] = csub(a, c);
    y[1] = cadd(b, d);
    y[3] = csub(b, d);
}

template <int S, typename T>
__device__ __forceinline__ void dft5(cx_t<T>* y) {
    using C = cx_t<T>;
    const T c1 = T(0.30901699437494742410229341718282);    // cos(2 pi / 5)
    const T c2 = T(-0.80901699437494742410229341718282);   // cos(4 pi / 5)
    const T s1 = T(0.95105651629515357211643933337938);    // sin(2 pi / 5)
    const T s2 = T(0.58778525229247312916870595463907);    // sin(4 pi / 5)
    C t1 = cadd(y[1], y[4]), d1 = csub(y[1], y[4]);
    C t2 = cadd(y[2], y[3]), d2 = csub(y[2], y[3]);
    C m1 = C{y[0].x + c1 * t1.x + c2 * t2.x, y[0].y + c1 * t1.y + c2 * t2.y};
    C m2 = C{y[0].x + c2 * t1.x + c1 * t2.x, y[0].y + c2 * t1.y + c1 * t2.y};
    C r1 = cmul_si<S>(C{s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y});
    C r2 = cmul_si<S>(C{s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y});
    y[0] = cadd(y[0], cadd(t1, t2));
    y[1] = cadd(m1, r1);
    y[4] = csub(m1, r1);
    y[2] = cadd(m2, r2);
    y[3] = csub(m2, r2);
}

// Other prime radices (7, 11, 13) by direct summation with table twiddles
// W_p^{uq} = W_n^{(uq mod p) n/p}; P is a compile-time constant so the
// working set stays in registers.
template <int P, int S, typename T>
__device__ __forceinline__ void dftp(cx_t<T>* y, const cx_t<T>* tw, int n) {
    using C = cx_t<T>;
    C out[P];
    const int st = n / P;
#pragma unroll
    for (int q = 0; q < P; ++q) {
        C acc = y[0];
#pragma unroll
        for (int u = 1; u < P; ++u) acc = cadd(acc, cmul(y[u], twid<S>(tw, ((u * q) % P) * st)));
        out[q] = acc;
    }
#pragma unroll
    for (int q = 0; q < P; ++q) y[q] = out[q];
}

template <int P, int S, bool DIT, typename T>
__device__ __forceinline__ void butterfly(cx_t<T>* buf, int base, int es, int k1, int tstep,
                                          const cx_t<T>* tw, int n) {
    using C = cx_t<T>;
    C y[P];
#pragma unroll
    for (int u = 0; u < P; ++u) y[u] = buf[base + u * es];
    if (DIT && k1 != 0) {
#pragma unroll
        for (int u = 1; u < P; ++u) y[u] = cmul(y[u], twid<S>(tw, u * k1 * tstep));
    }
    if constexpr (P == 2) dft2<S, T>(y);
    else if constexpr (P == 3) dft3<S, T>(y);
    else if constexpr (P == 4) dft4<S, T>(y);
    else if constexpr (P == 5) dft5<S, T>(y);
    else dftp<P, S, T>(y, tw, n);
    if (!DIT && k1 != 0) {
#pragma unroll
        for (int u = 1; u < P; ++u) y[u] = cmul(y[u], twid<S>(tw, u * k1 * tstep));
    }
#pragma unroll
    for (int u = 0; u < P; ++u) buf[base + u * es] = y[u];
}

// One radix stage over `nlines` lines: element e of line l sits at
// l*ls + e*es.  L = product of the radices already combined (DIT) / still
// to combine below this stage (DIF).
template <int S, bool DIT, typename T>
__device__ void fft_stage(cx_t<T>* buf, int n, int p, int L, int es, int ls, int nlines,
                          const cx_t<T>* tw) {
    const int nbf = n / p;
    const int total = nlines * nbf;
    const int Lt = L * p;
    const int tstep = n / Lt;
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
        int line = t / nbf;
        int r = t - line * nbf;
        int b = r / L;
        int k1 = r - b * L;
        int base = line * ls + (b * Lt + k1) * es;
        int ues = L * es;
        switch (p) {
            case 2: butterfly<2, S, DIT, T>(buf, base, ues, k1, tstep, tw, n); break;
            case 3: butterfly<3, S, DIT, T>(buf, base, ues, k1, tstep, tw, n); break;
            case 4: butterfly<4, S, DIT, T>(buf, base, ues, k1, tstep, tw, n); break;
            case 5: butterfly<5, S, DIT, T>(buf, base, ues, k1, tstep, tw, n); break;
            case 7: butterfly<7, S, DIT, T>(buf, base, ues, k1, tstep, tw, n); break;
            case 11: butterfly<11, S, DIT, T>(buf, base, ues, k1, tstep, tw, n); break;
            default: butterfly<13, S, DIT, T>(buf, base, ues, k1, tstep, tw, n); break;
        }
    }
}

// Full transform along one axis of LDS data.  Ends with a barrier.
template <int S, bool DIT, typename T>
__device__ void fft_lds(cx_t<T>* buf, const FftAxis& ax, int es, int ls, int nlines,
                        const cx_t<T>* tw) {
    if (ax.n <= 1) return;
    if (DIT) {
        int L = 1;
        for (int t = 0; t < ax.ns; ++t) {
            fft_stage<S, true, T>(buf, ax.n, ax.p[t], L, es, ls, nlines, tw);
            __syncthreads();
            L *= ax.p[t];
        }
    } else {
        int L = ax.n;
        for (int t = ax.ns - 1; t >= 0; --t) {
            L /= ax.p[t];
            fft_stage<S, false, T>(buf, ax.n, ax.p[t], L, es, ls, nlines, tw);
            __syncthreads();
        }
    }
}

}  // namespace effm
