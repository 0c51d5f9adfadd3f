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

// ---- compile-time twiddles ------------------------------------------------
// cos / sin of 2 pi k / P evaluated at compile time: the angle is reduced to
// an octant of |x| <= pi/4 and summed as a Taylor series to full double
// precision.
namespace ce {
constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double taylor_sin(double x) {
    double t = x, s = x;
    for (int k = 1; k < 14; ++k) { t *= -x * x / ((2 * k) * (2 * k + 1)); s += t; }
    return s;
}
constexpr double taylor_cos(double x) {
    double t = 1.0, s = 1.0;
    for (int k = 1; k < 14; ++k) { t *= -x * x / ((2 * k - 1) * (2 * k)); s += t; }
    return s;
}
// cos(2 pi k / P) and sin(2 pi k / P) via the octant of 8k/P
constexpr double cos2pi(int k, int P) {
    k = ((k % P) + P) % P;
    // angle = 2 pi k / P = (pi/4) * (8k/P); octant o = floor(8k/P), rem r
    int o = (8 * k) / P;
    double r = (kPi / 4) * ((double)(8 * k - o * P) / P);   // in [0, pi/4)
    switch (o) {
        case 0: return taylor_cos(r);
        case 1: return taylor_sin(kPi / 4 - r);
        case 2: return -taylor_sin(r);
        case 3: return -taylor_cos(kPi / 4 - r);
        case 4: return -taylor_cos(r);
        case 5: return -taylor_sin(kPi / 4 - r);
        case 6: return taylor_sin(r);
        default: return taylor_cos(kPi / 4 - r);
    }
}
constexpr double sin2pi(int k, int P) { return cos2pi(k * 4 - P, 4 * P); }
}  // namespace ce

// y_q <- sum_u y_u exp(S 2 pi i u q / P) for any P used by the plans.
template <int P, int S, typename T>
__device__ __forceinline__ void dft(cx_t<T>* y);

// Radix 7 by the symmetric pairs (y_u +- y_{7-u}), like dft5.
template <int S, typename T>
__device__ __forceinline__ void dft7(cx_t<T>* y) {
    using C = cx_t<T>;
    const T c1 = T(0.62348980185873353052500488400424);    // cos(2 pi / 7)
    const T c2 = T(-0.22252093395631440428890256449679);   // cos(4 pi / 7)
    const T c3 = T(-0.90096886790241912623610231950745);   // cos(6 pi / 7)
    const T s1 = T(0.78183148246802980870844452667406);    // sin(2 pi / 7)
    const T s2 = T(0.97492791218182360701813168299393);    // sin(4 pi / 7)
    const T s3 = T(0.43388373911755812047576833284836);    // sin(6 pi / 7)
    C t1 = cadd(y[1], y[6]), d1 = csub(y[1], y[6]);
    C t2 = cadd(y[2], y[5]), d2 = csub(y[2], y[5]);
    C t3 = cadd(y[3], y[4]), d3 = csub(y[3], y[4]);
    C m1 = C{y[0].x + c1 * t1.x + c2 * t2.x + c3 * t3.x, y[0].y + c1 * t1.y + c2 * t2.y + c3 * t3.y};
    C m2 = C{y[0].x + c2 * t1.x + c3 * t2.x + c1 * t3.x, y[0].y + c2 * t1.y + c3 * t2.y + c1 * t3.y};
    C m3 = C{y[0].x + c3 * t1.x + c1 * t2.x + c2 * t3.x, y[0].y + c3 * t1.y + c1 * t2.y + c2 * t3.y};
    // sin(2 pi u q / 7) for q = 1, 2, 3 over u = 1, 2, 3
    C r1 = cmul_si<S>(C{s1 * d1.x + s2 * d2.x + s3 * d3.x, s1 * d1.y + s2 * d2.y + s3 * d3.y});
    C r2 = cmul_si<S>(C{s2 * d1.x - s3 * d2.x - s1 * d3.x, s2 * d1.y - s3 * d2.y - s1 * d3.y});
    C r3 = cmul_si<S>(C{s3 * d1.x - s1 * d2.x + s2 * d3.x, s3 * d1.y - s1 * d2.y + s2 * d3.y});
    y[0] = cadd(y[0], cadd(t1, cadd(t2, t3)));
    y[1] = cadd(m1, r1);
    y[6] = csub(m1, r1);
    y[2] = cadd(m2, r2);
    y[5] = csub(m2, r2);
    y[3] = cadd(m3, r3);
    y[4] = csub(m3, r3);
}

// Any odd prime P (11, 13 in the plans) by the symmetric pairs
// t_u = y_u + y_{P-u}, d_u = y_u - y_{P-u}, u = 1 .. (P-1)/2:
//   y_q, y_{P-q} = y_0 + sum_u cos(2 pi u q / P) t_u
//                  +- S i sum_u sin(2 pi u q / P) d_u.
// The recursion over (q, u) makes every cosine and sine a compile-time
// constant.
template <int P, int S, typename T, int Q, int U>
struct OddPairs {
    static constexpr int H = (P - 1) / 2;
    static __device__ __forceinline__ void sums(cx_t<T>& m, cx_t<T>& r, const cx_t<T>* t, const cx_t<T>* d) {
        constexpr T cq = (T)ce::cos2pi(U * Q, P);
        constexpr T sq = (T)ce::sin2pi(U * Q, P);
        m.x += cq * t[U - 1].x;
        m.y += cq * t[U - 1].y;
        r.x += sq * d[U - 1].x;
        r.y += sq * d[U - 1].y;
        if constexpr (U < H) OddPairs<P, S, T, Q, U + 1>::sums(m, r, t, d);
    }
    static __device__ __forceinline__ void outputs(cx_t<T>* y, const cx_t<T>* t, const cx_t<T>* d) {
        using C = cx_t<T>;
        C m = y[0], r = C{0, 0};
        OddPairs<P, S, T, Q, 1>::sums(m, r, t, d);
        r = cmul_si<S>(r);
        y[Q] = cadd(m, r);
        y[P - Q] = csub(m, r);
        if constexpr (Q < H) OddPairs<P, S, T, Q + 1, 1>::outputs(y, t, d);
    }
};

template <int P, int S, typename T>
__device__ __forceinline__ void dft_odd(cx_t<T>* y) {
    using C = cx_t<T>;
    constexpr int H = (P - 1) / 2;
    C t[H], d[H];
#pragma unroll
    for (int u = 1; u <= H; ++u) {
        t[u - 1] = cadd(y[u], y[P - u]);
        d[u - 1] = csub(y[u], y[P - u]);
    }
    C sum = y[0];
#pragma unroll
    for (int u = 0; u < H; ++u) sum = cadd(sum, t[u]);
    OddPairs<P, S, T, 1, 1>::outputs(y, t, d);      // reads y[0] before it is replaced
    y[0] = sum;
}

// Composite P = P1 * P2 by one Cooley-Tukey split in registers:
// u = P2 u1 + u2, q = q1 + P1 q2:
//   Z[u2][q1] = DFT_P1 over u1 of y[P2 u1 + u2], times w^(u2 q1),
//   X[q1 + P1 q2] = DFT_P2 over u2 of Z[u2][q1].
template <int P1, int P2, int S, typename T>
__device__ __forceinline__ void dft_split(cx_t<T>* y) {
    using C = cx_t<T>;
    constexpr int P = P1 * P2;
    C z[P];
#pragma unroll
    for (int u2 = 0; u2 < P2; ++u2) {
        C a[P1];
#pragma unroll
        for (int u1 = 0; u1 < P1; ++u1) a[u1] = y[P2 * u1 + u2];
        dft<P1, S, T>(a);
#pragma unroll
        for (int q1 = 0; q1 < P1; ++q1) {
            if (u2 * q1 != 0) {
                const C w = C{(T)ce::cos2pi(u2 * q1, P), (T)(S * ce::sin2pi(u2 * q1, P))};
                a[q1] = cmul(a[q1], w);
            }
            z[u2 * P1 + q1] = a[q1];
        }
    }
#pragma unroll
    for (int q1 = 0; q1 < P1; ++q1) {
        C b[P2];
#pragma unroll
        for (int u2 = 0; u2 < P2; ++u2) b[u2] = z[u2 * P1 + q1];
        dft<P2, S, T>(b);
#pragma unroll
        for (int q2 = 0; q2 < P2; ++q2) y[q1 + P1 * q2] = b[q2];
    }
}

// Inverse of a modulo m (a, m coprime), at compile time.
constexpr int inv_mod(int a, int m) {
    for (int x = 1; x < m; ++x)
        if ((a * x) % m == 1) return x;
    return 1;
}
constexpr int gcd_c(int a, int b) { return b == 0 ? a : gcd_c(b, a % b); }

// Composite P = P1 * P2 with gcd(P1, P2) = 1 by the prime-factor (Good-
// Thomas) mapping, which needs no twiddle factors: input u = (P2 u1 + P1 u2)
// mod P, output q = (P2 a q1 + P1 b q2) mod P with a = P2^-1 mod P1,
// b = P1^-1 mod P2 (q = q1 mod P1 = q2 mod P2), since then
// u q = P2 u1 q1 + P1 u2 q2 (mod P) and the DFT separates into DFT_P1 over
// u1 and DFT_P2 over u2.
template <int P1, int P2, int S, typename T>
__device__ __forceinline__ void dft_pfa(cx_t<T>* y) {
    using C = cx_t<T>;
    constexpr int P = P1 * P2;
    constexpr int A = inv_mod(P2 % P1, P1), B = inv_mod(P1 % P2, P2);
    C z[P];
#pragma unroll
    for (int u2 = 0; u2 < P2; ++u2) {
        C a[P1];
#pragma unroll
        for (int u1 = 0; u1 < P1; ++u1) a[u1] = y[(P2 * u1 + P1 * u2) % P];
        dft<P1, S, T>(a);
#pragma unroll
        for (int q1 = 0; q1 < P1; ++q1) z[u2 * P1 + q1] = a[q1];
    }
#pragma unroll
    for (int q1 = 0; q1 < P1; ++q1) {
        C b[P2];
#pragma unroll
        for (int u2 = 0; u2 < P2; ++u2) b[u2] = z[u2 * P1 + q1];
        dft<P2, S, T>(b);
#pragma unroll
        for (int q2 = 0; q2 < P2; ++q2) y[(P2 * A * q1 + P1 * B * q2) % P] = b[q2];
    }
}

template <int P, int S, typename T>
__device__ __forceinline__ void dft(cx_t<T>* y) {
    if constexpr (P == 1) {
    } else if constexpr (P == 2) dft2<S, T>(y);
    else if constexpr (P == 3) dft3<S, T>(y);
    else if constexpr (P == 4) dft4<S, T>(y);
    else if constexpr (P == 5) dft5<S, T>(y);
    else if constexpr (P == 6) dft_pfa<2, 3, S, T>(y);
    else if constexpr (P == 8) dft_split<2, 4, S, T>(y);
    else if constexpr (P == 9) dft_split<3, 3, S, T>(y);
    else if constexpr (P == 10) dft_pfa<2, 5, S, T>(y);
    else if constexpr (P == 12) dft_pfa<3, 4, S, T>(y);
    else if constexpr (P == 15) dft_pfa<3, 5, S, T>(y);
    else if constexpr (P == 16) dft_split<4, 4, S, T>(y);
    else if constexpr (P == 7) dft7<S, T>(y);
    else if constexpr (P == 11 || P == 13) dft_odd<P, S, T>(y);
}

template <int P, int S, bool DIT, typename T>
__device__ __forceinline__ void butterfly(cx_t<T>* buf, int base, int es, int k1, int tstep,
                                          const cx_t<T>* tw) {
    using C = cx_t<T>;
    C y[P];
#pragma unroll
    for (int u = 0; u < P; ++u) y[u] = buf[base + u * es];
    if (DIT && k1 != 0) {
#pragma unroll
        for (int u = 1; u < P; ++u) y[u] = cmul(y[u], twid<S>(tw, u * k1 * tstep));
    }
    dft<P, S, T>(y);
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
    // Strided axes (es > 1, lines adjacent) put consecutive lanes on
    // consecutive lines so a wave's 16-byte LDS accesses are contiguous;
    // the contiguous axis keeps butterflies of one line on adjacent lanes.
    const bool line_fast = es != 1;
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
        int line, r;
        if (line_fast) {
            r = t / nlines;
            line = t - r * nlines;
        } else {
            line = t / nbf;
            r = t - line * nbf;
        }
        int b = r / L;
        int k1 = r - b * L;
        int base = line * ls + (b * Lt + k1) * es;
        int ues = L * es;
        switch (p) {
#define EFFM_RADIX(P) case P: butterfly<P, S, DIT, T>(buf, base, ues, k1, tstep, tw); break;
            EFFM_RADIX(2) EFFM_RADIX(3) EFFM_RADIX(4) EFFM_RADIX(5) EFFM_RADIX(6)
            EFFM_RADIX(7) EFFM_RADIX(8) EFFM_RADIX(9) EFFM_RADIX(10) EFFM_RADIX(11)
            EFFM_RADIX(12) EFFM_RADIX(13)
#undef EFFM_RADIX
            default: break;
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
