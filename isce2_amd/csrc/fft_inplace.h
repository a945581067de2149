// In-place two-factor FFTs (n = R1*R2) over many signals held in LDS.
//
// fft_nat2dr: natural-order input, output X[k1 + R1*k2] at q = k2 + R2*k1
//   pass 1: DFT_R1 over n1 (stride R2), twiddle w_n^{n2*k1}
//   pass 2: DFT_R2 over n2 (contiguous blocks of R2)
// fft_dr2nat: input Z[k1 + R1*k2] at q = k2 + R2*k1, natural-order output
//   pass A: DFT_R2 over k2 (contiguous), twiddle w_n^{m2*k1}
//   pass B: DFT_R1 over k1 (stride R2)
// Signal s, storage index q lives at buf[s*ss + q*es]. Consecutive threads
// take consecutive signals, so rows (es=1) with an odd pitch and columns
// (ss=1) are both free of LDS bank-conflict hot spots. Each pass ends with a
// workgroup barrier. DIR is the exponent sign; no scaling.
#pragma once
#include "cx.h"
#include "fft_lds.h"

namespace ampcor {

template <class T, int DIR>
__device__ __forceinline__ void dft5(cx<T> *x) {
    const T c1 = T(0.30901699437494742410), c2 = T(-0.80901699437494742410);
    const T s1 = T(DIR) * T(0.95105651629515357212), s2 = T(DIR) * T(0.58778525229247312917);
    const cx<T> a = x[1] + x[4], b = x[1] - x[4], c = x[2] + x[3], d = x[2] - x[3];
    const cx<T> m1 = {x[0].x + c1 * a.x + c2 * c.x, x[0].y + c1 * a.y + c2 * c.y};
    const cx<T> m2 = {x[0].x + c2 * a.x + c1 * c.x, x[0].y + c2 * a.y + c1 * c.y};
    // i * (s1 b + s2 d) and i * (s2 b - s1 d)
    const cx<T> r1 = {-(s1 * b.y + s2 * d.y), s1 * b.x + s2 * d.x};
    const cx<T> r2 = {-(s2 * b.y - s1 * d.y), s2 * b.x - s1 * d.x};
    x[0] = x[0] + a + c;
    x[1] = m1 + r1;
    x[4] = m1 - r1;
    x[2] = m2 + r2;
    x[3] = m2 - r2;
}

template <class T, int DIR>
__device__ __forceinline__ void dft9(cx<T> *x) {
    // 3 x 3: n = n2 + 3 n1, k = k1 + 3 k2
    const T c1 = T(0.76604444311897803520), s1 = T(0.64278760968653932632);
    const T c2 = T(0.17364817766693034885), s2 = T(0.98480775301220805936);
    const T c4 = T(-0.93969262078590838405), s4 = T(0.34202014332566873304);
    cx<T> y[9];
#pragma unroll
    for (int n2 = 0; n2 < 3; ++n2) {
        cx<T> v[3] = {x[n2], x[n2 + 3], x[n2 + 6]};
        dft3<T, DIR>(v);
        y[n2 * 3 + 0] = v[0]; y[n2 * 3 + 1] = v[1]; y[n2 * 3 + 2] = v[2];   // y[n2][k1]
    }
    // twiddles w9^{n2 k1}: (n2,k1) = (1,1):1 (1,2):2 (2,1):2 (2,2):4
    y[4] = cmul(y[4], cx<T>{c1, T(DIR) * s1});
    y[5] = cmul(y[5], cx<T>{c2, T(DIR) * s2});
    y[7] = cmul(y[7], cx<T>{c2, T(DIR) * s2});
    y[8] = cmul(y[8], cx<T>{c4, T(DIR) * s4});
#pragma unroll
    for (int k1 = 0; k1 < 3; ++k1) {
        cx<T> v[3] = {y[k1], y[3 + k1], y[6 + k1]};
        dft3<T, DIR>(v);
        x[k1] = v[0]; x[k1 + 3] = v[1]; x[k1 + 6] = v[2];
    }
}

// cos and sin of 2 pi m / R for m in [0, R), evaluated at compile time (Taylor
// series after reduction to [-pi, pi]; 30 terms reach double precision there)
namespace trig_detail {
constexpr double kPi = 3.14159265358979323846;
constexpr double reduce(double x) {
    while (x > kPi) x -= 2 * kPi;
    while (x < -kPi) x += 2 * kPi;
    return x;
}
constexpr double csin(double x) {
    x = reduce(x);
    double term = x, sum = x;
    for (int n = 1; n < 30; ++n) {
        term *= -x * x / double((2 * n) * (2 * n + 1));
        sum += term;
    }
    return sum;
}
constexpr double ccos(double x) {
    x = reduce(x);
    double term = 1.0, sum = 1.0;
    for (int n = 1; n < 30; ++n) {
        term *= -x * x / double((2 * n - 1) * (2 * n));
        sum += term;
    }
    return sum;
}
}  // namespace trig_detail

template <int R>
struct RootTable {
    double c[R], s[R];
    constexpr RootTable() : c(), s() {
        for (int m = 0; m < R; ++m) {
            c[m] = trig_detail::ccos(2 * trig_detail::kPi * m / R);
            s[m] = trig_detail::csin(2 * trig_detail::kPi * m / R);
        }
    }
};

// odd-length DFT by the conjugate symmetry of its kernel (any odd R): with
// a_r = x_r + x_{R-r}, b_r = x_r - x_{R-r} (r = 1..M, M = (R-1)/2),
//   X_q, X_{R-q} = x_0 + sum_r a_r cos(2 pi rq/R) +- DIR i sum_r b_r sin(2 pi rq/R)
// (M^2 real multiply-adds per component instead of an FFT's R log R: cheap for
// the radices 11..33 used here)
template <class T, int DIR, int R>
__device__ __forceinline__ void dft_oddc(cx<T> *x) {
    constexpr int M = (R - 1) / 2;
    constexpr RootTable<R> tg{};
    cx<T> a[M + 1], b[M + 1];
    const cx<T> x0 = x[0];
    cx<T> sum = x0;
#pragma unroll
    for (int r = 1; r <= M; ++r) {
        a[r] = x[r] + x[R - r];
        b[r] = x[r] - x[R - r];
        sum = sum + a[r];
    }
    x[0] = sum;
#pragma unroll
    for (int q = 1; q <= M; ++q) {
        cx<T> A = x0, B = {T(0), T(0)};
#pragma unroll
        for (int r = 1; r <= M; ++r) {
            const int m = (r * q) % R;
            const T c = T(tg.c[m]), sn = T(tg.s[m]);
            A.x = fma(a[r].x, c, A.x);
            A.y = fma(a[r].y, c, A.y);
            B.x = fma(b[r].x, sn, B.x);
            B.y = fma(b[r].y, sn, B.y);
        }
        const cx<T> iB = {-T(DIR) * B.y, T(DIR) * B.x};
        x[q] = A + iB;
        x[R - q] = A - iB;
    }
}

template <class T, int DIR, int R>
__device__ __forceinline__ void dft_r(cx<T> *x);

// smallest odd factor >= 3 of an odd R (R itself when prime)
template <int R>
constexpr int odd_factor() {
    for (int a = 3; a * a <= R; a += 2)
        if (R % a == 0) return a;
    return R;
}

// DFT of composite length R = A * B in registers: n = B n1 + n2, k = k1 + A k2;
// A-point DFTs over stride B, twiddles w_R^{n2 k1}, B-point DFTs, natural order out
template <class T, int DIR, int A, int B>
__device__ __forceinline__ void dft_nested(cx<T> *x) {
    constexpr int R = A * B;
    constexpr RootTable<R> tg{};
    cx<T> y[R];
#pragma unroll
    for (int n2 = 0; n2 < B; ++n2) {
        cx<T> v[A];
#pragma unroll
        for (int n1 = 0; n1 < A; ++n1) v[n1] = x[B * n1 + n2];
        dft_r<T, DIR, A>(v);
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) {
            const int m = (n2 * k1) % R;
            y[n2 * A + k1] = m == 0 ? v[k1] : cmul(v[k1], cx<T>{T(tg.c[m]), T(DIR) * T(tg.s[m])});
        }
    }
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) {
        cx<T> v[B];
#pragma unroll
        for (int n2 = 0; n2 < B; ++n2) v[n2] = y[n2 * A + k1];
        dft_r<T, DIR, B>(v);
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) x[k1 + A * k2] = v[k2];
    }
}

template <class T, int DIR, int R>
__device__ __forceinline__ void dft_r(cx<T> *x) {
    if constexpr (R == 2) dft2<T, DIR>(x[0], x[1]);
    else if constexpr (R == 3) dft3<T, DIR>(x);
    else if constexpr (R == 4) dft4<T, DIR>(x[0], x[1], x[2], x[3]);
    else if constexpr (R == 5) dft5<T, DIR>(x);
    else if constexpr (R == 8) dft8<T, DIR>(x);
    else if constexpr (R == 9) dft9<T, DIR>(x);
    else if constexpr (R == 16) dft16<T, DIR>(x);
    else if constexpr (R % 16 == 0 && R > 16) dft_nested<T, DIR, 16, R / 16>(x);
    else if constexpr (R % 8 == 0 && R > 16) dft_nested<T, DIR, 8, R / 8>(x);
    else if constexpr (R % 2 == 1 && odd_factor<R>() == R) dft_oddc<T, DIR, R>(x);
    else if constexpr (R % 2 == 1) dft_nested<T, DIR, odd_factor<R>(), R / odd_factor<R>()>(x);
    else static_assert(R == 2, "unsupported radix");
}

// digit-reversed storage index of bin k
template <int R1, int R2>
__host__ __device__ __forceinline__ int dr_pos(int k) { return k / R1 + R2 * (k % R1); }

// post (optional): bin k is multiplied by post[k] as it is stored
template <class T, int DIR, int R1, int R2, int NTH>
__device__ __forceinline__ void fft_nat2dr(cx<T> *buf, int nsig, int ss, int es, const cx<T> *tw,
                                           const cx<T> *post = nullptr) {
    constexpr int n = R1 * R2;
    #pragma unroll 1
    for (int t = threadIdx.x; t < nsig * R2; t += NTH) {
        const int s = t % nsig, n2 = t / nsig;
        cx<T> *b = buf + s * ss + n2 * es;
        cx<T> x[R1];
#pragma unroll
        for (int n1 = 0; n1 < R1; ++n1) x[n1] = b[(R2 * n1) * es];
        dft_r<T, DIR, R1>(x);
        // no branch on n2 == 0 (tw[0] is exactly 1): the R1 - 1 table loads go out together
#pragma unroll
        for (int k1 = 1; k1 < R1; ++k1) x[k1] = cmul(x[k1], twid<T, DIR>(tw, n2 * k1));
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) b[(R2 * k1) * es] = x[k1];
    }
    __syncthreads();
    #pragma unroll 1
    for (int t = threadIdx.x; t < nsig * R1; t += NTH) {
        const int s = t % nsig, k1 = t / nsig;
        cx<T> *b = buf + s * ss + (R2 * k1) * es;
        cx<T> x[R2];
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2) x[n2] = b[n2 * es];
        dft_r<T, DIR, R2>(x);
        if (post) {
#pragma unroll
            for (int k2 = 0; k2 < R2; ++k2) x[k2] = cmul(x[k2], post[k1 + R1 * k2]);
        }
#pragma unroll
        for (int k2 = 0; k2 < R2; ++k2) b[k2 * es] = x[k2];
    }
    __syncthreads();
    (void)n;
}

template <class T, int DIR, int R1, int R2, int NTH>
__device__ __forceinline__ void fft_dr2nat(cx<T> *buf, int nsig, int ss, int es, const cx<T> *tw) {
    #pragma unroll 1
    for (int t = threadIdx.x; t < nsig * R1; t += NTH) {
        const int s = t % nsig, k1 = t / nsig;
        cx<T> *b = buf + s * ss + (R2 * k1) * es;
        cx<T> x[R2];
#pragma unroll
        for (int k2 = 0; k2 < R2; ++k2) x[k2] = b[k2 * es];
        dft_r<T, DIR, R2>(x);
#pragma unroll
        for (int m2 = 1; m2 < R2; ++m2) x[m2] = cmul(x[m2], twid<T, DIR>(tw, m2 * k1));
#pragma unroll
        for (int m2 = 0; m2 < R2; ++m2) b[m2 * es] = x[m2];
    }
    __syncthreads();
    #pragma unroll 1
    for (int t = threadIdx.x; t < nsig * R2; t += NTH) {
        const int s = t % nsig, m2 = t / nsig;
        cx<T> *b = buf + s * ss + m2 * es;
        cx<T> x[R1];
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) x[k1] = b[(R2 * k1) * es];
        dft_r<T, DIR, R1>(x);
#pragma unroll
        for (int m1 = 0; m1 < R1; ++m1) b[(R2 * m1) * es] = x[m1];
    }
    __syncthreads();
}

}  // namespace ampcor
