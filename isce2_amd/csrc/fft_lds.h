// LDS-resident mixed-radix FFTs for the engine (no library FFT).
//
// Stockham auto-sort passes over radices {16, 8, 4, 2, 3, 5, 7, 11, 13, 17};
// a length with any other prime factor runs as a direct O(n^2) DFT. Each
// thread performs whole radix-R butterflies in registers, LDS holds the
// signals between passes (ping-pong buffers). Signals are addressed as
// buf[s*ss + e*es] (signal s, element e), so rows (es=1) and interleaved
// column tiles (es=ncols, ss=1) use the same code.
//
// Sign: DIR=-1 computes X[k] = sum x[j] e^{-2 pi i jk/n}, DIR=+1 the +i sum;
// neither scales. Twiddles come from a per-length table tw[k] = e^{-2 pi i k/n}
// computed in FP64 on the host and rounded to T.
#pragma once
#include "cx.h"

namespace ampcor {

struct FftDesc {
    int n;
    int npass;        // 0 with direct=1
    int direct;       // 1: O(n^2) DFT (prime factor > 17)
    int radix[14];
    const void *tw;   // cx<T>[n] in device memory
};

template <class T, int DIR>
__device__ __forceinline__ cx<T> twid(const cx<T> *tw, int k) {
    cx<T> w = tw[k];
    return DIR < 0 ? w : conj(w);
}

// ---- fixed-radix butterflies (in registers) --------------------------------
template <class T, int DIR>
__device__ __forceinline__ void dft2(cx<T> &a, cx<T> &b) {
    cx<T> t = a;
    a = t + b;
    b = t - b;
}

template <class T, int DIR>
__device__ __forceinline__ void dft4(cx<T> &x0, cx<T> &x1, cx<T> &x2, cx<T> &x3) {
    cx<T> a0 = x0 + x2, a1 = x0 - x2, a2 = x1 + x3, a3 = mul_i<T, DIR>(x1 - x3);
    x0 = a0 + a2;
    x2 = a0 - a2;
    x1 = a1 + a3;
    x3 = a1 - a3;
}

template <class T, int DIR>
__device__ __forceinline__ void dft8(cx<T> *x) {
    const T h = T(0.70710678118654752440);
    cx<T> e0 = x[0], e1 = x[2], e2 = x[4], e3 = x[6];
    cx<T> o0 = x[1], o1 = x[3], o2 = x[5], o3 = x[7];
    dft4<T, DIR>(e0, e1, e2, e3);
    dft4<T, DIR>(o0, o1, o2, o3);
    // w8^1 = h(1 + DIR i), w8^2 = DIR i, w8^3 = h(-1 + DIR i)
    cx<T> t1 = cmul(o1, cx<T>{h, T(DIR) * h});
    cx<T> t2 = mul_i<T, DIR>(o2);
    cx<T> t3 = cmul(o3, cx<T>{-h, T(DIR) * h});
    x[0] = e0 + o0; x[4] = e0 - o0;
    x[1] = e1 + t1; x[5] = e1 - t1;
    x[2] = e2 + t2; x[6] = e2 - t2;
    x[3] = e3 + t3; x[7] = e3 - t3;
}

template <class T, int DIR>
__device__ __forceinline__ void dft16(cx<T> *x) {
    // 4 x 4: columns (stride 4) first, twiddle w16^{q*k}, then rows
    const T c1 = T(0.92387953251128675613), s1 = T(0.38268343236508977173), h = T(0.70710678118654752440);
    cx<T> y[16];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        cx<T> a = x[q], b = x[q + 4], c = x[q + 8], d = x[q + 12];
        dft4<T, DIR>(a, b, c, d);
        y[q] = a; y[q + 4] = b; y[q + 8] = c; y[q + 12] = d;   // y[q + 4k] = X_q[k]
    }
    // twiddles w16^{q*k}, w16^m = cos(2 pi m/16) + DIR i sin(2 pi m/16)
    const T cs[8] = {T(1), c1, h, s1, T(0), -s1, -h, -c1};
    const T sn[8] = {T(0), s1, h, c1, T(1), c1, h, s1};
#pragma unroll
    for (int q = 1; q < 4; ++q)
#pragma unroll
        for (int k = 1; k < 4; ++k) {
            const int m = q * k;   // < 16, here <= 9
            y[q + 4 * k] = cmul(y[q + 4 * k], cx<T>{cs[m & 7] * (m >= 8 ? T(-1) : T(1)),
                                                    T(DIR) * sn[m & 7] * (m >= 8 ? T(-1) : T(1))});
        }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        cx<T> a = y[4 * k], b = y[4 * k + 1], c = y[4 * k + 2], d = y[4 * k + 3];
        dft4<T, DIR>(a, b, c, d);
        x[k] = a; x[k + 4] = b; x[k + 8] = c; x[k + 12] = d;
    }
}

template <class T, int DIR>
__device__ __forceinline__ void dft3(cx<T> *x) {
    const T c = T(-0.5), s = T(DIR) * T(0.86602540378443864676);
    cx<T> a = x[1] + x[2], b = x[1] - x[2];
    cx<T> m = {x[0].x + c * a.x, x[0].y + c * a.y};
    cx<T> r = {-s * b.y, s * b.x};   // i*s*b
    x[0] = x[0] + a;
    x[1] = m + r;
    x[2] = m - r;
}

// odd-prime DFT using the conjugate symmetry of the kernel: with
// a_r = x_r + x_{R-r}, b_r = x_r - x_{R-r} (r = 1..M, M = (R-1)/2),
//   X_q     = x_0 + sum_r a_r cos(2 pi rq/R) + i DIR sum_r b_r sin(2 pi rq/R)
//   X_{R-q} = x_0 + sum_r a_r cos(2 pi rq/R) - i DIR sum_r b_r sin(2 pi rq/R)
// The M distinct cos/sin values are read once from the length-n table.
template <class T, int DIR, int R>
__device__ __forceinline__ void dft_odd(cx<T> *x, const cx<T> *tw, int n) {
    constexpr int M = (R - 1) / 2;
    const int st = n / R;
    T c[M + 1], s[M + 1];
#pragma unroll
    for (int k = 1; k <= M; ++k) {
        const cx<T> w = tw[k * st];           // e^{-2 pi i k/R}
        c[k] = w.x;
        s[k] = -w.y;                          // sin(2 pi k/R) > 0
    }
    cx<T> a[M + 1], b[M + 1];
    cx<T> sum = x[0];
#pragma unroll
    for (int r = 1; r <= M; ++r) {
        a[r] = x[r] + x[R - r];
        b[r] = x[r] - x[R - r];
        sum = sum + a[r];
    }
    const cx<T> x0 = x[0];
    x[0] = sum;
#pragma unroll
    for (int q = 1; q <= M; ++q) {
        cx<T> A = x0, B = {T(0), T(0)};
#pragma unroll
        for (int r = 1; r <= M; ++r) {
            const int m = (r * q) % R;
            const T cv = m <= M ? c[m] : c[R - m];
            const T sv = m <= M ? s[m] : -s[R - m];
            A.x += a[r].x * cv; A.y += a[r].y * cv;
            B.x += b[r].x * sv; B.y += b[r].y * sv;
        }
        // i DIR B
        const cx<T> iB = DIR > 0 ? cx<T>{-B.y, B.x} : cx<T>{B.y, -B.x};
        x[q] = A + iB;
        x[R - q] = A - iB;
    }
}

template <class T, int DIR, int R>
__device__ __forceinline__ void dft_radix(cx<T> *x, const cx<T> *tw, int n) {
    if constexpr (R == 2) dft2<T, DIR>(x[0], x[1]);
    else if constexpr (R == 4) dft4<T, DIR>(x[0], x[1], x[2], x[3]);
    else if constexpr (R == 8) dft8<T, DIR>(x);
    else if constexpr (R == 16) dft16<T, DIR>(x);
    else if constexpr (R == 3) dft3<T, DIR>(x);
    else dft_odd<T, DIR, R>(x, tw, n);
}

// one Stockham pass of radix R: ns = product of the radices already applied
template <class T, int DIR, int R>
__device__ void stockham_pass(const cx<T> *src, cx<T> *dst, const cx<T> *tw, int n, int ns,
                              int nsig, int es, int ss) {
    const int m = n / R;
    const int total = nsig * m;
    // strided signals (columns, es > 1) sit side by side in LDS: consecutive
    // threads take consecutive signals so a wave's accesses are contiguous;
    // contiguous signals (rows) take consecutive butterflies of one signal
    const bool across = es != 1;
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
        const int s = across ? t % nsig : t / m;
        const int j = across ? t / nsig : t - s * m;
        const int k = j % ns;
        const cx<T> *in = src + s * ss;
        cx<T> *out = dst + s * ss;
        const int step = (n / (ns * R)) * k;
        cx<T> x[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            cx<T> v = in[(j + r * m) * es];
            // no branch on k == 0 (tw[0] is exactly 1): the table loads go out together
            if (r > 0) v = cmul(v, twid<T, DIR>(tw, r * step));
            x[r] = v;
        }
        dft_radix<T, DIR, R>(x, tw, n);
        const int base = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) out[(base + r * ns) * es] = x[r];
    }
}

template <class T, int DIR>
__device__ void direct_dft(const cx<T> *src, cx<T> *dst, const cx<T> *tw, int n, int nsig, int es, int ss) {
    const int total = nsig * n;
    const bool across = es != 1;   // thread order as in stockham_pass
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
        const int s = across ? t % nsig : t / n;
        const int q = across ? t / nsig : t - s * n;
        const cx<T> *in = src + s * ss;
        cx<T> acc = {T(0), T(0)};
        int m = 0;
        for (int j = 0; j < n; ++j) {
            acc = acc + cmul(in[j * es], twid<T, DIR>(tw, m));
            m += q;
            if (m >= n) m -= n;
        }
        dst[s * ss + q * es] = acc;
    }
}

// Transform nsig signals held in `a`; `b` is scratch of the same extent.
// Returns the buffer (a or b) that holds the result. Ends with __syncthreads().
template <class T, int DIR>
__device__ cx<T> *fft_lds(cx<T> *a, cx<T> *b, const FftDesc &d, int nsig, int es, int ss) {
    const cx<T> *tw = static_cast<const cx<T> *>(d.tw);
    if (d.direct) {
        direct_dft<T, DIR>(a, b, tw, d.n, nsig, es, ss);
        __syncthreads();
        return b;
    }
    cx<T> *src = a, *dst = b;
    int ns = 1;
    for (int p = 0; p < d.npass; ++p) {
        const int R = d.radix[p];
        switch (R) {
            case 16: stockham_pass<T, DIR, 16>(src, dst, tw, d.n, ns, nsig, es, ss); break;
            case 8: stockham_pass<T, DIR, 8>(src, dst, tw, d.n, ns, nsig, es, ss); break;
            case 4: stockham_pass<T, DIR, 4>(src, dst, tw, d.n, ns, nsig, es, ss); break;
            case 2: stockham_pass<T, DIR, 2>(src, dst, tw, d.n, ns, nsig, es, ss); break;
            case 3: stockham_pass<T, DIR, 3>(src, dst, tw, d.n, ns, nsig, es, ss); break;
            case 5: stockham_pass<T, DIR, 5>(src, dst, tw, d.n, ns, nsig, es, ss); break;
            case 7: stockham_pass<T, DIR, 7>(src, dst, tw, d.n, ns, nsig, es, ss); break;
            case 11: stockham_pass<T, DIR, 11>(src, dst, tw, d.n, ns, nsig, es, ss); break;
            case 13: stockham_pass<T, DIR, 13>(src, dst, tw, d.n, ns, nsig, es, ss); break;
            default: stockham_pass<T, DIR, 17>(src, dst, tw, d.n, ns, nsig, es, ss); break;
        }
        __syncthreads();
        cx<T> *t = src; src = dst; dst = t;
        ns *= R;
    }
    return src;
}

}  // namespace ampcor
