// Fused per-window kernels: one workgroup per window keeps each stage's
// intermediates in LDS/registers instead of HBM.
//
//   k_coarse     chip gather + |z| + template mean + FP64 window sums + raw
//                correlation + normalization + argmax + covariance + SNR +
//                zoom start/shift            (cuAmpcorChunk.cpp:18-106)
//   k_oversample chip gather + deramp + x2 band-limited interpolation + |z|
//                (+ FP64 mean)               (cuAmpcorChunk.cpp:115-137)
//   k_fine       FP64 window sums + oversampled correlation + normalization +
//                trim + 16x surface interpolation + argmax + sub-pixel offset
//                                            (cuAmpcorChunk.cpp:148-207)
//
// Correlations use row FFTs of length L >= search width plus a direct sum
// over template rows per frequency bin:
//   c[u,v] = sum_i (1/L) IFFT_L( conj(T_i) S_{i+u} )[v]
// which equals the reference's padded 2-D FFT correlation (cuCorrFrequency.cu:63-79)
// for every valid lag, with only S x L/2 complex values resident in LDS.
#pragma once
#include <cfloat>
#include "cx.h"
#include "fft_lds.h"

namespace ampcor {
namespace fused {

constexpr int NT = 256;   // threads per workgroup

using cf = cx<float>;

__device__ __forceinline__ float pixel_amp(const void *img, int mode, int64_t off) {
    if (mode == 2) return fabsf(static_cast<const float *>(img)[off]);
    const float2 z = static_cast<const float2 *>(img)[off];
    return sqrtf(z.x * z.x + z.y * z.y);
}

__device__ __forceinline__ cf pixel_cx(const void *img, int mode, int64_t off) {
    if (mode == 2) return {static_cast<const float *>(img)[off], 0.f};
    const float2 z = static_cast<const float2 *>(img)[off];
    if (mode == 1) return {sqrtf(z.x * z.x + z.y * z.y), 0.f};
    return {z.x, z.y};
}

__device__ __forceinline__ double block_sum(double v, double *sh) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) s += sh[i];
    __syncthreads();
    return s;
}

// ---------------------------------------------------------------------------
// FP64 window sums -> normalization factors nrm[u][v] = 1/sqrt((s2 - s1^2/n) t2 + FLT_EPSILON)
// (cuCorrNormalization.cu:402-423). Waves own contiguous lag rows; lanes own
// columns (CPL per lane); vertical sums slide down, horizontal sums come from
// wave prefix scans. No barriers inside.
// ---------------------------------------------------------------------------
template <int CPL>
__device__ void window_norms(const float *amp, int pitch, int SW, int TH, int TW, int NLh, int NLw,
                             double t2, float *nrm) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int per = (NLh + NT / 64 - 1) / (NT / 64);
    const int u_beg = wave * per, u_end = min(NLh, u_beg + per);
    if (u_beg >= u_end) return;
    double g1[CPL], g2[CPL];
#pragma unroll
    for (int m = 0; m < CPL; ++m) {
        g1[m] = 0.0; g2[m] = 0.0;
        const int c = lane + 64 * m;
        if (c < SW)
            for (int i = 0; i < TH; ++i) {
                const double v = amp[(u_beg + i) * pitch + c];
                g1[m] += v;
                g2[m] += v * v;
            }
    }
    const double inv_n = 1.0 / double(TH * TW);
    for (int u = u_beg; u < u_end; ++u) {
        double p1[CPL], p2[CPL];
        double c1 = 0.0, c2 = 0.0;
#pragma unroll
        for (int m = 0; m < CPL; ++m) {
            double x1 = g1[m], x2 = g2[m];
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const double y1 = __shfl_up(x1, d), y2 = __shfl_up(x2, d);
                if (lane >= d) { x1 += y1; x2 += y2; }
            }
            x1 += c1; x2 += c2;
            p1[m] = x1; p2[m] = x2;
            c1 = __shfl(x1, 63);
            c2 = __shfl(x2, 63);
        }
        // window sums for v = lane + 64*k: P(v+TW-1) - P(v-1)
        for (int k = 0; k * 64 < NLw; ++k) {
            const int v = lane + 64 * k;
            const int hi = min(v + TW - 1, SW - 1), lo = v - 1;
            double ph1 = 0, ph2 = 0, pl1 = 0, pl2 = 0;
#pragma unroll
            for (int m = 0; m < CPL; ++m) {
                const double a1 = __shfl(p1[m], hi & 63), a2 = __shfl(p2[m], hi & 63);
                const double b1 = __shfl(p1[m], lo & 63), b2 = __shfl(p2[m], lo & 63);
                if ((hi >> 6) == m) { ph1 = a1; ph2 = a2; }
                if (lo >= 0 && (lo >> 6) == m) { pl1 = b1; pl2 = b2; }
            }
            if (v < NLw) {
                const double s1 = ph1 - pl1, s2 = ph2 - pl2;
                const double norm2 = (s2 - s1 * s1 * inv_n) * t2;
                nrm[u * NLw + v] = float(1.0 / sqrt(norm2 + double(FLT_EPSILON)));
            }
        }
        if (u + 1 < u_end) {
#pragma unroll
            for (int m = 0; m < CPL; ++m) {
                const int c = lane + 64 * m;
                if (c < SW) {
                    const double vn = amp[(u + TH) * pitch + c], vo = amp[u * pitch + c];
                    g1[m] += vn - vo;
                    g2[m] += vn * vn - vo * vo;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// real row FFTs: up to 8 real rows (two per complex transform) -> packed half
// spectra rows of L/2 complex, slot 0 = (DC, Nyquist). rowval(r, x) gives the
// real input (x < L). work holds 2*4*L complex.
// ---------------------------------------------------------------------------
template <int L, class RowFn>
__device__ void rfft_rows8(RowFn rowval, int nr, cf *work, const FftDesc &d, cf *dst) {
    constexpr int H = L / 2;
    for (int i = threadIdx.x; i < 4 * L; i += NT) {
        const int j = i / L, x = i - j * L;
        const float a = (2 * j < nr) ? rowval(2 * j, x) : 0.f;
        const float b = (2 * j + 1 < nr) ? rowval(2 * j + 1, x) : 0.f;
        work[i] = {a, b};
    }
    __syncthreads();
    const cf *Z = fft_lds<float, -1>(work, work + 4 * L, d, 4, 1, L);
    for (int i = threadIdx.x; i < 4 * H; i += NT) {
        const int j = i / H, k = i - j * H;
        const cf *z = Z + j * L;
        cf A, B;
        if (k == 0) {
            A = {z[0].x, z[H].x};
            B = {z[0].y, z[H].y};
        } else {
            const cf zk = z[k], zc = conj(z[L - k]);
            A = (zk + zc) * 0.5f;
            const cf D = (zk - zc) * 0.5f;
            B = {D.y, -D.x};
        }
        if (2 * j < nr) dst[(2 * j) * H + k] = A;
        if (2 * j + 1 < nr) dst[(2 * j + 1) * H + k] = B;
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Correlation of template rows (staged in batches of 8 by fill(i0, nr), then read
// by trow(i, x) from any thread) with
// the search spectra resident in `sp` ([SH][L/2] packed). Lags u < NLh,
// v < NLw. Thread (f, b): bin slot f = tid % H, lag block b = tid / H holding
// K consecutive lags. Result (unnormalized) -> corr[NLh][NLw]. `sp` is
// reused as scratch for the inverse transforms.
// ---------------------------------------------------------------------------
template <int L, int K, class FillFn, class TRowFn>
__device__ void correlate_rows(FillFn fill, TRowFn trow, int TH, int SH, int NLh, int NLw, cf *sp, cf *work,
                               cf *tb, const FftDesc &d, float *corr) {
    constexpr int H = L / 2;
    const int f = threadIdx.x % H;
    const int blk = threadIdx.x / H;
    const int u0 = blk * K;
    const bool active = (blk < NT / H) && (u0 < NLh);
    cf q[K];
#pragma unroll
    for (int k = 0; k < K; ++k) q[k] = {0.f, 0.f};

    for (int i0 = 0; i0 < TH; i0 += 8) {
        const int nr = min(8, TH - i0);
        fill(i0, nr);          // stage template rows i0..i0+nr-1 where trow can read them
        __syncthreads();
        rfft_rows8<L>([&](int r, int x) { return trow(i0 + r, x); }, nr, work, d, tb);
        if (active) {
            cf sw[K + 7];
#pragma unroll
            for (int m = 0; m < K + 7; ++m) {
                const int row = i0 + u0 + m;
                sw[m] = (row < SH) ? sp[row * H + f] : cf{0.f, 0.f};
            }
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                if (i < nr) {
                    const cf tv = tb[i * H + f];
                    if (f == 0) {
#pragma unroll
                        for (int k = 0; k < K; ++k) {
                            q[k].x += tv.x * sw[i + k].x;
                            q[k].y += tv.y * sw[i + k].y;
                        }
                    } else {
#pragma unroll
                        for (int k = 0; k < K; ++k) {
                            const cf s = sw[i + k];
                            q[k].x += tv.x * s.x + tv.y * s.y;
                            q[k].y += tv.x * s.y - tv.y * s.x;
                        }
                    }
                }
            }
        }
        __syncthreads();   // tb is rewritten by the next batch
    }
    // spectra no longer needed: store Q[u][f] in sp
    if (active) {
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (u0 + k < NLh) sp[(u0 + k) * H + f] = q[k];
    }
    __syncthreads();
    // inverse: pack lag pairs (u, u+1) into one complex transform, 4 at a time
    const int npair = (NLh + 1) / 2;
    const float invL = 1.0f / float(L);
    for (int p0 = 0; p0 < npair; p0 += 4) {
        const int np = min(4, npair - p0);
        for (int i = threadIdx.x; i < 4 * L; i += NT) {
            const int j = i / L, k = i - j * L;
            const int ua = 2 * (p0 + j), ub = ua + 1;
            cf y = {0.f, 0.f};
            if (j < np) {
                const cf *qa = sp + ua * H;
                const cf *qb = (ub < NLh) ? sp + ub * H : nullptr;
                if (k == 0) y = {qa[0].x, qb ? qb[0].x : 0.f};
                else if (k == H) y = {qa[0].y, qb ? qb[0].y : 0.f};
                else if (k < H) {
                    const cf a = qa[k], b = qb ? qb[k] : cf{0.f, 0.f};
                    y = {a.x - b.y, a.y + b.x};
                } else {
                    const cf a = qa[L - k], b = qb ? qb[L - k] : cf{0.f, 0.f};
                    y = {a.x + b.y, b.x - a.y};
                }
            }
            work[i] = y;
        }
        __syncthreads();
        const cf *Y = fft_lds<float, +1>(work, work + 4 * L, d, 4, 1, L);
        for (int i = threadIdx.x; i < 4 * NLw; i += NT) {
            const int j = i / NLw, v = i - j * NLw;
            const int ua = 2 * (p0 + j);
            if (j < np) {
                const cf y = Y[j * L + v];
                corr[ua * NLw + v] = y.x * invL;
                if (ua + 1 < NLh) corr[(ua + 1) * NLw + v] = y.y * invL;
            }
        }
        __syncthreads();
    }
}

// in-place (bottom-up) conversion of float amplitudes amp[SH][SW] (pitch SW)
// into packed half spectra rows [SH][L/2] complex occupying the same memory
template <int L>
__device__ void search_spectra_inplace(float *amp, int SH, int SW, cf *work, const FftDesc &d) {
    cf *sp = reinterpret_cast<cf *>(amp);
    const int nb = (SH + 7) / 8;
    for (int b = nb - 1; b >= 0; --b) {
        const int r0 = b * 8, nr = min(8, SH - r0);
        rfft_rows8<L>([&](int r, int x) { return x < SW ? amp[(r0 + r) * SW + x] : 0.f; }, nr, work, d,
                      sp + r0 * (L / 2));
    }
}

// argmax with the reference tie rule over n values in LDS; returns (idx, val) to all threads
__device__ __forceinline__ void block_argmax(const float *v, int n, int &idx, float &val, float *shv,
                                             uint64_t *shk) {
    float bv = -FLT_MAX;
    uint64_t bk = ~uint64_t(0);
    for (int i = threadIdx.x; i < n; i += NT) {
        const float x = v[i];
        const uint64_t k = maxloc_key(uint32_t(i));
        if (x > -FLT_MAX && better(x, k, bv, bk)) { bv = x; bk = k; }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const float ov = __shfl_xor(bv, o);
        const uint64_t ok = __shfl_xor(bk, o);
        if (better(ov, ok, bv, bk)) { bv = ov; bk = ok; }
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) { shv[threadIdx.x >> 6] = bv; shk[threadIdx.x >> 6] = bk; }
    __syncthreads();
    bv = shv[0]; bk = shk[0];
#pragma unroll
    for (int i = 1; i < NT / 64; ++i)
        if (better(shv[i], shk[i], bv, bk)) { bv = shv[i]; bk = shk[i]; }
    __syncthreads();
    idx = (bk == ~uint64_t(0)) ? 0 : int(bk & 0xffffffffu);
    val = bv;
}

}  // namespace fused
}  // namespace ampcor
