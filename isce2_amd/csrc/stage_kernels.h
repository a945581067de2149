// Batched per-stage kernels of the staged pipeline (one launch per stage over a
// chunk of windows, intermediates in HBM). Each kernel names the reference
// function whose arithmetic it performs (paths under contrib/PyCuAmpcor/src/).
#pragma once
#include <cfloat>
#include "cx.h"
#include "fft_inplace.h"
#include "fft_lds.h"

namespace ampcor {

// ---------------------------------------------------------------------------
// 1D FFTs over rows / columns of batched row-major images in HBM
// ---------------------------------------------------------------------------
// rows: `nrows` contiguous rows of length d.n (row pitch d.n), `g` rows per block
template <class T, int DIR>
__global__ void __launch_bounds__(256) k_fft_rows(cx<T> *data, int nrows, int g, FftDesc d) {
    extern __shared__ unsigned char smem_raw[];
    cx<T> *a = reinterpret_cast<cx<T> *>(smem_raw);
    cx<T> *b = a + g * d.n;
    const int r0 = blockIdx.x * g;
    const int nr = min(g, nrows - r0);
    const int tot = nr * d.n;
    cx<T> *base = data + size_t(r0) * d.n;
    for (int i = threadIdx.x; i < tot; i += blockDim.x) a[i] = base[i];
    __syncthreads();
    cx<T> *res = fft_lds<T, DIR>(a, b, d, nr, 1, d.n);
    for (int i = threadIdx.x; i < tot; i += blockDim.x) base[i] = res[i];
}

// rows of length d.n selected per image: of the `sel` selected rows of each
// [H][d.n] image, rows s < split are row s and the others row off1 + s - split
// (the non-zero rows of a quadrant-padded spectrum, or the leading rows an
// inverse correlation keeps); `g` selected rows per block
template <class T, int DIR>
__global__ void __launch_bounds__(256) k_fft_rows_sel(cx<T> *data, int nimg, int H, int sel, int split, int off1,
                                                      int g, FftDesc d) {
    extern __shared__ unsigned char smem_raw[];
    cx<T> *a = reinterpret_cast<cx<T> *>(smem_raw);
    cx<T> *b = a + g * d.n;
    const int t0 = blockIdx.x * g;
    const int nr = min(g, nimg * sel - t0);
    auto row = [&](int k) {
        const int t = t0 + k, img = t / sel, sr = t - img * sel;
        const int r = sr < split ? sr : off1 + (sr - split);
        return data + (size_t(img) * H + r) * d.n;
    };
    for (int i = threadIdx.x; i < nr * d.n; i += blockDim.x) {
        const int k = i / d.n, c = i - k * d.n;
        a[i] = row(k)[c];
    }
    __syncthreads();
    cx<T> *res = fft_lds<T, DIR>(a, b, d, nr, 1, d.n);
    for (int i = threadIdx.x; i < nr * d.n; i += blockDim.x) {
        const int k = i / d.n, c = i - k * d.n;
        row(k)[c] = res[i];
    }
}

// columns: images [nimg][h][w] with h = d.n; `c` columns per block
template <class T, int DIR>
__global__ void __launch_bounds__(512) k_fft_cols(cx<T> *data, int w, int c, FftDesc d) {
    extern __shared__ unsigned char smem_raw[];
    cx<T> *a = reinterpret_cast<cx<T> *>(smem_raw);
    cx<T> *b = a + c * d.n;
    const int h = d.n;
    const int col0 = blockIdx.x * c;
    const int nc = min(c, w - col0);
    cx<T> *img = data + size_t(blockIdx.y) * h * w + col0;
    const int tot = h * nc;
    for (int i = threadIdx.x; i < tot; i += blockDim.x) {
        const int r = i / nc, cc = i - r * nc;
        a[r * c + cc] = img[size_t(r) * w + cc];
    }
    __syncthreads();
    cx<T> *res = fft_lds<T, DIR>(a, b, d, nc, c, 1);
    for (int i = threadIdx.x; i < tot; i += blockDim.x) {
        const int r = i / nc, cc = i - r * nc;
        img[size_t(r) * w + cc] = res[r * c + cc];
    }
}

// Inverse 2-D FFT of the quadrant padding of h x w spectra into H x W
// (cuArraysPadding.cu:59-103 followed by the oversampler's second transform,
// cuOverSampler.cpp:72-78 and :141-148), in two passes that never store the
// padding. Row pass: each of the 2*(h/2) non-zero padded rows is built in LDS
// from the small spectrum (scaled by 1/(h*w), zeros between the quadrants) and
// transformed; it lands at its padded row of `out` ([nimg][H][d.n]), whose
// all-zero rows are left unwritten. `g` rows per block.
template <class T, int DIR>
__global__ void __launch_bounds__(256) k_fft_rows_quad(const cx<T> *in, int nimg, int h, int w, cx<T> *out, int H,
                                                       int g, FftDesc d) {
    extern __shared__ unsigned char smem_raw[];
    cx<T> *a = reinterpret_cast<cx<T> *>(smem_raw);
    cx<T> *b = a + g * d.n;
    const int W = d.n, h2 = h / 2, w2 = w / 2, sel = 2 * h2;
    const T f = T(1) / T(h * w);
    const int t0 = blockIdx.x * g;
    const int nr = min(g, nimg * sel - t0);
    for (int i = threadIdx.x; i < nr * W; i += blockDim.x) {
        const int k = i / W, q = i - k * W;
        const int t = t0 + k, img = t / sel, s = t - img * sel;
        const int sr = s < h2 ? s : h - sel + s;   // rows [0, h2) and [h - h2, h) of the spectrum
        const int sc = q < w2 ? q : (q >= W - w2 ? q - W + w : -1);
        a[i] = sc >= 0 ? in[(size_t(img) * h + sr) * w + sc] * f : cx<T>{T(0), T(0)};
    }
    __syncthreads();
    cx<T> *res = fft_lds<T, DIR>(a, b, d, nr, 1, W);
    for (int i = threadIdx.x; i < nr * W; i += blockDim.x) {
        const int k = i / W, q = i - k * W;
        const int t = t0 + k, img = t / sel, s = t - img * sel;
        const int p = s < h2 ? s : H - sel + s;    // padded rows [0, h2) and [H - h2, H)
        out[(size_t(img) * H + p) * W + q] = res[i];
    }
}

// Column pass of the same inverse over [nimg][d.n][w] images whose rows
// [h2, d.n - h2) are zero and not read. Only a real image is kept: |z| (OUT 1,
// the oversampled chips' amplitude, cuArraysCopy.cu:510-529) or Re z (OUT 2,
// the oversampled surface, cuArraysCopy.cu:326-339), written to `out`
// ([nimg][d.n][w]). `c` columns per block.
template <class T, int DIR, int OUT>
__global__ void __launch_bounds__(512) k_fft_cols_quad(const cx<T> *data, int w, int c, int h2, FftDesc d, T *out) {
    extern __shared__ unsigned char smem_raw[];
    cx<T> *a = reinterpret_cast<cx<T> *>(smem_raw);
    cx<T> *b = a + c * d.n;
    const int h = d.n;
    const int col0 = blockIdx.x * c;
    const int nc = min(c, w - col0);
    const cx<T> *img = data + size_t(blockIdx.y) * h * w + col0;
    const int tot = h * nc;
    for (int i = threadIdx.x; i < tot; i += blockDim.x) {
        const int r = i / nc, cc = i - r * nc;
        a[r * c + cc] = (r < h2 || r >= h - h2) ? img[size_t(r) * w + cc] : cx<T>{T(0), T(0)};
    }
    __syncthreads();
    const cx<T> *res = fft_lds<T, DIR>(a, b, d, nc, c, 1);
    T *o = out + size_t(blockIdx.y) * h * w + col0;
    for (int i = threadIdx.x; i < tot; i += blockDim.x) {
        const int r = i / nc, cc = i - r * nc;
        const cx<T> v = res[r * c + cc];
        o[size_t(r) * w + cc] = OUT == 1 ? amag(v) : v.x;
    }
}

// ---------------------------------------------------------------------------
// chip gather (cuArraysCopy.cu:70-165): mode 0 complex copy, 1 (|z|,0), 2 real -> (x,0);
// with `amp`, also the amplitude of the gathered chip (cuArraysAbs, cuArraysCopy.cu:510-529)
// ---------------------------------------------------------------------------
template <class T>
__global__ void k_gather(const void *img, int64_t pitch, const int32_t *sd, const int32_t *sa,
                         int h, int w, int mode, cx<T> *out, T *amp) {
    const int win = blockIdx.y;
    const int tot = h * w;
    const int64_t r0 = sd[win], c0 = sa[win];
    cx<T> *o = out + size_t(win) * tot;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < tot; i += gridDim.x * blockDim.x) {
        const int r = i / w, c = i - r * w;
        const int64_t off = (r0 + r) * pitch + c0 + c;
        cx<T> v;
        if (mode == 2) {
            v = {T(static_cast<const float *>(img)[off]), T(0)};
        } else {
            const float2 z = static_cast<const float2 *>(img)[off];
            v = {T(z.x), T(z.y)};
            if (mode == 1) v = {T(fast_sqrt(z.x * z.x + z.y * z.y)), T(0)};
        }
        o[i] = v;
        if (amp) amp[size_t(win) * tot + i] = amag(v);
    }
}

// No-sync bounds check: copy the start arrays ([4][n] out), replacing every
// window whose chips leave their image by window (0, 0) and flagging it
template <class I>
__global__ void k_sanitize_starts(const I *rd, const I *ra, const I *sd, const I *sa, int n, int64_t ref_h,
                                  int64_t ref_w, int64_t sec_h, int64_t sec_w, int wh, int ww, int sh, int sw,
                                  I *out, int32_t *bad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t r0 = rd[i], a0 = ra[i], s0 = sd[i], t0 = sa[i];
    const bool ok = r0 >= 0 && a0 >= 0 && r0 + wh <= ref_h && a0 + ww <= ref_w && s0 >= 0 && t0 >= 0 &&
                    s0 + sh <= sec_h && t0 + sw <= sec_w && ref_h >= wh && ref_w >= ww && sec_h >= sh && sec_w >= sw;
    out[i] = ok ? I(r0) : I(0);
    out[n + i] = ok ? I(a0) : I(0);
    out[2 * n + i] = ok ? I(s0) : I(0);
    out[3 * n + i] = ok ? I(t0) : I(0);
    bad[i] = ok ? 0 : 1;
}

// NaN outputs for the windows k_sanitize_starts flagged
template <class T>
__global__ void k_mark_bad(const int32_t *bad, int n, T *offset, T *snr, T *cov) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !bad[i]) return;
    const T q = T(__builtin_nanf(""));
    offset[2 * i] = q; offset[2 * i + 1] = q;
    snr[i] = q;
    cov[3 * i] = q; cov[3 * i + 1] = q; cov[3 * i + 2] = q;
}

// block-wide FP64 sum (256 threads)
// block sum for blocks of up to 64 * (size of sh) threads
__device__ __forceinline__ double block_sum256(double v, double *sh) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wv] = v;
    __syncthreads();
    double s = 0.0;
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; ++i) s += sh[i];
    __syncthreads();
    return s;
}

// mean subtraction per image, FP64 sum (cuCorrNormalization.cu:98-139)
template <class T>
__global__ void __launch_bounds__(1024) k_sub_mean(T *data, int size, double *ssd) {
    __shared__ double sh[16];
    T *img = data + size_t(blockIdx.x) * size;
    double s1 = 0.0, s2 = 0.0;
    for (int i = threadIdx.x; i < size; i += blockDim.x) {
        const double v = double(img[i]);
        s1 += v;
        s2 += v * v;
    }
    s1 = block_sum256(s1, sh);
    const T mean = T(s1 / size);
    for (int i = threadIdx.x; i < size; i += blockDim.x) img[i] -= mean;
    if (ssd) {
        // sum of squares of the mean-subtracted image (the normalization's template term)
        s2 = block_sum256(s2, sh);
        if (threadIdx.x == 0) ssd[blockIdx.x] = fmax(s2 - s1 * s1 / size, 0.0);
    }
}

// ---------------------------------------------------------------------------
// correlation via one complex 2D FFT of z = t_pad + i*s (cuCorrFrequency.cu:63-111)
// ---------------------------------------------------------------------------
template <class T>
__global__ void k_pack_corr(const T *tmpl, int th, int tw, const T *srch, int sh, int sw,
                            cx<T> *z, int Lh, int Lw) {
    const int win = blockIdx.y;
    const int tot = Lh * Lw;
    const T *t = tmpl + size_t(win) * th * tw;
    const T *s = srch + size_t(win) * sh * sw;
    cx<T> *o = z + size_t(win) * tot;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < tot; i += gridDim.x * blockDim.x) {
        const int r = i / Lw, c = i - r * Lw;
        const T a = (r < th && c < tw) ? t[r * tw + c] : T(0);
        const T b = (r < sh && c < sw) ? s[r * sw + c] : T(0);
        o[i] = {a, b};
    }
}

// P[k] = conj(T[k]) S[k] / L^2 with T = (Z[k] + conj Z[-k])/2, S = (Z[k] - conj Z[-k])/(2i);
// P[-k] = conj(P[k]). One thread per (k, -k) pair.
template <class T>
__global__ void k_spec_product(cx<T> *z, int Lh, int Lw, T scale) {
    const int win = blockIdx.y;
    cx<T> *Z = z + size_t(win) * Lh * Lw;
    const int tot = Lh * Lw;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < tot; i += gridDim.x * blockDim.x) {
        const int r = i / Lw, c = i - r * Lw;
        const int rn = r ? Lh - r : 0, cn = c ? Lw - c : 0;
        const int j = rn * Lw + cn;
        if (j < i) continue;   // pair handled by the smaller index
        const cx<T> a = Z[i], bc = conj(Z[j]);
        const cx<T> tk = (a + bc) * T(0.5);
        const cx<T> d = (a - bc) * T(0.5);
        const cx<T> sk = {d.y, -d.x};   // d / i
        const cx<T> p = cmulc(tk, sk) * scale;
        Z[i] = p;
        if (j != i) Z[j] = conj(p);
    }
}

// ---------------------------------------------------------------------------
// Row-spectrum correlation over NL lag rows:
//   c[u][v] = (1/L) IFFT_L( sum_t conj(T_t) S_{t+u} )[v],   0 <= u < NL,
// equal to the zero-padded 2-D correlation (cuCorrFrequency.cu:63-79) at every
// valid lag for any row length L >= the search width. Row half spectra come
// from k_rowspec, are summed over template rows per bin and lag (k_mac_lags),
// and only the NL lag rows go back through a row IFFT (k_lagrows_real).
// ---------------------------------------------------------------------------

// Half spectra of the rows of [nimg][h][w] real images ([nimg][h][L/2+1]):
// rows 2j and 2j+1 are loaded as one complex row z = r_2j + i r_2j+1 of length
// L = d.n (zero beyond w and h), transformed in LDS, and split with
// A = (Z[k] + conj Z[-k]) / 2, B = (Z[k] - conj Z[-k]) / (2i). `g` row pairs per block.
template <class T>
__global__ void __launch_bounds__(256) k_rowspec(const T *img, int nimg, int h, int w, int g, FftDesc d,
                                                 cx<T> *spec) {
    extern __shared__ unsigned char smem_raw[];
    cx<T> *a = reinterpret_cast<cx<T> *>(smem_raw);
    cx<T> *b = a + g * d.n;
    const int L = d.n, np = (h + 1) / 2, H1 = L / 2 + 1;
    const int t0 = blockIdx.x * g;
    const int nr = min(g, nimg * np - t0);
    for (int i = threadIdx.x; i < nr * L; i += blockDim.x) {
        const int k = i / L, x = i - k * L;
        const int t = t0 + k, win = t / np, r = 2 * (t - win * np);
        const T *im = img + size_t(win) * h * w;
        a[i] = {x < w ? im[r * w + x] : T(0), (x < w && r + 1 < h) ? im[(r + 1) * w + x] : T(0)};
    }
    __syncthreads();
    const cx<T> *Z = fft_lds<T, -1>(a, b, d, nr, 1, L);
    for (int i = threadIdx.x; i < nr * H1; i += blockDim.x) {
        const int k = i / H1, kk = i - k * H1;
        const int t = t0 + k, win = t / np, j = t - win * np;
        const cx<T> za = Z[k * L + kk], bc = conj(Z[k * L + (kk ? L - kk : 0)]);
        cx<T> *sp = spec + size_t(win) * h * H1;
        sp[size_t(2 * j) * H1 + kk] = (za + bc) * T(0.5);
        if (2 * j + 1 < h) {
            const cx<T> dd = (za - bc) * T(0.5);
            sp[size_t(2 * j + 1) * H1 + kk] = {dd.y, -dd.x};
        }
    }
}

// The same half spectra with the row transform in place (fft_nat2dr, length
// L = R1 * R2, read back from digit-reversed order), G row pairs per block of
// odd LDS pitch
template <class T, int R1, int R2, int G, int NTH = 256>
__global__ void __launch_bounds__(NTH) k_rowspec_f(const T *img, int nimg, int h, int w, const cx<T> *tw,
                                                   const double *mean, cx<T> *spec) {
    constexpr int L = R1 * R2, PR = L + 1, H1 = L / 2 + 1;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T> *C = reinterpret_cast<cx<T> *>(smem_raw);   // [G][PR]
    __shared__ T rowmu[G];                             // the pair's image mean, subtracted on load when given
    __shared__ cx<T> twl[L];                           // twiddles, read from LDS
    for (int i = threadIdx.x; i < L; i += NTH) twl[i] = tw[i];
    const int np = (h + 1) / 2;
    const int t0 = blockIdx.x * G;
    const int nr = min(G, nimg * np - t0);
    if (mean) {
        if (threadIdx.x < nr) rowmu[threadIdx.x] = T(mean[(t0 + threadIdx.x) / np]);
        __syncthreads();
    }
    for (int i = threadIdx.x; i < nr * L; i += NTH) {
        const int k = i / L, x = i - k * L;
        const int t = t0 + k, win = t / np, r = 2 * (t - win * np);
        const T *im = img + size_t(win) * h * w;
        cx<T> v = {x < w ? im[r * w + x] : T(0), (x < w && r + 1 < h) ? im[(r + 1) * w + x] : T(0)};
        if (mean) {
            const T mu = rowmu[k];
            v = {x < w ? v.x - mu : T(0), (x < w && r + 1 < h) ? v.y - mu : T(0)};
        }
        C[k * PR + x] = v;
    }
    __syncthreads();
    fft_nat2dr<T, -1, R1, R2, NTH>(C, nr, PR, 1, twl);
    for (int i = threadIdx.x; i < nr * H1; i += NTH) {
        const int k = i / H1, kk = i - k * H1;
        const int t = t0 + k, win = t / np, j = t - win * np;
        const cx<T> za = C[k * PR + dr_pos<R1, R2>(kk)], bc = conj(C[k * PR + dr_pos<R1, R2>(kk ? L - kk : 0)]);
        cx<T> *sp = spec + size_t(win) * h * H1;
        sp[size_t(2 * j) * H1 + kk] = (za + bc) * T(0.5);
        if (2 * j + 1 < h) {
            const cx<T> dd = (za - bc) * T(0.5);
            sp[size_t(2 * j + 1) * H1 + kk] = {dd.y, -dd.x};
        }
    }
}

// the inverse row FFT of the lag rows and k_crop_real in one pass: rows t of
// length d.n (t = window * nl + lag) -> out[t][v] = Re, v < nlw; `g` rows per block
template <class T>
__global__ void __launch_bounds__(256) k_lagrows_real(const cx<T> *q, int nrows, int g, FftDesc d, int nlw, T *out) {
    extern __shared__ unsigned char smem_raw[];
    cx<T> *a = reinterpret_cast<cx<T> *>(smem_raw);
    cx<T> *b = a + g * d.n;
    const int r0 = blockIdx.x * g;
    const int nr = min(g, nrows - r0);
    const cx<T> *base = q + size_t(r0) * d.n;
    for (int i = threadIdx.x; i < nr * d.n; i += blockDim.x) a[i] = base[i];
    __syncthreads();
    const cx<T> *res = fft_lds<T, 1>(a, b, d, nr, 1, d.n);
    for (int i = threadIdx.x; i < nr * nlw; i += blockDim.x) {
        const int k = i / nlw, v = i - k * nlw;
        out[size_t(r0 + k) * nlw + v] = res[k * d.n + v].x;
    }
}

// q[u][k] = scale * sum_t conj(T_t[k]) S_{t+u}[k] for lags u0 <= u < u0 + K (u < nl),
// written as full Hermitian rows of length L for the inverse row FFT. A workgroup
// of NQ waves per (64 bins, lag group, window): wave w walks the w-th contiguous
// part of the search rows, keeping the template rows of its K lags in a sliding
// register window, so each thread's chain of row loads is 1/NQ of the whole;
// the NQ partial sums are added in wave order through LDS.
template <class T, int K, int NQ>
__global__ void __launch_bounds__(64 * NQ) k_mac_lags(const cx<T> *tspec, int th, const cx<T> *sspec, int sh, int nl,
                                                      int L, cx<T> *q, T scale) {
    __shared__ cx<T> part[NQ > 1 ? NQ - 1 : 1][K][64];
    const int H1 = L / 2 + 1, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int k = blockIdx.x * 64 + lane, u0 = blockIdx.y * K, win = blockIdx.z;
    const int r_end = min(sh, u0 + K - 1 + th), per = (r_end - u0 + NQ - 1) / NQ;
    const int ra = u0 + w * per, rb = min(r_end, ra + per);
    cx<T> acc[K], tw[K];
#pragma unroll
    for (int j = 0; j < K; ++j) acc[j] = tw[j] = cx<T>{T(0), T(0)};
    if (k < H1) {
        const cx<T> *ts = tspec + size_t(win) * th * H1 + k;
        const cx<T> *ss = sspec + size_t(win) * sh * H1 + k;
        // the template rows the lags meet just before row ra: tw[m] = T_{ra - u0 - 1 - m}
#pragma unroll
        for (int m = 0; m + 1 < K; ++m) {
            const int t = ra - u0 - 1 - m;
            if (t >= 0 && t < th) tw[m] = ts[size_t(t) * H1];
        }
        #pragma unroll 2
        for (int r = ra; r < rb; ++r) {
#pragma unroll
            for (int j = K - 1; j > 0; --j) tw[j] = tw[j - 1];     // tw[j] = T_{r - u0 - j}
            tw[0] = r - u0 < th ? ts[size_t(r - u0) * H1] : cx<T>{T(0), T(0)};
            const cx<T> sv = ss[size_t(r) * H1];
#pragma unroll
            for (int j = 0; j < K; ++j) acc[j] = cmac_conj(acc[j], tw[j], sv);
        }
    }
    if (NQ > 1) {
        if (w > 0) {
#pragma unroll
            for (int j = 0; j < K; ++j) part[w - 1][j][lane] = acc[j];
        }
        __syncthreads();
        if (w > 0) return;
#pragma unroll
        for (int p = 0; p + 1 < NQ; ++p)
#pragma unroll
            for (int j = 0; j < K; ++j) acc[j] = acc[j] + part[p][j][lane];
    }
    if (k >= H1) return;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        if (u0 + j < nl) {   // no early exit: the unrolled loop keeps acc[] in registers
            cx<T> *row = q + (size_t(win) * nl + u0 + j) * L;
            const cx<T> v = acc[j] * scale;
            row[k] = v;
            if (k && L - k != k) row[L - k] = conj(v);
        }
    }
}

// real part of the top-left oh x ow block
template <class T>
__global__ void k_crop_real(const cx<T> *z, int Lh, int Lw, T *out, int oh, int ow) {
    const int win = blockIdx.y;
    const int tot = oh * ow;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < tot; i += gridDim.x * blockDim.x) {
        const int r = i / ow, c = i - r * ow;
        out[size_t(win) * tot + i] = z[size_t(win) * Lh * Lw + r * Lw + c].x;
    }
}

// ---------------------------------------------------------------------------
// normalization (formula of cuCorrNormalization.cu:230-314 and
// cuCorrNormalizationSAT.cu:237-240) with FP64 window sums computed directly:
// per-column vertical sums over the template height, slid down one lag row at
// a time, and horizontal window sums from a block prefix scan of them. One
// 256-thread block per window; dynamic LDS 4*sw + 2 doubles. Used for search
// chips wider than the two-pass path below handles.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void block_scan256(double &x1, double &x2, double *sh4) {
    // inclusive scan of (x1, x2) over the 256 threads; sh4 holds 8 doubles
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const double y1 = __shfl_up(x1, d), y2 = __shfl_up(x2, d);
        if (lane >= d) { x1 += y1; x2 += y2; }
    }
    __syncthreads();
    if (lane == 63) { sh4[wv] = x1; sh4[4 + wv] = x2; }
    __syncthreads();
    for (int k = 0; k < wv; ++k) { x1 += sh4[k]; x2 += sh4[4 + k]; }
}

template <class T>
__global__ void __launch_bounds__(256) k_normalize(T *corr, int ch, int cw, const T *tmpl, int th, int tw,
                                                   const T *srch, int sh, int sw) {
    __shared__ double sh_red[8];
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double *V1 = reinterpret_cast<double *>(smem_raw);
    double *V2 = V1 + sw;
    double *P1 = V2 + sw;            // [sw + 1] prefix sums
    double *P2 = P1 + sw + 1;
    const int win = blockIdx.x;
    const T *t = tmpl + size_t(win) * th * tw;
    const T *s = srch + size_t(win) * sh * sw;
    double t2 = 0.0;
    for (int i = threadIdx.x; i < th * tw; i += blockDim.x) {
        const double v = double(t[i]);
        t2 += v * v;
    }
    t2 = block_sum256(t2, sh_red);
    for (int c = threadIdx.x; c < sw; c += blockDim.x) {
        double a1 = 0.0, a2 = 0.0;
        for (int r = 0; r < th; ++r) {
            const double v = double(s[size_t(r) * sw + c]);
            a1 += v;
            a2 += v * v;
        }
        V1[c] = a1;
        V2[c] = a2;
    }
    if (threadIdx.x == 0) { P1[0] = 0.0; P2[0] = 0.0; }
    __syncthreads();
    const int per = (sw + blockDim.x - 1) / blockDim.x;     // columns per thread, contiguous
    const int c0 = threadIdx.x * per, c1 = min(sw, c0 + per);
    const double inv = 1.0 / double(th * tw);
    T *cc = corr + size_t(win) * ch * cw;
    for (int u = 0; u < ch; ++u) {
        double x1 = 0.0, x2 = 0.0;
        for (int c = c0; c < c1; ++c) { x1 += V1[c]; x2 += V2[c]; }
        block_scan256(x1, x2, sh_red);
        double e1 = x1, e2 = x2;                              // inclusive end of this thread's segment
        for (int c = c1 - 1; c >= c0; --c) {
            P1[c + 1] = e1; P2[c + 1] = e2;
            e1 -= V1[c]; e2 -= V2[c];
        }
        __syncthreads();
        for (int v = threadIdx.x; v < cw; v += blockDim.x) {
            const double s1 = P1[v + tw] - P1[v], s2 = P2[v + tw] - P2[v];
            const double norm2 = (s2 - s1 * s1 * inv) * t2;
            cc[u * cw + v] = T(double(cc[u * cw + v]) / sqrt(norm2 + double(FLT_EPSILON)));
        }
        if (u + 1 < ch) {
            for (int c = threadIdx.x; c < sw; c += blockDim.x) {
                const double vn = double(s[size_t(u + th) * sw + c]), vo = double(s[size_t(u) * sw + c]);
                V1[c] += vn - vo;
                V2[c] += vn * vn - vo * vo;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// normalization in two barrier-free passes (same formula and FP64 sums as
// k_normalize): k_colsums slides the vertical window sums of s and s^2 down
// every column (one thread per column) into V[w][u][c]; k_norm_rows takes one
// lag row per wave, forms the horizontal window sums from a wave prefix scan of
// V and scales the correlation row. ssd[w] is the template's sum of squared
// deviations (k_sub_mean).
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(64) k_colsums(const T *srch, int sh, int sw, int th, int nl, double *V1,
                                                double *V2) {
    const int w = blockIdx.y, c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= sw) return;
    const T *s = srch + size_t(w) * sh * sw + c;
    double b1[4] = {0.0, 0.0, 0.0, 0.0}, b2[4] = {0.0, 0.0, 0.0, 0.0};
    int r = 0;
    for (; r + 4 <= th; r += 4) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const double v = double(s[size_t(r + k) * sw]);
            b1[k] += v;
            b2[k] += v * v;
        }
    }
    for (; r < th; ++r) {
        const double v = double(s[size_t(r) * sw]);
        b1[0] += v;
        b2[0] += v * v;
    }
    double a1 = (b1[0] + b1[1]) + (b1[2] + b1[3]), a2 = (b2[0] + b2[1]) + (b2[2] + b2[3]);
    double *o1 = V1 + size_t(w) * nl * sw + c, *o2 = V2 + size_t(w) * nl * sw + c;
    o1[0] = a1;
    o2[0] = a2;
    for (int u = 1; u < nl; ++u) {
        const double vn = double(s[size_t(u + th - 1) * sw]), vo = double(s[size_t(u - 1) * sw]);
        a1 += vn - vo;
        a2 += vn * vn - vo * vo;
        o1[size_t(u) * sw] = a1;
        o2[size_t(u) * sw] = a2;
    }
}

template <class T, int CPL>
__global__ void __launch_bounds__(256) k_norm_rows(T *corr, int n, int ch, int cw, int th, int tw, const double *V1,
                                                   const double *V2, int sw, const double *ssd) {
    const int lane = threadIdx.x & 63;
    const int g = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int w = g / ch, u = g - w * ch;
    if (w >= n) return;
    const double *v1 = V1 + (size_t(w) * ch + u) * sw, *v2 = V2 + (size_t(w) * ch + u) * sw;
    // inclusive prefix over columns: p[m] at column lane + 64 m
    double p1[CPL], p2[CPL];
    double c1 = 0.0, c2 = 0.0;
#pragma unroll
    for (int m = 0; m < CPL; ++m) {
        const int c = lane + 64 * m;
        double x1 = c < sw ? v1[c] : 0.0, x2 = c < sw ? v2[c] : 0.0;
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
    const double t2 = ssd[w], inv = 1.0 / double(th * tw);
    T *cr = corr + (size_t(w) * ch + u) * cw;
    for (int k = 0; k * 64 < cw; ++k) {
        const int v = lane + 64 * k;
        const int hi = min(v + tw - 1, sw - 1), lo = v - 1;
        double ph1 = 0, ph2 = 0, pl1 = 0, pl2 = 0;
#pragma unroll
        for (int m = 0; m < CPL; ++m) {
            const double a1 = __shfl(p1[m], hi & 63), a2 = __shfl(p2[m], hi & 63);
            const double b1 = __shfl(p1[m], lo & 63), b2 = __shfl(p2[m], lo & 63);
            if ((hi >> 6) == m) { ph1 = a1; ph2 = a2; }
            if (lo >= 0 && (lo >> 6) == m) { pl1 = b1; pl2 = b2; }
        }
        if (v < cw) {
            const double s1 = ph1 - pl1, s2 = ph2 - pl2;
            const double norm2 = (s2 - s1 * s1 * inv) * t2;
            cr[v] = T(double(cr[v]) / sqrt(norm2 + double(FLT_EPSILON)));
        }
    }
}

// ---------------------------------------------------------------------------
// argmax with the reference tie rule (cuOffset.cu:23-84), one block per surface
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(256) k_maxloc(const T *surf, int h, int w, int32_t *loc, T *val, float *margin) {
    __shared__ T shv[4], shv2[4];
    __shared__ uint64_t shk[4];
    const T *s = surf + size_t(blockIdx.x) * h * w;
    Top2<T> top;
    for (int i = threadIdx.x; i < h * w; i += blockDim.x) top.add(s[i], maxloc_key(uint32_t(i)));
    top.template reduce<256>(shv, shv2, shk);
    if (threadIdx.x == 0) {
        const int idx = top.index();
        loc[2 * blockIdx.x] = idx / w;
        loc[2 * blockIdx.x + 1] = idx % w;
        if (val) val[blockIdx.x] = top.v1;
        if (margin) margin[blockIdx.x] = top.margin();
    }
}

// ---------------------------------------------------------------------------
// covariance, SNR, zoom start/shift; one thread per window
// (cuEstimateStats.cu:20-110, cuArraysCopy.cu:249-284, cuCorrNormalization.cu:143-192,
//  cuOffset.cu:190-241)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void adjust_axis(int maxloc, int old_range, int new_range, int &start, int &shift) {
    start = maxloc - new_range;
    shift = 0;
    const int rbound = 2 * (old_range - new_range);
    if (start < 0) { shift = -start; start = 0; }
    else if (start > rbound) { shift = start - rbound; start = rbound; }
}

// Covariance and SNR of one window's normalized surface, computed by one wave:
// lanes share the SNR patch sum, lane 0 writes both results.
template <class T>
__device__ void window_stats(const T *c, int nx, int ny, int px, int py, double peak, int template_size,
                             int patch_h, int patch_w, float *cov3, float *snr) {
    const int lane = threadIdx.x & 63;
    // SNR over the patch centred on the peak (offset -size/2), zero outside
    double sum = 0.0;
    int count = 0;
    for (int k = lane; k < patch_h * patch_w; k += 64) {
        const int i = k / patch_w, j = k - i * patch_w;
        const int x = px + i - patch_h / 2, y = py + j - patch_w / 2;
        if (x >= 0 && x < nx && y >= 0 && y < ny) {
            const double v = double(c[x * ny + y]);
            sum += v * v;
            ++count;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        sum += __shfl_xor(sum, o);
        count += __shfl_xor(count, o);
    }
    if (lane != 0) return;
    // covariance
    double cxx = 99.0, cyy = 99.0, cxy = 0.0;
    if (!(px - 1 < 0 || py - 1 < 0 || px + 1 >= nx || py + 1 >= ny)) {
        auto g = [&](int dx, int dy) { return double(c[(px + dx) * ny + py + dy]); };
        double dxx = -(g(1, 0) + g(-1, 0) - 2.0 * g(0, 0)) * template_size;
        double dyy = -(g(0, 1) + g(0, -1) - 2.0 * g(0, 0)) * template_size;
        double dxy = (g(1, 1) + g(-1, -1) - g(1, -1) - g(-1, 1)) * 0.25 * template_size;
        double n2 = fmax(1.0 - peak, 0.0);
        const double n4 = n2 * n2 * 0.5 * template_size;
        n2 *= 2.0;
        const double u = dxy * dxy - dxx * dyy;
        const double u2 = u * u;
        if (fabs(u) >= 1e-2) {
            cxx = (-n2 * u * dyy + n4 * (dyy * dyy + dxy * dxy)) / u2;
            cyy = (-n2 * u * dxx + n4 * (dxx * dxx + dxy * dxy)) / u2;
            cxy = (n2 * u * dxy - n4 * (dxx + dyy) * dxy) / u2;
        }
    }
    cov3[0] = float(cxx); cov3[1] = float(cyy); cov3[2] = float(cxy);
    const double mean = (sum - peak * peak) / double(count - 1);
    *snr = float(peak * peak / mean);
}

// one 64-thread block (one wave) per window
template <class T>
__global__ void __launch_bounds__(64) k_stats(const T *corr, int nx, int ny, int n, const int32_t *loc, const T *val,
                                              int template_size, int patch_h, int patch_w, int hs_down,
                                              int hs_across, int hz, float *cov, float *snr, int32_t *start,
                                              int32_t *shift) {
    const int w = blockIdx.x;
    const int px = loc[2 * w], py = loc[2 * w + 1];
    window_stats(corr + size_t(w) * nx * ny, nx, ny, px, py, double(val[w]), template_size, patch_h, patch_w,
                 cov + 3 * w, snr + w);
    if (threadIdx.x == 0) {
        int s0, h0, s1, h1;
        adjust_axis(px, hs_down, hz, s0, h0);
        adjust_axis(py, hs_across, hz, s1, h1);
        start[2 * w] = s0; start[2 * w + 1] = s1;
        shift[2 * w] = h0; shift[2 * w + 1] = h1;
    }
}

// ---------------------------------------------------------------------------
// deramp method 1 (cuDeramp.cu:67-148) with float2.h's operator*= (imaginary
// part reuses the updated real part); sums, phases and rotation in FP64. The
// phase is separable, c*phx + r*phy: k_deramp_phase writes one FP64 table per
// chip, e^{i r phy} for its h rows then e^{i c phx} for its w columns, and the
// forward row FFT applies the rotation as it loads the chip (deramp_apply).
// ---------------------------------------------------------------------------
// chip win of [nimg][ih][iw] images, cropped at `start` (row, column) when given
template <class T>
__device__ __forceinline__ const cx<T> *crop_origin(const cx<T> *src, int ih, int iw, const int32_t *start, int win) {
    const cx<T> *o = src + size_t(win) * ih * iw;
    return start ? o + start[2 * win] * iw + start[2 * win + 1] : o;
}

template <class T>
__global__ void __launch_bounds__(1024) k_deramp_phase(const cx<T> *src, int ih, int iw, const int32_t *start, int h,
                                                       int w, cx<double> *tab) {
    __shared__ double sh[16];
    const cx<T> *z = crop_origin(src, ih, iw, start, blockIdx.x);
    double ayr = 0, ayi = 0, axr = 0, axi = 0;
    for (int i = threadIdx.x; i < h * w; i += blockDim.x) {
        const int r = i / w, c = i - r * w;
        const cx<T> v = z[r * iw + c];
        const double zr = double(v.x), zi = double(v.y);
        if (c < w - 1) {   // z * conj(z_right)
            const cx<T> u = z[r * iw + c + 1];
            const double br = double(u.x), bi = double(u.y);
            ayr += zr * br + zi * bi;
            ayi += zi * br - zr * bi;
        }
        if (r < h - 1) {   // z * conj(z_below)
            const cx<T> u = z[(r + 1) * iw + c];
            const double br = double(u.x), bi = double(u.y);
            axr += zr * br + zi * bi;
            axi += zi * br - zr * bi;
        }
    }
    ayr = block_sum256(ayr, sh);
    ayi = block_sum256(ayi, sh);
    axr = block_sum256(axr, sh);
    axi = block_sum256(axi, sh);
    const double phy = atan2(ayi, ayr);
    const double phx = atan2(axi, axr);
    cx<double> *rowf = tab + size_t(blockIdx.x) * (h + w), *colf = rowf + h;
    for (int r = threadIdx.x; r < h; r += blockDim.x) sincos(double(r) * phy, &rowf[r].y, &rowf[r].x);
    for (int c = threadIdx.x; c < w; c += blockDim.x) sincos(double(c) * phx, &colf[c].y, &colf[c].x);
}

template <class T>
__device__ __forceinline__ cx<T> deramp_apply(cx<T> v, cx<double> rf, cx<double> cf) {
    const cx<double> f = cmul(cf, rf);
    const double x = double(v.x), y = double(v.y);
    const double nx = x * f.x - y * f.y;
    const double ny = y * f.x + nx * f.y;
    return {T(nx), T(ny)};
}

// forward row FFT of h x d.n chips read from [nimg][ih][iw] sources (cropped at
// `start` when given: the secondary zoom-in window, cuArraysCopy.cu:209-224),
// deramped on load when `tab` is given, written to dst [nimg][h][d.n] (dst may
// be the source itself when there is no crop); `g` rows per block
template <class T, int DIR>
__global__ void __launch_bounds__(256) k_fft_rows_chip(const cx<T> *src, int ih, int iw, const int32_t *start,
                                                       const cx<double> *tab, cx<T> *dst, int nimg, int h, int g,
                                                       FftDesc d) {
    extern __shared__ unsigned char smem_raw[];
    cx<T> *a = reinterpret_cast<cx<T> *>(smem_raw);
    cx<T> *b = a + g * d.n;
    const int w = d.n, r0 = blockIdx.x * g;
    const int nr = min(g, nimg * h - r0);
    for (int i = threadIdx.x; i < nr * w; i += blockDim.x) {
        const int k = i / w, c = i - k * w;
        const int t = r0 + k, win = t / h, r = t - win * h;
        cx<T> v = crop_origin(src, ih, iw, start, win)[r * iw + c];
        if (tab) {
            const cx<double> *rowf = tab + size_t(win) * (h + w);
            v = deramp_apply(v, rowf[r], rowf[h + c]);
        }
        a[i] = v;
    }
    __syncthreads();
    const cx<T> *res = fft_lds<T, DIR>(a, b, d, nr, 1, w);
    cx<T> *base = dst + size_t(r0) * w;
    for (int i = threadIdx.x; i < nr * w; i += blockDim.x) base[i] = res[i];
}

// ---------------------------------------------------------------------------
// x2 chip oversampling, separable form. The quadrant-padded inverse 2-D FFT of
// the reference's oversampler (cuOverSampler.cpp:72-78 with
// cuArraysPadding.cu:59-103) is the tensor product of one 1-D interpolation
// per axis, because the padding and both transforms factor by axis. Along an
// axis of N samples, output sample 2m is input sample m, and output sample
// 2m+1 is (1/N) FFT_-(FFT_+(x) e^{-i pi k_s/N})[m] with k_s = k for k < N/2 and
// k - N otherwise (the padding puts bin N/2 among the negative frequencies).
// Column pass (k_os2_cols): the odd output rows Y of each chip. Row pass
// (k_os2_rows): every output row (chip row m or Y row m), |.| at even columns
// and |.| of its interpolation at odd columns. Chips are read from
// [nimg][ih][iw] sources, cropped at `start` when given and deramped with
// `tab` when given, as in k_fft_rows_chip.
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ cx<T> half_shift(int k, int n) {
    const int ks = k < n / 2 ? k : k - n;
    double s, c;
    sincospi(-double(ks) / double(n), &s, &c);
    return {T(c / n), T(s / n)};
}

template <class T>
__device__ __forceinline__ cx<T> chip_pixel(const cx<T> *src, int ih, int iw, const int32_t *start,
                                            const cx<double> *tab, int win, int h, int w, int r, int c) {
    cx<T> v = crop_origin(src, ih, iw, start, win)[r * iw + c];
    if (tab) {
        const cx<double> *rowf = tab + size_t(win) * (h + w);
        v = deramp_apply(v, rowf[r], rowf[h + c]);
    }
    return v;
}

// `cb` chip columns per block (blockIdx.x), one chip per blockIdx.y; d.n = h;
// Y: [nimg][h][w]
template <class T>
__global__ void __launch_bounds__(256) k_os2_cols(const cx<T> *src, int ih, int iw, const int32_t *start,
                                                  const cx<double> *tab, int w, int cb, FftDesc d, cx<T> *Y) {
    extern __shared__ unsigned char smem_raw[];
    const int h = d.n;
    cx<T> *a = reinterpret_cast<cx<T> *>(smem_raw);
    cx<T> *b = a + cb * h;
    cx<T> *sh = b + cb * h;
    const int win = blockIdx.y, c0 = blockIdx.x * cb, nc = min(cb, w - c0);
    for (int i = threadIdx.x; i < h * nc; i += blockDim.x) {
        const int r = i / nc, c = i - r * nc;
        a[r * cb + c] = chip_pixel(src, ih, iw, start, tab, win, h, w, r, c0 + c);
    }
    for (int k = threadIdx.x; k < h; k += blockDim.x) sh[k] = half_shift<T>(k, h);
    __syncthreads();
    cx<T> *X = fft_lds<T, +1>(a, b, d, nc, cb, 1);
    cx<T> *other = X == a ? b : a;
    for (int i = threadIdx.x; i < h * nc; i += blockDim.x) {
        const int k = i / nc, c = i - k * nc;
        X[k * cb + c] = cmul(X[k * cb + c], sh[k]);
    }
    __syncthreads();
    const cx<T> *res = fft_lds<T, -1>(X, other, d, nc, cb, 1);
    cx<T> *y = Y + size_t(win) * h * w + c0;
    for (int i = threadIdx.x; i < h * nc; i += blockDim.x) {
        const int r = i / nc, c = i - r * nc;
        y[size_t(r) * w + c] = res[r * cb + c];
    }
}

// `g` output rows per block over the nimg * 2h rows; d.n = w; out: [nimg][2h][2w]
// real, |.| (OUT 1) of the oversampled chip
template <class T>
__global__ void __launch_bounds__(256) k_os2_rows(const cx<T> *src, int ih, int iw, const int32_t *start,
                                                  const cx<double> *tab, const cx<T> *Y, int nimg, int h, int g,
                                                  FftDesc d, T *out) {
    extern __shared__ unsigned char smem_raw[];
    const int w = d.n;
    cx<T> *a = reinterpret_cast<cx<T> *>(smem_raw);
    cx<T> *b = a + g * w;
    cx<T> *sh = b + g * w;
    T *amp = reinterpret_cast<T *>(sh + w);
    const int t0 = blockIdx.x * g;
    const int nr = min(g, nimg * 2 * h - t0);
    for (int i = threadIdx.x; i < nr * w; i += blockDim.x) {
        const int k = i / w, c = i - k * w;
        const int t = t0 + k, win = t / (2 * h), tr = t - win * 2 * h, m = tr >> 1;
        const cx<T> v = (tr & 1) ? Y[(size_t(win) * h + m) * w + c] : chip_pixel(src, ih, iw, start, tab, win, h, w, m, c);
        a[i] = v;
        amp[i] = amag(v);
    }
    for (int k = threadIdx.x; k < w; k += blockDim.x) sh[k] = half_shift<T>(k, w);
    __syncthreads();
    cx<T> *X = fft_lds<T, +1>(a, b, d, nr, 1, w);
    cx<T> *other = X == a ? b : a;
    for (int i = threadIdx.x; i < nr * w; i += blockDim.x) {
        const int c = i % w;
        X[i] = cmul(X[i], sh[c]);
    }
    __syncthreads();
    const cx<T> *res = fft_lds<T, -1>(X, other, d, nr, 1, w);
    T *o = out + size_t(t0) * 2 * w;
    for (int i = threadIdx.x; i < nr * w; i += blockDim.x) {
        const int k = i / w, c = i - k * w;
        *reinterpret_cast<cx<T> *>(o + size_t(k) * 2 * w + 2 * c) = cx<T>{amp[i], amag(res[i])};
    }
}

// The same two passes for chip sides N = R1 * R2 with the transforms in place
// (fft_nat2dr, which applies the shift factors as it stores each bin, then
// fft_dr2nat: the spectrum stays in digit-reversed order), one chip line per
// LDS row of odd pitch.

// CB chip columns per block (blockIdx.x), one chip per blockIdx.y; chip height
// R1 * R2; Y: [nimg][R1 * R2][w]
template <class T, int R1, int R2, int CB, int NTH = 256>
__global__ void __launch_bounds__(NTH) k_os2_cols_f(const cx<T> *src, int ih, int iw, const int32_t *start,
                                                    const cx<double> *tab, int w, const cx<T> *tw, cx<T> *Y) {
    constexpr int N = R1 * R2, PC = CB + 1;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T> *C = reinterpret_cast<cx<T> *>(smem_raw);   // [N][PC]: column c at C[c + r * PC]
    cx<T> *sh = C + N * PC;                            // shift factors in bin order
    __shared__ cx<T> twl[N];                           // twiddles, read from LDS
    const int win = blockIdx.y, c0 = blockIdx.x * CB, nc = min(CB, w - c0);
    for (int i = threadIdx.x; i < N * CB; i += NTH) {
        const int r = i / CB, c = i - r * CB;
        if (c < nc) C[r * PC + c] = chip_pixel(src, ih, iw, start, tab, win, N, w, r, c0 + c);
    }
    for (int k = threadIdx.x; k < N; k += NTH) {
        sh[k] = half_shift<T>(k, N);
        twl[k] = tw[k];
    }
    __syncthreads();
    // the shift multiplies each bin as the forward transform stores it
    fft_nat2dr<T, +1, R1, R2, NTH>(C, nc, 1, PC, twl, sh);
    fft_dr2nat<T, -1, R1, R2, NTH>(C, nc, 1, PC, twl);
    cx<T> *y = Y + size_t(win) * N * w + c0;
    for (int i = threadIdx.x; i < N * CB; i += NTH) {
        const int r = i / CB, c = i - r * CB;
        if (c < nc) y[size_t(r) * w + c] = C[r * PC + c];
    }
}

// G output rows per block over the nimg * 2h rows; chip width R1 * R2;
// out: [nimg][2h][2 R1 R2] amplitudes. With `part` (G dividing 2h, so that a
// block holds rows of one chip only), each block also writes the FP64 sums of
// its amplitudes and of their squares to part[2 * block], part[2 * block + 1].
template <class T, int R1, int R2, int G, int NTH = 256>
__global__ void __launch_bounds__(NTH) k_os2_rows_f(const cx<T> *src, int ih, int iw, const int32_t *start,
                                                    const cx<double> *tab, const cx<T> *Y, int nimg, int h,
                                                    const cx<T> *tw, T *out, double *part) {
    constexpr int N = R1 * R2, PR = N + 1;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T> *C = reinterpret_cast<cx<T> *>(smem_raw);   // [G][PR]
    cx<T> *sh = C + G * PR;                            // shift factors in bin order
    T *amp = reinterpret_cast<T *>(sh + N);            // [G][N]
    __shared__ cx<T> twl[N];                           // twiddles, read from LDS
    const int t0 = blockIdx.x * G;
    const int nr = min(G, nimg * 2 * h - t0);
    for (int i = threadIdx.x; i < nr * N; i += NTH) {
        const int k = i / N, c = i - k * N;
        const int t = t0 + k, win = t / (2 * h), tr = t - win * 2 * h, m = tr >> 1;
        const cx<T> v = (tr & 1) ? Y[(size_t(win) * h + m) * N + c] : chip_pixel(src, ih, iw, start, tab, win, h, N, m, c);
        C[k * PR + c] = v;
        amp[i] = amag(v);
    }
    for (int k = threadIdx.x; k < N; k += NTH) {
        sh[k] = half_shift<T>(k, N);
        twl[k] = tw[k];
    }
    __syncthreads();
    // the shift multiplies each bin as the forward transform stores it
    fft_nat2dr<T, +1, R1, R2, NTH>(C, nr, PR, 1, twl, sh);
    fft_dr2nat<T, -1, R1, R2, NTH>(C, nr, PR, 1, twl);
    T *o = out + size_t(t0) * 2 * N;
    double s1 = 0.0, s2 = 0.0;
    for (int i = threadIdx.x; i < nr * N; i += NTH) {
        const int k = i / N, c = i - k * N;
        const T a0 = amp[i], a1 = amag(C[k * PR + c]);
        *reinterpret_cast<cx<T> *>(o + size_t(k) * 2 * N + 2 * c) = cx<T>{a0, a1};
        s1 += double(a0) + double(a1);
        s2 += double(a0) * double(a0) + double(a1) * double(a1);
    }
    if (part) {
        __shared__ double red[16];
        s1 = block_sum256(s1, red);
        s2 = block_sum256(s2, red);
        if (threadIdx.x == 0) {
            part[2 * blockIdx.x] = s1;
            part[2 * blockIdx.x + 1] = s2;
        }
    }
}

// idx[0..cap) = the first cap windows (in order) whose margins are below the
// thresholds (NaN counts as below), then the first windows that are not, for
// cap <= n. One workgroup of 1024 threads: per 1024-window slice, a ballot
// per wave and the wave totals give each flagged window its slot.
static __global__ void __launch_bounds__(1024) k_select_near(const float *mc, const float *mf, int n, float tc,
                                                             float tf, int cap, int32_t *idx) {
    __shared__ int wtot[16];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    int filled = 0;                     // slots taken so far (the same in every thread)
    for (int pass = 0; pass < 2 && filled < cap; ++pass) {
        for (int c0 = 0; c0 < n && filled < cap; c0 += blockDim.x) {
            const int i = c0 + threadIdx.x;
            bool flag = false;
            if (i < n) {
                const bool near = !(mc[i] >= tc) || !(mf[i] >= tf);
                flag = pass == 0 ? near : !near;
            }
            const unsigned long long m = __ballot(flag);
            const int before = __popcll(m & ((1ull << lane) - 1ull));
            if (lane == 0) wtot[wv] = __popcll(m);
            __syncthreads();
            int off = 0, tot = 0;
            for (int k = 0; k < nw; ++k) {
                const int t = wtot[k];
                if (k < wv) off += t;
                tot += t;
            }
            const int slot = filled + off + before;
            if (flag && slot < cap) idx[slot] = i;
            filled += tot;
            __syncthreads();
        }
        if (filled > cap) filled = cap;
    }
}

// mean and sum of squared deviations of each [size] image from nb block sums
// (k_os2_rows_f's `part`): the values k_sub_mean finds, for the consumer to
// subtract the mean as it loads the image (k_rowspec_f's `mean`)
static __global__ void k_mean_parts(const double *part, int nb, int size, int nimg, double *mean, double *ssd) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nimg) return;
    double s1 = 0.0, s2 = 0.0;
    for (int j = 0; j < nb; ++j) {
        s1 += part[2 * (size_t(w) * nb + j)];
        s2 += part[2 * (size_t(w) * nb + j) + 1];
    }
    mean[w] = s1 / size;
    if (ssd) ssd[w] = fmax(s2 - s1 * s1 / size, 0.0);
}

template <class T>
__global__ void k_trim_complex(const T *in, int ih, int iw, cx<T> *out, int z) {
    const int win = blockIdx.y;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < z * z; i += gridDim.x * blockDim.x) {
        const int r = i / z, c = i - r * z;
        out[size_t(win) * z * z + i] = {in[size_t(win) * ih * iw + r * iw + c], T(0)};
    }
}

template <class T>
__global__ void k_trim_real(const T *in, int ih, int iw, T *out, int z) {
    const int win = blockIdx.y;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < z * z; i += gridDim.x * blockDim.x) {
        const int r = i / z, c = i - r * z;
        out[size_t(win) * z * z + i] = in[size_t(win) * ih * iw + r * iw + c];
    }
}

// windowed-sinc surface oversampling (cuSincOverSampler.cu:84-195): zero output,
// then the (2*2*covs+1)^2 patch around out/2 + shift*factor
template <class T>
__global__ void k_sinc(const T *in, int nin, T *out, int nout, const int32_t *shift, int factor,
                       const float *filt, int covs, int dec, int intplength) {
    const int win = blockIdx.y;
    const int rng = 2 * covs, st = nout / 2 - rng, size = 2 * rng + 1;
    T *o = out + size_t(win) * nout * nout;
    const T *a = in + size_t(win) * nin * nin;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nout * nout; i += gridDim.x * blockDim.x) o[i] = T(0);
    __syncthreads();   // same block writes zeros then values only for its own indices below
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < size * size; i += gridDim.x * blockDim.x) {
        const int ix = i / size, iy = i - ix * size;
        int ox = ix + st + shift[2 * win] * factor;
        if (ox >= nout) ox -= nout;
        int oy = iy + st + shift[2 * win + 1] * factor;
        if (oy >= nout) oy -= nout;
        const int xo = ox / covs, yo = oy / covs;
        const int xf = int((float(ox) / covs - xo) * dec);
        const int yf = int((float(oy) / covs - yo) * dec);
        double acc = 0.0, wsum = 0.0;
        for (int p = 0; p < intplength; ++p) {
            int x = xo - p + intplength / 2;
            if (x < 0) x += nin;
            if (x >= nin) x -= nin;
            const double wx = filt[p * dec + xf];
            for (int q = 0; q < intplength; ++q) {
                int y = yo - q + intplength / 2;
                if (y < 0) y += nin;
                if (y >= nin) y -= nin;
                const double wgt = wx * double(filt[q * dec + yf]);
                wsum += wgt;
                acc += double(a[x * nin + y]) * wgt;
            }
        }
        o[ox * nout + oy] = T(acc / wsum);
    }
}

// final offset (cuOffset.cu:137-146)
// ---------------------------------------------------------------------------
// 16x16 -> 256x256 surface oversampling (cuOverSampler.cpp:141-148, real part
// cuArraysCopy.cu:326-339), argmax (cuOffset.cu:23-84) and sub-pixel offset
// (cuOffset.cu:137-187) in one kernel, one 256-thread block per window. The
// padded 256x256 spectrum has only 16x16 non-zero bins, so with p = 16a + b,
// q = 16c + d and signed bins ks:
//   G[k1][16c+d]   = DFT16_k2( A[k1][k2] e^{-2 pi i d ks2/256} )[c]
//   out[16a+b][q]  = Re DFT16_k1( G[k1][q] e^{-2 pi i b ks1/256} )[a]
// where A = (+i 2-D DFT of the trimmed patch) / 256. Needs zoom = 16 and a
// surface oversampling factor of 16. LDS: (2*256 + 16*256) cx<T>.
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(256) k_surface(const T *corr, int ch, int cw, const int32_t *start, int hs_down,
                                                 int hs_across, float ratio, const cx<T> *tw256, int32_t *zidx,
                                                 float *margin, float *offset) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T> *a16 = reinterpret_cast<cx<T> *>(smem_raw);
    cx<T> *tws = a16 + 256;
    cx<T> *G = tws + 256;
    __shared__ T shv[4], shv2[4];
    __shared__ uint64_t shk[4];
    const int win = blockIdx.x;
    const T *c = corr + size_t(win) * ch * cw;
    {
        const int r = threadIdx.x >> 4, q = threadIdx.x & 15;
        a16[threadIdx.x] = {c[r * cw + q] * T(1.0 / 256.0), T(0)};
        tws[threadIdx.x] = tw256[threadIdx.x];
    }
    __syncthreads();
    if (threadIdx.x < 16) {
        cx<T> x[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) x[q] = a16[threadIdx.x * 16 + q];
        dft16<T, +1>(x);
#pragma unroll
        for (int q = 0; q < 16; ++q) a16[threadIdx.x * 16 + q] = x[q];
    }
    __syncthreads();
    if (threadIdx.x < 16) {
        cx<T> x[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) x[r] = a16[r * 16 + threadIdx.x];
        dft16<T, +1>(x);
#pragma unroll
        for (int r = 0; r < 16; ++r) a16[r * 16 + threadIdx.x] = x[r];
    }
    __syncthreads();
    {
        const int k1 = threadIdx.x >> 4, dd = threadIdx.x & 15;
        cx<T> x[16];
#pragma unroll
        for (int k2 = 0; k2 < 16; ++k2) {
            const int ks = k2 < 8 ? k2 : k2 - 16;
            x[k2] = cmul(a16[k1 * 16 + k2], tws[(dd * ks + 4096) & 255]);
        }
        dft16<T, -1>(x);
#pragma unroll
        for (int cc = 0; cc < 16; ++cc) G[k1 * 256 + 16 * cc + dd] = x[cc];
    }
    __syncthreads();
    // Column pruning: every value of column q is at most B_q = sum_k1 |G[k1][q]|
    // and the surface passes through the 16x16 samples, so its maximum is at
    // least their largest value amax. Columns with B_q (1 + 2^-10) + 2^-11 |a|max
    // < amax hold neither the argmax nor any value within 2^-12 |a|max of it, and
    // are skipped (argmax, tie rule and margins below that are unchanged; NaN
    // bounds keep their column).
    __shared__ uint64_t masks[4];
    __shared__ int cols[256];
    T amax = -INFINITY, amag = T(0);
    for (int i = threadIdx.x & 63; i < 256; i += 64) {
        const T v = c[(i >> 4) * cw + (i & 15)];
        amax = fmax(amax, v);
        amag = fmax(amag, fabs(v));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        amax = fmax(amax, __shfl_xor(amax, o));
        amag = fmax(amag, __shfl_xor(amag, o));
    }
    bool keep;
    {
        T bq = T(0);
#pragma unroll
        for (int k1 = 0; k1 < 16; ++k1) {
            const cx<T> g = G[k1 * 256 + threadIdx.x];
            bq += ampcor::amag(g);
        }
        keep = !(bq * T(1.0 + 0x1p-10) + T(0x1p-11) * amag < amax);
    }
    const uint64_t wm = __ballot(keep);
    if ((threadIdx.x & 63) == 0) masks[threadIdx.x >> 6] = wm;
    __syncthreads();
    int ncols = 0, before = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        const int n = __popcll(masks[w]);
        if (w < (threadIdx.x >> 6)) before += n;
        ncols += n;
    }
    if (keep) cols[before + __popcll(wm & ((uint64_t(1) << (threadIdx.x & 63)) - 1))] = threadIdx.x;
    __syncthreads();
    Top2<T> top;
    #pragma unroll 1
    for (int t = threadIdx.x; t < 16 * ncols; t += 256) {
        const int b = t / ncols, qq = cols[t - b * ncols];
        cx<T> x[16];
#pragma unroll
        for (int k1 = 0; k1 < 16; ++k1) {
            const int ks = k1 < 8 ? k1 : k1 - 16;
            x[k1] = cmul(G[k1 * 256 + qq], tws[(b * ks + 4096) & 255]);
        }
        dft16<T, -1>(x);
#pragma unroll
        for (int aa = 0; aa < 16; ++aa) top.add_idx(x[aa].x, uint32_t((16 * aa + b) * 256 + qq));
    }
    top.template reduce<256>(shv, shv2, shk);
    if (threadIdx.x == 0) {
        const int idx = top.index();
        const int px = idx >> 8, py = idx & 255;
        zidx[2 * win] = px;
        zidx[2 * win + 1] = py;
        offset[2 * win] = ratio * float(px) + float(start[2 * win]) - float(hs_down);
        offset[2 * win + 1] = ratio * float(py) + float(start[2 * win + 1]) - float(hs_across);
        if (margin) margin[win] = top.margin();
    }
}

static __global__ void k_subpixel(const int32_t *start, const int32_t *zidx, int n, float ratio, int hs_down,
                           int hs_across, float *offset) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n) return;
    offset[2 * w] = ratio * float(zidx[2 * w]) + float(start[2 * w]) - float(hs_down);
    offset[2 * w + 1] = ratio * float(zidx[2 * w + 1]) + float(start[2 * w + 1]) - float(hs_across);
}

// ---------------------------------------------------------------------------
// FP64 sub-pixel refinement of the oversampled-surface peak. No counterpart in
// the reference, whose offset is the bare argmax on the 1/(surf_ovs*raw_ovs)
// grid (cuOffset.cu:137-187). The surface is the FFT interpolation of the
// trimmed Z x Z surface a (spectrum zero-padded to N = Z*ovs with the Nyquist
// bin on the negative side, real part; cuOverSampler.cpp:141-148), so at any
// integer (X, Y) of the N x N grid
//   s(X, Y) = Re (1/Z^2) sum_{r,c} a[r][c] D(r, X) D(c, Y),
//   D(r, X) = sum_{f=-Z/2}^{Z/2-1} e^{2 pi i f (r ovs - X)/N},
// which is evaluated in FP64 at the peak (px, py) and its four axis neighbours
// (wrapping, as the interpolation is periodic); a parabola through each axis'
// three values gives the vertex d in [-1/2, 1/2], and the refined offset is
// (p + d)/(ovs * raw_ovs) + start - hs. One wave per window; the trimmed
// surface is the top-left Z x Z of surf ([n][ch][cw]); LDS: refine_lds(Z).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double parabola_vertex(double fm, double f0, double fp) {
    const double den = fm - 2.0 * f0 + fp;
    const double d = den < 0.0 ? 0.5 * (fm - fp) / den : 0.0;
    return fmin(0.5, fmax(-0.5, d));
}

__host__ __device__ inline size_t refine_lds(int Z) { return sizeof(double) * (size_t(Z) * Z + 5) + 16 * 9 * size_t(Z); }

template <class T>
__global__ void __launch_bounds__(64) k_refine(const T *surf, int ch, int cw, int Z, int ovs, double ratio,
                                               const int32_t *zidx, const int32_t *start, int hs_d, int hs_a,
                                               const cx<double> *tw, double *out) {
    extern __shared__ __align__(16) unsigned char smem[];
    cx<double> *D = reinterpret_cast<cx<double> *>(smem);    // [2 axes][3][Z]
    cx<double> *B = D + 6 * Z;                                 // [3][Z]
    double *a = reinterpret_cast<double *>(B + 3 * Z);         // [Z][Z]
    double *s5 = a + Z * Z;
    const int w = blockIdx.x, lane = threadIdx.x;
    const int N = Z * ovs;
    const T *src = surf + size_t(w) * ch * cw;
    for (int i = lane; i < Z * Z; i += 64) a[i] = double(src[(i / Z) * cw + i % Z]);
    const int px = zidx[2 * w], py = zidx[2 * w + 1];
    // D(r, p - 1 + t) for both axes (e^{+2 pi i k/N} = conj(tw[k])); the table
    // index f*m mod N advances by m from f = -Z/2 on
    for (int i = lane; i < 6 * Z; i += 64) {
        const int axis = i / (3 * Z), j = i - axis * 3 * Z, t = j / Z, r = j - t * Z;
        const int X = (axis ? py : px) - 1 + t;
        const int m = ((r * ovs - X) % N + N) % N;
        int k = int((int64_t(N - Z / 2) * m) % N);
        cx<double> acc{0.0, 0.0};
        for (int f = 0; f < Z; ++f) {
            acc.x += tw[k].x;
            acc.y -= tw[k].y;
            k += m;
            if (k >= N) k -= N;
        }
        D[i] = acc;
    }
    __syncthreads();
    // B[t][r] = sum_c a[r][c] D(c, py - 1 + t)
    for (int i = lane; i < 3 * Z; i += 64) {
        const int t = i / Z, r = i - t * Z;
        const cx<double> *dy = D + (3 + t) * Z;
        cx<double> acc{0.0, 0.0};
        for (int c = 0; c < Z; ++c) {
            acc.x += a[r * Z + c] * dy[c].x;
            acc.y += a[r * Z + c] * dy[c].y;
        }
        B[i] = acc;
    }
    __syncthreads();
    // the centre and its four axis neighbours: (tx, ty) = (1,1) (0,1) (2,1) (1,0) (1,2)
    if (lane < 5) {
        const int tx = lane == 1 ? 0 : (lane == 2 ? 2 : 1);
        const int ty = lane == 3 ? 0 : (lane == 4 ? 2 : 1);
        double acc = 0.0;
        for (int r = 0; r < Z; ++r) {
            const cx<double> d = D[tx * Z + r], b = B[ty * Z + r];
            acc += d.x * b.x - d.y * b.y;
        }
        s5[lane] = acc / (double(Z) * Z);
    }
    __syncthreads();
    if (lane == 0) {
        const double dx = parabola_vertex(s5[1], s5[0], s5[2]);
        const double dy = parabola_vertex(s5[3], s5[0], s5[4]);
        out[2 * w] = (double(px) + dx) * ratio + double(start[2 * w]) - double(hs_d);
        out[2 * w + 1] = (double(py) + dy) * ratio + double(start[2 * w + 1]) - double(hs_a);
    }
}

}  // namespace ampcor
