// Fused per-window pipeline: the fast path for square windows of 32 or 64
// pixels with 2x raw oversampling, a 16x16 zoom surface and 16x FFT surface
// oversampling (the reference's defaults). Three kernels per chunk of windows:
// k_coarse -> k_oversample (reference chip; search zoom chip) -> k_fine.
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <vector>
#include "engine.h"
#include "fused_kernels.h"
#include "stage_kernels.h"

namespace ampcor {
namespace fused {

__host__ __device__ inline int al16(int x) { return (x + 15) & ~15; }

// ---------------------------------------------------------------------------
// coarse stage (cuAmpcorChunk.cpp:18-106), one workgroup per window
// ---------------------------------------------------------------------------
// template rows per spectrum chunk of the coarse stage: the window-64 layout
// then fits two workgroups per CU (77 KB of LDS each)
__host__ __device__ constexpr int coarse_tc(int W) { return W > 32 ? 32 : W; }

struct CoarseLds {
    int spec, nrm, red, tw, total, y_off, corr_off;   // bytes (spec, nrm, red, tw, total); cf units (y_off, corr_off)
    __host__ __device__ CoarseLds(int W, int L, int SH, int NLh, int NLw) {
        const int P = L + 1, H = L / 2;
        const int nS = (SH + 1) / 2, nTc = coarse_tc(W) / 2;
        y_off = (NLh * H + 1) & ~1;                   // lag spectra Q, then the inverse pairs Y
        const int npair = (NLh + 1) / 2;
        corr_off = y_off + npair * P;                 // then the normalized surface
        int cf_units = (nS + nTc) * P;                // search + one template chunk
        const int need = corr_off + (NLh * NLw + 1) / 2;
        if (need > cf_units) cf_units = need;
        spec = 0;
        nrm = al16(spec + cf_units * 8);
        red = al16(nrm + NLh * NLw * 4);
        tw = red + 32 * 8;                            // the row transform's twiddles
        total = tw + L * 8;
    }
};

struct CoarseArgs {
    const void *ref; int64_t ref_pitch;
    const void *sec; int64_t sec_pitch;
    int mode;                      // 0 complex, 1 |z| chips, 2 real image
    const int32_t *ref_d, *ref_a, *sec_d, *sec_a;
    int hs_d, hs_a, hz, patch_h, patch_w;
    int32_t *maxloc; float *maxval; int32_t *start, *shift; float *snr, *cov;
    float *margin;                 // [n] top-1 minus top-2 of the normalized surface, or null
};

template <int W, int R1, int R2, int K>
__global__ void __launch_bounds__(NT, (K >= 9 ? 2 : 4)) k_coarse(CoarseArgs a, const cf *tw_g, int w0) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int L = R1 * R2, H = L / 2, P = L + 1, TC = coarse_tc(W), nTc = TC / 2;
    static_assert(W % TC == 0, "template chunks cover the window");
    // mac_rows reads up to K - 2 rows past the search (only discarded lags use them):
    // they land in the template chunk right after the search pairs
    static_assert(nTc >= K / 2 + 1, "the template chunk follows the search rows");
    static_assert(2 * NORM_BLOCKS * L <= nTc * P, "window-sum scratch fits the template chunk region");
    const int SH = W + 2 * a.hs_d, SW = W + 2 * a.hs_a;
    const int NLh = SH - W + 1, NLw = SW - W + 1;
    const int nS = (SH + 1) / 2;
    const CoarseLds lay(W, L, SH, NLh, NLw);
    cf *S = reinterpret_cast<cf *>(smem + lay.spec);
    cf *Tb = S + nS * P;
    float *nrm = reinterpret_cast<float *>(smem + lay.nrm);
    float *corr = reinterpret_cast<float *>(S + lay.corr_off);
    double *red = reinterpret_cast<double *>(smem + lay.red);
    const int win = w0 + xcd_window(blockIdx.x, gridDim.x);
    // twiddles read from LDS (the transforms' table loads wait one at a time)
    cf *twl = reinterpret_cast<cf *>(smem + lay.tw);
    for (int i = threadIdx.x; i < L; i += NT) twl[i] = tw_g[i];
    const cf *tw = twl;

    // search amplitudes, row pairs packed as (row 2j, row 2j+1), and the
    // template's FP64 sum and sum of squares, gathers unrolled for memory parallelism
    const int64_t rd = a.ref_d[win], ra = a.ref_a[win];
    const int64_t sd = a.sec_d[win], sa = a.sec_a[win];
    // template pixels are requested first and held raw in registers, so that
    // their loads overlap the search chip's
    constexpr int TI = W * W / NT;
    static_assert(W * W % NT == 0, "template pixels divide evenly over the workgroup");
    static_assert(NT % L == 0 && NT % W == 0, "fixed column per thread in the gathers");
    float2 traw[TI];
#pragma unroll
    for (int k = 0; k < TI; ++k) {
        const int i = threadIdx.x + k * NT, r = i / W, x = i - r * W;
        traw[k] = pixel_raw(a.ref, a.mode, (rd + r) * a.ref_pitch + ra + x);
    }
    {
        // every thread keeps one column x of the search chip and walks its row
        // pairs with a running pointer
        const int x = threadIdx.x % L;
        const int j0 = threadIdx.x / L;
        constexpr int JS = NT / L;
        const bool in = x < SW;
        const int64_t step = int64_t(2 * JS) * a.sec_pitch;
        int64_t off = (sd + 2 * j0) * a.sec_pitch + sa + (in ? x : 0);
        auto put = [&](auto amp_at) {
#pragma unroll 4
            for (int j = j0; j < nS; j += JS, off += step) {
                cf v = {0.f, 0.f};
                if (in) {
                    v.x = amp_at(off);
                    if (2 * j + 1 < SH) v.y = amp_at(off + a.sec_pitch);
                }
                S[j * P + x] = v;
            }
        };
        if (a.mode == 2) {
            const float *img = static_cast<const float *>(a.sec);
            put([&](int64_t o) { return fabsf(img[o]); });
        } else {
            const float2 *img = static_cast<const float2 *>(a.sec);
            put([&](int64_t o) {
                const float2 z = img[o];
                return fast_sqrt(z.x * z.x + z.y * z.y);
            });
        }
    }
    // template amplitudes, held for the chunk loads below
    float tamp[TI];
    double sum = 0.0, sum2 = 0.0;
#pragma unroll
    for (int k = 0; k < TI; ++k) {
        tamp[k] = raw_amp(traw[k], a.mode);
        const double v = tamp[k];
        sum += v;
        sum2 += v * v;
    }
    sum = block_sum(sum, red);
    sum2 = block_sum(sum2, red);
    const double mean = sum / double(W * W);
    const double t2 = fmax(sum2 - sum * sum / double(W * W), 0.0);   // sum of squared deviations
    // window sums from 16-row block sums down the whole search (the template
    // chunk's region, not loaded yet, holds them: 2 * ceil(SH/16) * SW doubles)
    static_assert(W % NORM_BLK == 0, "template height a multiple of the norm block");
    static_assert(2 * ((L + NORM_BLK - 1) / NORM_BLK) * L <= nTc * P, "block sums fit the template chunk region");
    window_norms_blk<(L + 63) / 64, NT, W % 64 == 0 ? W / 64 : 0>(
        [&](int r, int c) -> double { return reinterpret_cast<const float *>(S)[2 * ((r >> 1) * P + c) + (r & 1)]; },
        SH, SW, W, W, NLh, NLw, t2, nrm, reinterpret_cast<double *>(Tb));
    __syncthreads();

    const int f = threadIdx.x % H, blk = threadIdx.x / H, u0 = blk * K;
    const bool active = (blk < NT / H) && (u0 < NLh);
    cf q[K];
#pragma unroll
    for (int k = 0; k < K; ++k) q[k] = {0.f, 0.f};
    #pragma unroll 1
    for (int c0 = 0; c0 < W; c0 += TC) {
        // template rows [c0, c0 + TC) minus the FP64 mean, packed in pairs, from
        // the amplitudes held in registers; zero beyond column W
#pragma unroll
        for (int k = 0; k < TI; ++k) {
            const int i = threadIdx.x + k * NT, r = i / W, x = i - r * W;
            if (r >= c0 && r < c0 + TC) {
                float *slot = reinterpret_cast<float *>(Tb + ((r - c0) >> 1) * P + x) + (r & 1);
                *slot = float(double(tamp[k]) - mean);
            }
        }
        for (int i = threadIdx.x; i < nTc * (L - W); i += NT) {
            const int j = i / (L - W), x = W + i - j * (L - W);
            Tb[j * P + x] = cf{0.f, 0.f};
        }
        __syncthreads();
        if (c0 == 0) {
            // row spectra of the search and of the first template chunk in one pass
            fft_nat2dr<float, -1, R1, R2, NT>(S, nS + nTc, P, 1, tw);
            split_pairs_inplace<R1, R2>(S, nS + nTc, P);
        } else {
            fft_nat2dr<float, -1, R1, R2, NT>(Tb, nTc, P, 1, tw);
            split_pairs_inplace<R1, R2>(Tb, nTc, P);
        }
        if (active) mac_rows<H, K>(q, S, P, Tb, P, c0, c0, c0 + TC, u0, SH, f);
        __syncthreads();
    }
    cf *Qs = S;
    if (active) {
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (u0 + k < NLh) Qs[(u0 + k) * H + f] = q[k];
    }
    __syncthreads();
    inverse_lags<R1, R2>(Qs, NLh, NLw, S + lay.y_off, P, tw, nrm, corr);

    const Top2<float> top = block_argmax(corr, NLh * NLw, red);
    const int idx = top.index();
    const float val = top.v1;
    const int px = idx / NLw, py = idx - px * NLw;
    if (threadIdx.x < 64)      // wave 0: SNR patch sum over its lanes, covariance on lane 0
        window_stats(corr, NLh, NLw, px, py, double(val), W * W, a.patch_h, a.patch_w, a.cov + 3 * win,
                     a.snr + win);
    if (threadIdx.x == 0) {
        if (a.margin) a.margin[win] = top.margin();
        int s0, h0, s1, h1;
        adjust_axis(px, a.hs_d, a.hz, s0, h0);
        adjust_axis(py, a.hs_a, a.hz, s1, h1);
        a.maxloc[2 * win] = px; a.maxloc[2 * win + 1] = py;
        a.maxval[win] = val;
        a.start[2 * win] = s0; a.start[2 * win + 1] = s1;
        a.shift[2 * win] = h0; a.shift[2 * win + 1] = h1;
    }
}

// ---------------------------------------------------------------------------
// x2 chip oversampling (cuOverSampler.cpp:72-78, cuArraysPadding.cu:59-103,
// cuDeramp.cu:67-148): per axis, odd samples = (1/N) FFT_-(FFT_+(x) e^{-i pi k_s/N}),
// even samples = x. Output pixel (2r+p, 2c+q) comes from:
//   p=0,q=0: X;  p=1,q=0: columns of X interpolated;
//   p=0,q=1: Y = rows of X interpolated;  p=1,q=1: columns of Y interpolated.
// Only one N x N complex chip lives in LDS (so several workgroups share a CU):
// the chip is gathered (and deramped) twice, once for the X column pass and
// once for the Y passes. |.| goes to HBM.
// ---------------------------------------------------------------------------
struct OsArgs {
    const void *img; int64_t pitch; int mode;
    const int32_t *sd, *sa, *start;   // start may be null
    int deramp;
    const cf *hsh;                    // [N] e^{-i pi k_s/N} / N
    float *out;                       // [n][4][N][N]: pixel (2r+p, 2c+q) in plane 2p+q at (r, c)
    double *mean;                     // [n][2]: mean and sum of squared deviations of |.| (MEAN kernels)
};

// chip (pitch N+1), shift table, deramp row/column rotation factors, reduction scratch, twiddles
__host__ __device__ inline int os_lds_bytes(int N) { return al16(N * (N + 1) * 8 + N * 8) + 2 * N * 8 + 48 * 8 + N * 8; }

template <int R1, int R2, bool MEAN, int NTH = NT>
__global__ void __launch_bounds__(NTH, (R1 * R2 == 64 && MEAN) ? 8 : 1) k_oversample(OsArgs a, const cf *tw_g, int w0) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int N = R1 * R2, PC = N + 1;
    static_assert(2 * N <= NTH, "one thread per deramp row and column factor");
    cf *C = reinterpret_cast<cf *>(smem);
    cf *hsh = C + N * PC;
    cf *rowf = reinterpret_cast<cf *>(smem + al16(N * PC * 8 + N * 8));   // e^{i r phy}
    cf *colf = rowf + N;                                                  // e^{i c phx}
    double *red = reinterpret_cast<double *>(colf + N);
    cf *twl = reinterpret_cast<cf *>(red + 48);                           // twiddles, read from LDS
    const cf *tw = twl;
    const int win = w0 + xcd_window(blockIdx.x, gridDim.x);
    int64_t r0 = a.sd[win], c0 = a.sa[win];
    if (a.start) { r0 += a.start[2 * win]; c0 += a.start[2 * win + 1]; }
    // deramp of pixel (r, c) (float2's operator*= (float2.h:71-75): the imaginary part
    // uses the new real part); the phases and their sines and cosines are FP64, the
    // factors and the rotation FP32
    auto deramp_px = [&](cf z, int r, int c) -> cf {
        const cf f = cmul(colf[c], rowf[r]);
        const float nx = z.x * f.x - z.y * f.y;
        return {nx, z.y * f.x + nx * f.y};
    };
    // chip into LDS; rot: deramped as it loads (the factors are in place)
    auto gather = [&](bool rot) {
#pragma unroll 4
        for (int i = threadIdx.x; i < N * N; i += NTH) {
            const int r = i / N, c = i - r * N;
            const cf z = pixel_cx(a.img, a.mode, (r0 + r) * a.pitch + c0 + c);
            C[r * PC + c] = rot ? deramp_px(z, r, c) : z;
        }
    };
    double *psum = red + 16;             // [NTH/64][2] per-wave FP64 sums of |.| and |.|^2 (MEAN)
    static_assert(16 + 2 * (NTH / 64) <= 48, "wave slots fit the reduction scratch");
    if (MEAN && threadIdx.x < 2 * (NTH / 64)) psum[threadIdx.x] = 0.0;
    gather(false);
    for (int i = threadIdx.x; i < N; i += NTH) {
        hsh[i] = a.hsh[i];
        twl[i] = tw_g[i];
    }
    __syncthreads();
    if (a.deramp == 1) {
        double ayr = 0, ayi = 0, axr = 0, axi = 0;
        for (int i = threadIdx.x; i < N * N; i += NTH) {
            const int r = i / N, c = i - r * N;
            const cf z = C[r * PC + c];
            const double zr = z.x, zi = z.y;
            if (c < N - 1) {
                const cf b = C[r * PC + c + 1];
                ayr += zr * b.x + zi * b.y; ayi += zi * b.x - zr * b.y;
            }
            if (r < N - 1) {
                const cf b = C[(r + 1) * PC + c];
                axr += zr * b.x + zi * b.y; axi += zi * b.x - zr * b.y;
            }
        }
        ayr = block_sum<NTH>(ayr, red); ayi = block_sum<NTH>(ayi, red);
        axr = block_sum<NTH>(axr, red); axi = block_sum<NTH>(axi, red);
        const double phy = atan2(ayi, ayr), phx = atan2(axi, axr);
        // the phase c*phx + r*phy is separable: one FP64 sincos per row and per column
        if (threadIdx.x < 2 * N) {
            const int k = threadIdx.x < N ? threadIdx.x : threadIdx.x - N;
            double sn, cs;
            sincos(double(k) * (threadIdx.x < N ? phy : phx), &sn, &cs);
            (threadIdx.x < N ? rowf : colf)[k] = {float(cs), float(sn)};
        }
        __syncthreads();
    }
    float *out = a.out + size_t(win) * 4 * N * N;
    // FP64 sums of |.| and |.|^2: each emit reduces its own over the wave and
    // adds them to the wave's LDS slot, so no FP64 accumulator stays live
    // through the transforms (the reference-chip kernel is held to 64 VGPRs)
    static_assert(N % 4 == 0, "four output columns per store");
    // |C| -> output pixels (2r+p, 2c+q), four columns per thread; rot: C is deramped
    // in place first (each thread rewrites the pixels it reads)
    auto emit = [&](int p, int q, bool rot) {
        double s1 = 0.0, s2 = 0.0;
        float *plane = out + size_t(2 * p + q) * N * N;           // whole lines per phase
        for (int i = threadIdx.x; i < N * N / 4; i += NTH) {
            const int r = (4 * i) / N, c = 4 * i - r * N;
            float amp[4];
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                cf v = C[r * PC + c + m];
                if (rot) {
                    v = deramp_px(v, r, c + m);
                    C[r * PC + c + m] = v;
                }
                amp[m] = fast_sqrt(v.x * v.x + v.y * v.y);
                if (MEAN) {
                    s1 += amp[m];
                    s2 += double(amp[m]) * amp[m];
                }
            }
            *reinterpret_cast<float4 *>(plane + r * N + c) = make_float4(amp[0], amp[1], amp[2], amp[3]);
        }
        if (MEAN) {
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                s1 += __shfl_xor(s1, o);
                s2 += __shfl_xor(s2, o);
            }
            if ((threadIdx.x & 63) == 0) {
                psum[2 * (threadIdx.x >> 6)] += s1;
                psum[2 * (threadIdx.x >> 6) + 1] += s2;
            }
        }
    };
    // the half-sample shift multiplies each bin as the forward transform stores it,
    // except in the 64-VGPR reference kernel, where its own pass spills less
    constexpr bool FOLD = !(N == 64 && MEAN);
    auto interp_cols = [&]() {          // C <- C at half-row positions (every column)
        fft_nat2dr<float, +1, R1, R2, NTH>(C, N, 1, PC, tw, FOLD ? hsh : nullptr);
        if (!FOLD) {
            for (int i = threadIdx.x; i < N * N; i += NTH) {
                const int k = i / N, c = i - k * N;
                cf &z = C[dr_pos<R1, R2>(k) * PC + c];
                z = cmul(z, hsh[k]);
            }
            __syncthreads();
        }
        fft_dr2nat<float, -1, R1, R2, NTH>(C, N, 1, PC, tw);
    };
    // X: even rows, even columns; then its column interpolation: odd rows, even columns
    emit(0, 0, a.deramp == 1);
    __syncthreads();
    interp_cols();
    emit(1, 0, false);
    __syncthreads();
    // Y = X interpolated along rows: even rows, odd columns; then its columns: odd rows, odd columns
    gather(a.deramp == 1);
    __syncthreads();
    fft_nat2dr<float, +1, R1, R2, NTH>(C, N, PC, 1, tw, FOLD ? hsh : nullptr);
    if (!FOLD) {
        for (int i = threadIdx.x; i < N * N; i += NTH) {
            const int r = i / N, k = i - r * N;
            cf &z = C[r * PC + dr_pos<R1, R2>(k)];
            z = cmul(z, hsh[k]);
        }
        __syncthreads();
    }
    fft_dr2nat<float, -1, R1, R2, NTH>(C, N, PC, 1, tw);
    emit(0, 1, false);
    __syncthreads();
    interp_cols();
    emit(1, 1, false);
    if (MEAN) {
        // FP64 mean and sum of squared deviations (sum2 - sum^2 / count)
        __syncthreads();
        double sum = 0.0, sum2 = 0.0;
#pragma unroll
        for (int w = 0; w < NTH / 64; ++w) {
            sum += psum[2 * w];
            sum2 += psum[2 * w + 1];
        }
        const double cnt = double(4 * N * N);
        if (threadIdx.x == 0) {
            a.mean[2 * win] = sum / cnt;
            a.mean[2 * win + 1] = fmax(sum2 - sum * sum / cnt, 0.0);
        }
    }
}

// ---------------------------------------------------------------------------
// fine stage (cuAmpcorChunk.cpp:148-207): oversampled correlation, normalization,
// trim to 16x16, 16x surface interpolation (16-sparse spectrum evaluated with
// 16-point DFTs), argmax with the reference tie rule, sub-pixel offset. The
// whole template's row spectra sit in LDS next to the search's (for window 64:
// 72 + 64 packed signals of 144 bins, 158 KB), so both are transformed and
// split together and the per-bin sums run over all template rows in one pass.
// ---------------------------------------------------------------------------

struct FineLds {
    int spec, tchunk, nrm, corr, red, total, y_off;
    __host__ __device__ FineLds(int P, int NL, int T) {
        const int L = P, PP = L + 1, H = L / 2;
        const int nS = P / 2;
        y_off = (NL * H + 1) & ~1;
        int s_units = nS * PP;
        const int npair = (NL + 1) / 2;
        if (y_off + npair * PP > s_units) s_units = y_off + npair * PP;
        if (s_units < 16 * 256) s_units = 16 * 256;            // surface G[16][256]
        int t_units = (T / 2) * PP;                           // template row spectra
        if (t_units < 2 * 256) t_units = 2 * 256;             // 16x16 surface + twiddles
        spec = 0;
        tchunk = al16(s_units * 8);
        nrm = al16(tchunk + t_units * 8);
        corr = al16(nrm + NL * NL * 4);
        red = al16(corr + NL * NL * 4);
        total = red + 64 * 8;          // block sums + workgroup argmax scratch (up to 1024 threads)
    }
};

struct FineArgs {
    const float *tamp; const double *tmean;   // [n][4][T/2][T/2] planar, [n][2] (mean, sum of squared deviations)
    const float *samp;                         // [n][4][P/2][P/2] planar (k_oversample's layout)
    const int32_t *start;
    int hs_d, hs_a;
    float ratio;                               // 1/(surf_ovs*raw_ovs)
    const cf *tw256;                           // e^{-2 pi i k/256}
    int32_t *zoom_idx; float *offset;
    float *margin;                             // [n] top-1 minus top-2 of the surface, or null
    float *surf;                               // [n][16][16] trimmed surface for the FP64 refinement, or null
};

// ---------------------------------------------------------------------------
// Surface stage of the fine kernels: trimmed 16x16 normalized surface (corr,
// row pitch NL) -> 16x16 +i DFT -> 16-sparse spectrum G[16][256] -> column
// bounds -> evaluation of the kept columns with 16-point DFTs -> argmax with the
// reference tie rule -> zoom index, offset, margin. LDS: G (16*256 cf), abuf
// (512 cf: the 16x16 spectrum and 256 twiddles), cols (256 ints), red (64 doubles).
// ---------------------------------------------------------------------------
template <int NTH>
__device__ __forceinline__ void fine_surface(const FineArgs &a, const float *corr, int NL, cf *G, cf *abuf, int *cols,
                                             double *red, int win) {
    constexpr int Z = 16, SO = 256;
    // trimmed 16x16 surface -> 2-D +i DFT / 256 (cuOverSampler.cpp:141-145)
    cf *a16 = abuf;          // [16][16]
    cf *tws = abuf + 256;    // e^{-2 pi i k/256}
    for (int i = threadIdx.x; i < Z * Z; i += NTH) {
        const int r = i / Z, c = i - r * Z;
        a16[i] = {corr[r * NL + c] * (1.0f / 256.0f), 0.f};
        if (a.surf) a.surf[size_t(win) * Z * Z + i] = corr[r * NL + c];
    }
    for (int i = threadIdx.x; i < SO; i += NTH) tws[i] = a.tw256[i];
    __syncthreads();
    if (threadIdx.x < 16) {
        cf x[16];
#pragma unroll
        for (int c = 0; c < 16; ++c) x[c] = a16[threadIdx.x * 16 + c];
        dft16<float, +1>(x);
#pragma unroll
        for (int c = 0; c < 16; ++c) a16[threadIdx.x * 16 + c] = x[c];
    }
    __syncthreads();
    if (threadIdx.x < 16) {
        cf x[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) x[r] = a16[r * 16 + threadIdx.x];
        dft16<float, +1>(x);
#pragma unroll
        for (int r = 0; r < 16; ++r) a16[r * 16 + threadIdx.x] = x[r];
    }
    __syncthreads();
    // G[k1][16c+dd] = DFT16_c( A[k1][k2] e^{-2 pi i dd ks2/256} ), ks2 = signed k2
    if (threadIdx.x < 256) {
        const int k1 = threadIdx.x >> 4, dd = threadIdx.x & 15;
        cf x[16];
#pragma unroll
        for (int k2 = 0; k2 < 16; ++k2) {
            const int ks = k2 < 8 ? k2 : k2 - 16;
            x[k2] = cmul(a16[k1 * 16 + k2], tws[(dd * ks + 4096) & 255]);
        }
        dft16<float, -1>(x);
#pragma unroll
        for (int c = 0; c < 16; ++c) G[k1 * SO + 16 * c + dd] = x[c];
    }
    __syncthreads();
    // Column pruning. Every value of surface column q is at most
    // B_q = sum_k1 |G[k1][q]|, and the surface passes through the trimmed
    // samples, so its maximum is at least their largest value amax. A column with
    // B_q (1 + 2^-10) + 2^-11 |a|max < amax can hold neither the argmax nor any
    // value within 2^-12 |a|max of it (the slack covers FP32 rounding many times
    // over): it is skipped, which leaves the argmax, the tie rule and every
    // margin below 2^-12 |a|max unchanged. NaN bounds keep their column.
    static_assert(NTH >= SO, "one thread per surface column for the bounds");
    float amax = -FLT_MAX, amag = 0.f;
    for (int i = threadIdx.x & 63; i < Z * Z; i += 64) {
        const float v = corr[(i / Z) * NL + (i % Z)];
        amax = fmaxf(amax, v);
        amag = fmaxf(amag, fabsf(v));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        amax = fmaxf(amax, __shfl_xor(amax, o));
        amag = fmaxf(amag, __shfl_xor(amag, o));
    }
    uint64_t *masks = reinterpret_cast<uint64_t *>(red + 48);   // ballots of the SO/64 bound waves
    bool keep = false;
    if (threadIdx.x < SO) {
        float bq = 0.f;
#pragma unroll
        for (int k1 = 0; k1 < 16; ++k1) {
            const cf g = G[k1 * SO + threadIdx.x];
            bq += fast_sqrt(g.x * g.x + g.y * g.y);
        }
        keep = !(bq * (1.0f + 0x1p-10f) + 0x1p-11f * amag < amax);
    }
    const uint64_t wm = __ballot(keep);
    if (threadIdx.x < SO && (threadIdx.x & 63) == 0) masks[threadIdx.x >> 6] = wm;
    __syncthreads();
    int ncols = 0, before = 0;
#pragma unroll
    for (int w = 0; w < SO / 64; ++w) {
        const int c = __popcll(masks[w]);
        if (w < (threadIdx.x >> 6)) before += c;
        ncols += c;
    }
    if (keep) cols[before + __popcll(wm & ((uint64_t(1) << (threadIdx.x & 63)) - 1))] = threadIdx.x;
    __syncthreads();
    // out[16aa+b][q] = Re DFT16_aa( G[k1][q] e^{-2 pi i b ks1/256} ) over the kept
    // columns q; task t = (b, column list entry); running argmax
    Top2<float> top;
    #pragma unroll 1
    for (int t = threadIdx.x; t < 16 * ncols; t += NTH) {
        const int b = t / ncols, qq = cols[t - b * ncols];
        // e^{-2 pi i b k/256} for k = 1..8 (negative bins use the conjugate)
        cf w[9];
#pragma unroll
        for (int k = 1; k <= 8; ++k) w[k] = tws[(b * k) & 255];
        cf x[16];
        x[0] = G[qq];
#pragma unroll
        for (int k1 = 1; k1 < 16; ++k1) {
            const cf tk = k1 < 8 ? w[k1] : conj(w[16 - k1]);
            x[k1] = cmul(G[k1 * SO + qq], tk);
        }
        dft16<float, -1>(x);
#pragma unroll
        for (int aa = 0; aa < 16; ++aa) top.add_idx(x[aa].x, uint32_t((16 * aa + b) * SO + qq));
    }
    // red: [0, 16) block sums, then NTH/64 top-1 values, runner-up values and keys
    top.reduce<NTH>(reinterpret_cast<float *>(red + 16), reinterpret_cast<float *>(red + 16) + 16,
                    reinterpret_cast<uint64_t *>(red + 32));
    if (threadIdx.x == 0) {
        const int idx = top.index();
        const int px = idx / SO, py = idx % SO;
        a.zoom_idx[2 * win] = px;
        a.zoom_idx[2 * win + 1] = py;
        a.offset[2 * win] = a.ratio * float(px) + float(a.start[2 * win]) - float(a.hs_d);
        a.offset[2 * win + 1] = a.ratio * float(py) + float(a.start[2 * win + 1]) - float(a.hs_a);
        if (a.margin) a.margin[win] = top.margin();
    }
}

template <int T, int R1, int R2, int K, int NTH>
__global__ void __launch_bounds__(NTH) k_fine(FineArgs a, const cf *tw, int w0) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int P = R1 * R2, L = P, PP = L + 1, H = L / 2, nS = P / 2;
    constexpr int NL = P - T + 1;      // 17
    constexpr int Z = NL - 1;          // 16 (trimmed surface)
    constexpr int SO = 256;            // oversampled surface side
    static_assert(Z == 16, "fused surface stage assumes a 16x16 zoom surface");
    // mac_rows reads up to K - 2 rows past the search; the template spectra follow it
    static_assert(T / 2 >= K / 2 + 1, "the template region follows the search rows");
    const FineLds lay(P, NL, T);
    cf *S = reinterpret_cast<cf *>(smem + lay.spec);
    cf *Tb = reinterpret_cast<cf *>(smem + lay.tchunk);
    float *nrm = reinterpret_cast<float *>(smem + lay.nrm);
    float *corr = reinterpret_cast<float *>(smem + lay.corr);
    double *red = reinterpret_cast<double *>(smem + lay.red);
    const int win = w0 + xcd_window(blockIdx.x, gridDim.x);
    const float *tp = a.tamp + size_t(win) * T * T;
    const float *sg = a.samp + size_t(win) * P * P;
    const double mean = a.tmean[2 * win], t2 = a.tmean[2 * win + 1];
    // the search chip's loads are all issued before any of its values is stored
    constexpr int SN = nS * L, SI = (SN + NTH - 1) / NTH;
    constexpr int TN = (T / 2) * L, TI = (TN + NTH - 1) / NTH;
    cf sv[SI], tv[TI];
#pragma unroll
    for (int k = 0; k < SI; ++k) {
        const int i = threadIdx.x + k * NTH, j = i / L, x = i - j * L;
        // planar oversampled chip: row 2j + p, column x -> plane 2p + (x & 1), (j, x >> 1)
        const float *s0 = sg + size_t(x & 1) * (P / 2) * (P / 2) + j * (P / 2) + (x >> 1);
        if (SN % NTH == 0 || i < SN) sv[k] = {s0[0], s0[size_t(2) * (P / 2) * (P / 2)]};
    }
#pragma unroll
    for (int k = 0; k < SI; ++k) {
        const int i = threadIdx.x + k * NTH, j = i / L, x = i - j * L;
        if (SN % NTH == 0 || i < SN) S[j * PP + x] = sv[k];
    }
    __syncthreads();
    // the template's loads go out once the search chip has arrived (they would
    // only delay it) and stay in flight through the window sums
#pragma unroll
    for (int k = 0; k < TI; ++k) {
        const int i = threadIdx.x + k * NTH, j = i / L, x = i - j * L;
        const float *t0 = tp + size_t(x & 1) * (T / 2) * (T / 2) + j * (T / 2) + (x >> 1);
        tv[k] = {0.f, 0.f};
        if ((TN % NTH == 0 || i < TN) && x < T) tv[k] = {t0[0], t0[size_t(2) * (T / 2) * (T / 2)]};
    }
    window_norms<(P + 63) / 64, NTH, T % 64 == 0 ? T / 64 : 0>(
        [&](int r, int c) -> double { return reinterpret_cast<const float *>(S)[2 * ((r >> 1) * PP + c) + (r & 1)]; },
        P, T, T, NL, NL, t2, nrm, reinterpret_cast<double *>(Tb));   // template not loaded yet
    __syncthreads();
    // template rows minus the FP64 mean, row pairs packed, zero beyond column T
#pragma unroll
    for (int k = 0; k < TI; ++k) {
        const int i = threadIdx.x + k * NTH, j = i / L, x = i - j * L;
        if (TN % NTH == 0 || i < TN)
            Tb[j * PP + x] = x < T ? cf{float(double(tv[k].x) - mean), float(double(tv[k].y) - mean)} : cf{0.f, 0.f};
    }
    __syncthreads();
    if (lay.tchunk == nS * PP * 8) {
        // search and template spectra are one evenly spaced set of signals
        fft_nat2dr<float, -1, R1, R2, NTH>(S, nS + T / 2, PP, 1, tw);
        split_pairs_inplace<R1, R2, NTH>(S, nS + T / 2, PP);
    } else {
        fft_nat2dr<float, -1, R1, R2, NTH>(S, nS, PP, 1, tw);
        fft_nat2dr<float, -1, R1, R2, NTH>(Tb, T / 2, PP, 1, tw);
        split_pairs_inplace<R1, R2, NTH>(S, nS, PP);
        split_pairs_inplace<R1, R2, NTH>(Tb, T / 2, PP);
    }

    const int f = threadIdx.x % H, blk = threadIdx.x / H, u0 = blk * K;
    const bool active = (blk < NTH / H) && (u0 < NL);
    cf q[K];
#pragma unroll
    for (int k = 0; k < K; ++k) q[k] = {0.f, 0.f};
    if (active) mac_rows<H, K>(q, S, PP, Tb, PP, 0, 0, T, u0, P, f);
    __syncthreads();
    cf *Qs = S;
    if (active) {
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (u0 + k < NL) Qs[(u0 + k) * H + f] = q[k];
    }
    __syncthreads();
    inverse_lags<R1, R2, NTH>(Qs, NL, NL, S + lay.y_off, PP, tw, nrm, corr);

    fine_surface<NTH>(a, corr, NL, S, Tb, reinterpret_cast<int *>(nrm), red, win);
}

// ---------------------------------------------------------------------------
// Fine stage, streamed over template rows: the computation of k_fine with the
// template taken TC rows at a time and only the search rows those rows meet in
// LDS (TC + NL - 1 rows; the last NL - 1 carry over to the next chunk), so a
// workgroup needs ~68 KB of LDS and two share a CU. The window sums of the
// normalization come from FP64 column sums accumulated as the search rows
// arrive, with the first and last NL - 1 rows kept for the slides.
// ---------------------------------------------------------------------------
struct Fine2Lds {
    int sr, tb, save, nrm, corr, red, tw, total;   // bytes
    __host__ __device__ Fine2Lds(int P, int NL, int TC) {
        const int PP = P + 1;
        const int srp = (TC + NL - 1 + 1) / 2, tcp = TC / 2;
        sr = 0;
        tb = sr + srp * PP * 8;                    // template pairs directly after the search pairs
        save = tb + tcp * PP * 8;
        nrm = al16(save + 2 * (NL - 1) * P * 4);
        corr = al16(nrm + NL * NL * 4);
        red = al16(corr + NL * NL * 4);
        tw = red + 64 * 8;                         // the row transform's twiddles
        total = tw + P * 8;
    }
};

template <int T, int R1, int R2, int K, int NTH, int TC>
__global__ void __launch_bounds__(NTH, 2) k_fine2(FineArgs a, const cf *tw_g, int w0) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int P = R1 * R2, L = P, PP = L + 1, H = L / 2;
    constexpr int NL = P - T + 1;            // lags per axis (17)
    constexpr int NCH = T / TC;              // template chunks
    constexpr int SROWS = TC + NL - 1;       // search rows a chunk meets
    constexpr int SRP = (SROWS + 1) / 2;     // as row pairs
    constexpr int KEEP = (NL - 1) / 2;       // pairs carried to the next chunk
    constexpr int TCP = TC / 2;
    constexpr int NS = NL - 1;               // rows kept at each end for the slides
    constexpr int Q = P / 2, TQ = T / 2;     // polyphase plane sides
    static_assert(T % TC == 0 && TC % 2 == 0 && NS % 2 == 0 && NS == 16, "even chunks, 16x16 zoom surface");
    static_assert(NCH > 1 && NS <= 2 * SRP && TC + NS <= 2 * SRP, "the kept rows lie in the first and last chunks");
    // mac_rows reads up to K - 2 rows past a chunk's search rows; the template pairs follow them
    static_assert(TCP >= K / 2 + 1, "the template chunk follows the search rows");
    static_assert(3 * P <= NTH, "one thread per column and row-pair phase");
    static_assert(6 * P * 8 <= (TC / 2) * (P + 1) * 8, "column-sum partials fit the template chunk region");
    const Fine2Lds lay(P, NL, TC);
    cf *SR = reinterpret_cast<cf *>(smem + lay.sr);
    cf *TB = reinterpret_cast<cf *>(smem + lay.tb);
    float *save = reinterpret_cast<float *>(smem + lay.save);   // rows [0, NS) then [T, T + NS), pitch P
    float *nrm = reinterpret_cast<float *>(smem + lay.nrm);
    float *corr = reinterpret_cast<float *>(smem + lay.corr);
    double *red = reinterpret_cast<double *>(smem + lay.red);
    // twiddles read from LDS (the transforms' table loads wait one at a time)
    cf *twl = reinterpret_cast<cf *>(smem + lay.tw);
    for (int i = threadIdx.x; i < P; i += NTH) twl[i] = tw_g[i];
    const cf *tw = twl;
    const int win = w0 + xcd_window(blockIdx.x, gridDim.x);
    const float *tp = a.tamp + size_t(win) * T * T;
    const float *sg = a.samp + size_t(win) * P * P;
    const double mean = a.tmean[2 * win], t2 = a.tmean[2 * win + 1];
    // planar oversampled chips: pixel (2j + p, 2c + q) sits in plane 2p + q at (j, c)
    auto sval = [&](int r, int x) -> float {
        return sg[size_t(2 * (r & 1) + (x & 1)) * Q * Q + (r >> 1) * Q + (x >> 1)];
    };
    auto tval = [&](int r, int x) -> float {
        return tp[size_t(2 * (r & 1) + (x & 1)) * TQ * TQ + (r >> 1) * TQ + (x >> 1)];
    };
    // FP64 column sums of s and s^2 over rows [0, T): thread (column, phase)
    // takes the row pairs whose index is phase mod 3, both rows of a pair from
    // one load; the three partials are added in phase order by the window sums
    const int cs_c = threadIdx.x % P, cs_ph = threadIdx.x / P;
    double cs1 = 0.0, cs2 = 0.0;

    const int f = threadIdx.x % H, blk = threadIdx.x / H, u0 = blk * K;
    const bool active = (blk < NTH / H) && (u0 < NL);
    cf q[K];
#pragma unroll
    for (int k = 0; k < K; ++k) q[k] = {0.f, 0.f};
    // the MAC lanes all take the complex form; slot 0's (DC, Nyquist) lag sums come
    // from one thread per lag and component in a wave the MAC leaves idle
    constexpr int HB = ((NL + K - 1) / K * H + 63) / 64 * 64;
    static_assert(NTH - HB >= 2 * NL, "an idle wave holds the slot-0 sums");
    const int hid = int(threadIdx.x) - HB;
    float dcn = 0.f;

    #pragma unroll 1
    for (int ch = 0; ch < NCH; ++ch) {
        const int c0 = ch * TC;
        const int first = ch ? KEEP : 0;     // first pair slot loaded in this chunk
        // search rows [c0 + 2 first, c0 + SROWS) into pair slots [first, SRP), eight
        // columns per work item: four 16-byte loads (even and odd columns of both rows)
        for (int i = threadIdx.x; i < (SRP - first) * (L / 8); i += NTH) {
            const int jj = first + i / (L / 8), x0 = 8 * (i - (i / (L / 8)) * (L / 8));
            const int r = c0 + 2 * jj;      // even: planes 0/1 hold row r, planes 2/3 row r + 1
            const float *b = sg + (r >> 1) * Q + (x0 >> 1);
            const float4 e0 = *reinterpret_cast<const float4 *>(b);
            const float4 o0 = *reinterpret_cast<const float4 *>(b + Q * Q);
            const float4 e1 = *reinterpret_cast<const float4 *>(b + 2 * Q * Q);
            const float4 o1 = *reinterpret_cast<const float4 *>(b + 3 * Q * Q);
            cf *d = SR + jj * PP + x0;
            d[0] = {e0.x, e1.x}; d[1] = {o0.x, o1.x}; d[2] = {e0.y, e1.y}; d[3] = {o0.y, o1.y};
            d[4] = {e0.z, e1.z}; d[5] = {o0.z, o1.z}; d[6] = {e0.w, e1.w}; d[7] = {o0.w, o1.w};
        }
        // template rows [c0, c0 + TC) minus the FP64 mean, zero beyond column T
        for (int i = threadIdx.x; i < TCP * (L / 8); i += NTH) {
            const int j = i / (L / 8), x0 = 8 * (i - j * (L / 8));
            cf *d = TB + j * PP + x0;
            if (x0 < T) {
                const float *b = tp + ((c0 >> 1) + j) * TQ + (x0 >> 1);
                const float4 e0 = *reinterpret_cast<const float4 *>(b);
                const float4 o0 = *reinterpret_cast<const float4 *>(b + TQ * TQ);
                const float4 e1 = *reinterpret_cast<const float4 *>(b + 2 * TQ * TQ);
                const float4 o1 = *reinterpret_cast<const float4 *>(b + 3 * TQ * TQ);
                auto m = [&](float v) { return float(double(v) - mean); };
                d[0] = {m(e0.x), m(e1.x)}; d[1] = {m(o0.x), m(o1.x)}; d[2] = {m(e0.y), m(e1.y)};
                d[3] = {m(o0.y), m(o1.y)}; d[4] = {m(e0.z), m(e1.z)}; d[5] = {m(o0.z), m(o1.z)};
                d[6] = {m(e0.w), m(e1.w)}; d[7] = {m(o0.w), m(o1.w)};
            } else {
#pragma unroll
                for (int k = 0; k < 8; ++k) d[k] = cf{0.f, 0.f};
            }
        }
        __syncthreads();
        // column sums of the new row pairs below T (T is even: no pair straddles it)
        if (cs_ph < 3) {
            const int g0 = (c0 >> 1) + first, g1 = min((c0 >> 1) + SRP, T / 2);
            int gp = g0 + (cs_ph - g0 % 3 + 3) % 3;
            const cf *src = SR + (gp - (c0 >> 1)) * PP + cs_c;
            #pragma unroll 2
            for (; gp < g1; gp += 3, src += 3 * PP) {
                const cf pv = *src;
                const double x = pv.x, y = pv.y;
                cs1 += x + y;
                cs2 += fma(x, x, y * y);
            }
        }
        // the first NS rows (chunk 0's first NS/2 pairs) and the last NS (rows
        // [T, P), the last chunk's) kept for the slides
        if (ch == 0 || ch == NCH - 1) {
            const int pr0 = ch == 0 ? 0 : T / 2 - (c0 >> 1), sr0 = ch == 0 ? 0 : NS;
            for (int i = threadIdx.x; i < (NS / 2) * P; i += NTH) {
                const int j = i / P, c = i - j * P;
                const cf pv = SR[(pr0 + j) * PP + c];
                save[(sr0 + 2 * j) * P + c] = pv.x;
                save[(sr0 + 2 * j + 1) * P + c] = pv.y;
            }
        }
        __syncthreads();
        // row spectra of the new search pairs and the template pairs (contiguous in LDS)
        fft_nat2dr<float, -1, R1, R2, NTH>(SR + first * PP, SRP - first + TCP, PP, 1, tw);
        split_pairs_inplace<R1, R2, NTH>(SR + first * PP, SRP - first + TCP, PP);
        if (active) mac_rows<H, K, false>(q, SR, PP, TB, PP, 0, 0, TC, u0, SROWS, f);
        else if (hid >= 0 && hid < 2 * NL) dcn = dc_lag_sum<H>(dcn, SR, PP, TB, PP, TC, 0, hid >> 1, SROWS, hid & 1);
        __syncthreads();
        if (ch + 1 < NCH) {
            // the last KEEP pairs are the next chunk's first
            for (int i = threadIdx.x; i < KEEP * PP; i += NTH) SR[i] = SR[(SRP - KEEP) * PP + i];
            __syncthreads();
        }
    }
    // the template region is free once the last chunk's sums are done
    double *part = reinterpret_cast<double *>(TB);
    if (cs_ph < 3) {
        part[cs_ph * P + cs_c] = cs1;
        part[(3 + cs_ph) * P + cs_c] = cs2;
    }
    cf *Qs = SR;
    if (active && f != 0) {
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (u0 + k < NL) Qs[(u0 + k) * H + f] = q[k];
    }
    if (hid >= 0 && hid < 2 * NL) reinterpret_cast<float *>(Qs + (hid >> 1) * H)[hid & 1] = dcn;
    __syncthreads();
    window_norms_rows<(P + 63) / 64, NTH, T % 64 == 0 ? T / 64 : 0>(
        [&](int r, int c) -> double { return double(r < NS ? save[r * P + c] : save[(NS + r - T) * P + c]); },
        part, 3, P, T, T, NL, NL, t2, nrm);
    inverse_lags<R1, R2, NTH>(Qs, NL, NL, SR + ((NL * H + 1) & ~1), PP, tw, nrm, corr);
    fine_surface<NTH>(a, corr, NL, SR, SR + 16 * 256, reinterpret_cast<int *>(nrm), red, win);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
int coarse_L(int SW) { return SW <= 64 ? 64 : (SW <= 128 ? 128 : 0); }
const int kK[] = {2, 3, 5, 6, 9};

int pick_k(int nl, int blocks) {
    for (int k : kK)
        if (k * blocks >= nl) return k;
    return 0;
}

template <int W, int R1, int R2>
void launch_coarse(int K, const CoarseArgs &a, const cf *tw, int n, int w0, size_t lds, hipStream_t st) {
    switch (K) {
        case 2: hipLaunchKernelGGL((k_coarse<W, R1, R2, 2>), dim3(n), dim3(NT), lds, st, a, tw, w0); break;
        case 3: hipLaunchKernelGGL((k_coarse<W, R1, R2, 3>), dim3(n), dim3(NT), lds, st, a, tw, w0); break;
        case 5: hipLaunchKernelGGL((k_coarse<W, R1, R2, 5>), dim3(n), dim3(NT), lds, st, a, tw, w0); break;
        case 6: hipLaunchKernelGGL((k_coarse<W, R1, R2, 6>), dim3(n), dim3(NT), lds, st, a, tw, w0); break;
        default: hipLaunchKernelGGL((k_coarse<W, R1, R2, 9>), dim3(n), dim3(NT), lds, st, a, tw, w0); break;
    }
}

}  // namespace fused

bool fused_supported(const ampcor_params &p, const ampcor_sizes &s) {
    using namespace fused;
    if (p.window_h != p.window_w || (p.window_h != 32 && p.window_h != 64)) return false;
    if (p.raw_ovs != 2 || s.zoom != 16 || p.surf_ovs != 16 || p.surf_method != 0) return false;
    const int L = coarse_L(s.search_w);
    if (L == 0) return false;
    if (pick_k(s.raw_corr_h, NT / (L / 2)) == 0) return false;
    if (CoarseLds(p.window_w, L, s.search_h, s.raw_corr_h, s.raw_corr_w).total > 160 * 1024) return false;
    // the window sums' block sums (doubles) live in the template chunk's region (cf)
    if (2 * ((s.search_h + NORM_BLK - 1) / NORM_BLK) * s.search_w > coarse_tc(p.window_w) / 2 * (L + 1)) return false;
    const int P = 2 * s.zoom_in_w;
    if (FineLds(P, P - 2 * p.window_w + 1, 2 * p.window_w).total > 160 * 1024) return false;
    return true;
}

// Tables the fused kernels read (twiddles, half-sample shifts). Built on first
// use; Engine::setup builds them for fused configurations, so that the
// allocation-free entry finds them in place.
FusedTables Engine::fused_tables() {
    using namespace fused;
    const int W = p.window_w, Zc = sz.zoom_in_w;
    // half-sample shift tables h[k] = e^{-i pi k_s/N}/N for the chip sizes
    auto hshift = [&](int N) -> const void * {
        const std::string key = "f_hsh" + std::to_string(N);
        auto &b = bufs[key];
        if (!b) {
            forbid_in_nosync("build a shift table");
            std::vector<float> h(2 * N);
            for (int k = 0; k < N; ++k) {
                const int ks = k < N / 2 ? k : k - N;
                const long double ang = -3.14159265358979323846264338327950288L * ks / N;
                h[2 * k] = float(cosl(ang) / N);
                h[2 * k + 1] = float(sinl(ang) / N);
            }
            b.reset(new DevBuf());
            b->reserve(h.size() * sizeof(float));
            HIP_OK(hipMemcpy(b->ptr, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
        }
        return b->ptr;
    };
    auto tw = [&](int n) { return static_cast<const void *>(plan(n).tw32.ptr); };
    FusedTables t;
    t.h_ref = hshift(W);
    t.h_zoom = hshift(Zc);
    t.tw1 = tw(coarse_L(sz.search_w));
    t.twW = tw(W);
    t.twZ = tw(Zc);
    t.tw2 = tw(2 * Zc);
    t.tw256 = tw(256);
    t.tw256_64 = plan(256).tw64.ptr;
    return t;
}

// One launch of each fused kernel over windows [w0, w0 + n) of the arrays in c
// (starts, stage outputs and final outputs indexed by absolute window; the
// oversampled-chip intermediates hold this launch's n windows only).
void Engine::launch_fused(const FusedTables &tb, const FusedLaunch &c, int n, int w0, hipStream_t st) {
    using namespace fused;
    const int W = p.window_w, Zc = sz.zoom_in_w;          // 64 -> 72, 32 -> 40
    const int T2 = 2 * W, P = 2 * Zc;
    const int L1 = coarse_L(sz.search_w);
    const int K1 = pick_k(sz.raw_corr_h, NT / (L1 / 2));
    const int mode = p.image_complex ? (p.deramp == 0 ? 1 : 0) : 2;
    const cf *tw1 = static_cast<const cf *>(tb.tw1), *twW = static_cast<const cf *>(tb.twW);
    const cf *twZ = static_cast<const cf *>(tb.twZ), *tw2 = static_cast<const cf *>(tb.tw2);
    const cf *tw256 = static_cast<const cf *>(tb.tw256);
    const cf *h_ref = static_cast<const cf *>(tb.h_ref), *h_zoom = static_cast<const cf *>(tb.h_zoom);
    float *tamp = c.tamp - size_t(w0) * T2 * T2;          // indexed by absolute window inside the kernels
    double *tmean = c.tmean - 2 * size_t(w0);
    float *samp = c.samp - size_t(w0) * P * P;

    CoarseArgs ca{c.ref, c.ref_pitch, c.sec, c.sec_pitch, mode, c.ref_d, c.ref_a, c.sec_d, c.sec_a,
                  p.hs_down, p.hs_across, sz.half_zoom, sz.corr_raw_zoom_h, sz.corr_raw_zoom_w,
                  c.loc, c.val, c.start, c.shift, c.snr, c.cov, c.mc};
    const size_t lds_c = CoarseLds(W, L1, sz.search_h, sz.raw_corr_h, sz.raw_corr_w).total;
    const size_t lds_or = os_lds_bytes(W), lds_oz = os_lds_bytes(Zc);
    const size_t lds_f = FineLds(P, P - T2 + 1, T2).total;
    const size_t lds_f2 = Fine2Lds(P, P - T2 + 1, 32).total;
    {
        timed("k_coarse", n, st, [&] {
            if (W == 64) launch_coarse<64, 8, 16>(K1, ca, tw1, n, w0, lds_c, st);
            else if (L1 == 128) launch_coarse<32, 8, 16>(K1, ca, tw1, n, w0, lds_c, st);
            else launch_coarse<32, 8, 8>(K1, ca, tw1, n, w0, lds_c, st);
        });
        OsArgs oref{c.ref, c.ref_pitch, mode, c.ref_d, c.ref_a, nullptr, p.deramp, h_ref, tamp, tmean};
        OsArgs osec{c.sec, c.sec_pitch, mode, c.sec_d, c.sec_a, c.start, p.deramp, h_zoom, samp, nullptr};
        FineArgs fa{tamp, tmean, samp, c.start, p.hs_down, p.hs_across, 1.0f / float(p.surf_ovs * p.raw_ovs),
                    tw256, c.zidx, c.offset, c.mf, c.surf ? c.surf - size_t(w0) * 256 : nullptr};
        if (W == 64) {
            timed("k_oversample_ref", n, st, [&] {
                hipLaunchKernelGGL((k_oversample<8, 8, true>), dim3(n), dim3(NT), lds_or, st, oref, twW, w0);
            });
            timed("k_oversample_search", n, st, [&] {
                hipLaunchKernelGGL((k_oversample<8, 9, false>), dim3(n), dim3(NT), lds_oz, st, osec, twZ, w0);
            });
            timed("k_fine", n, st, [&] {
                if (fine_kernel == 2)
                    hipLaunchKernelGGL((k_fine2<128, 9, 16, 3, NT, 32>), dim3(n), dim3(NT), lds_f2, st, fa, tw2, w0);
                else if (fine_threads == 1024)
                    hipLaunchKernelGGL((k_fine<128, 9, 16, 2, 1024>), dim3(n), dim3(1024), lds_f, st, fa, tw2, w0);
                else
                    hipLaunchKernelGGL((k_fine<128, 16, 9, 3, NT>), dim3(n), dim3(NT), lds_f, st, fa, tw2, w0);
            });
        } else {
            timed("k_oversample_ref", n, st, [&] {
                hipLaunchKernelGGL((k_oversample<8, 4, true>), dim3(n), dim3(NT), lds_or, st, oref, twW, w0);
            });
            timed("k_oversample_search", n, st, [&] {
                hipLaunchKernelGGL((k_oversample<8, 5, false>), dim3(n), dim3(NT), lds_oz, st, osec, twZ, w0);
            });
            timed("k_fine", n, st, [&] {
                hipLaunchKernelGGL((k_fine<64, 16, 5, 2, NT>), dim3(n), dim3(NT), lds_f, st, fa, tw2, w0);
            });
        }
        if (c.refined)
            hipLaunchKernelGGL(k_refine<float>, dim3(n), dim3(64), refine_lds(16), st, c.surf, 16, 16, 16,
                               p.surf_ovs, 1.0 / double(p.surf_ovs * p.raw_ovs), c.zidx + 2 * size_t(w0),
                               c.start + 2 * size_t(w0), p.hs_down, p.hs_across,
                               static_cast<const cx<double> *>(tb.tw256_64), c.refined + 2 * size_t(w0));
        HIP_OK(hipGetLastError());
    }
}

void Engine::run_fused(const RunArgs &a) {
    const int W = p.window_w, Zc = sz.zoom_in_w;
    const int T2 = 2 * W, P = 2 * Zc;
    const int chunk = std::min(a.chunk > 0 ? a.chunk : 32768, 65535);
    const int nb = std::min(chunk, a.n);
    const FusedTables tb = fused_tables();
    FusedLaunch c{};
    c.ref = a.ref; c.ref_pitch = a.ref_pitch; c.sec = a.sec; c.sec_pitch = a.sec_pitch;
    c.ref_d = a.ref_d; c.ref_a = a.ref_a; c.sec_d = a.sec_d; c.sec_a = a.sec_a;
    c.loc = a.maxloc ? a.maxloc : buf<int32_t>("f_loc", size_t(a.n) * 2);
    c.val = a.maxval ? a.maxval : buf<float>("f_val", a.n);
    c.start = a.start ? a.start : buf<int32_t>("f_start", size_t(a.n) * 2);
    c.shift = a.shift ? a.shift : buf<int32_t>("f_shift", size_t(a.n) * 2);
    c.zidx = a.zoom_idx ? a.zoom_idx : buf<int32_t>("f_zidx", size_t(a.n) * 2);
    c.tamp = buf<float>("f_tamp", size_t(nb) * T2 * T2);
    c.tmean = buf<double>("f_tmean", 2 * size_t(nb));
    c.samp = buf<float>("f_samp", size_t(nb) * P * P);
    const bool want_margin = precision == 1 || a.margin_coarse || a.margin_fine;
    c.mc = a.margin_coarse ? a.margin_coarse : (want_margin ? buf<float>("f_mc", a.n) : nullptr);
    c.mf = a.margin_fine ? a.margin_fine : (want_margin ? buf<float>("f_mf", a.n) : nullptr);
    c.offset = a.offset; c.snr = a.snr; c.cov = a.cov;
    if (a.refined) {
        c.surf = buf<float>("f_surf", size_t(nb) * 256);
        c.refined = a.refined;
    }
    for (int w0 = 0; w0 < a.n; w0 += nb) launch_fused(tb, c, std::min(nb, a.n - w0), w0, a.stream);
    if (precision == 1 && !a.nosync) rerun_fp64(a, c.mc, c.mf, c.start, c.shift);
}

// ---------------------------------------------------------------------------
// Allocation-free entry (ampcor_run_windows_noalloc). Every intermediate of the
// fused kernels lives in one file-scope device array; each GPU holds its own
// copy until the process ends, and every call on that GPU shares it (calls that
// overlap on one GPU race). Its address on each device is taken once, at context
// creation, with that device current.
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = size_t(256) << 20;   // 256 MiB per GPU
// In host code `scratch` names a host-side stand-in, never the device array:
// host code uses the address hipGetSymbolAddress returns for each device.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mu;
std::map<int, void *> g_scratch;                      // device -> address of that device's copy

size_t al256(size_t x) { return (x + 255) & ~size_t(255); }
}  // namespace

void *noalloc_scratch_init(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch.find(device);
    if (it != g_scratch.end()) return it->second;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        (void)hipGetLastError();
        throw Error(AMPCOR_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    }
    g_scratch[device] = ptr;
    return ptr;
}

// Per-window bytes of the carve-up below: sanitized starts, bad flags, stage
// outputs, margins, template mean, oversampled template and search amplitudes.
static size_t noalloc_window_bytes(int T2, int P) {
    return 16 + 4 + 8 * 4 + 4 + 8 + 16 + size_t(T2) * T2 * 4 + size_t(P) * P * 4;
}

void Engine::run_fused_noalloc(const RunArgs &a) {
    if (!scratch_ptr) throw Error(AMPCOR_E_DEVICE, "the no-alloc scratch of this device is not set up");
    const int W = p.window_w, Zc = sz.zoom_in_w;
    const int T2 = 2 * W, P = 2 * Zc;
    const size_t per = noalloc_window_bytes(T2, P);
    // 11 arrays, each padded to 256 bytes
    const size_t cap = (kScratchBytes - 11 * 256) / per;
    if (cap < 1)
        throw Error(AMPCOR_E_UNSUPPORTED, "one window needs " + std::to_string(per) + " bytes, more than the " +
                                              std::to_string(kScratchBytes) + "-byte no-alloc scratch");
    const int m = int(std::min<size_t>({cap, size_t(a.n), 65535}));
    const FusedTables tb = fused_tables();                 // built by ampcor_setup; never allocates here
    unsigned char *s = static_cast<unsigned char *>(scratch_ptr);
    auto take = [&](size_t bytes) {
        void *q = s;
        s += al256(bytes);
        return q;
    };
    int32_t *st4 = static_cast<int32_t *>(take(16 * size_t(m)));
    int32_t *bad = static_cast<int32_t *>(take(4 * size_t(m)));
    FusedLaunch c{};
    c.ref = a.ref; c.ref_pitch = a.ref_pitch; c.sec = a.sec; c.sec_pitch = a.sec_pitch;
    c.ref_d = st4; c.ref_a = st4 + m; c.sec_d = st4 + 2 * m; c.sec_a = st4 + 3 * m;
    c.loc = static_cast<int32_t *>(take(8 * size_t(m)));
    c.start = static_cast<int32_t *>(take(8 * size_t(m)));
    c.shift = static_cast<int32_t *>(take(8 * size_t(m)));
    c.zidx = static_cast<int32_t *>(take(8 * size_t(m)));
    c.val = static_cast<float *>(take(4 * size_t(m)));
    c.mc = static_cast<float *>(take(4 * size_t(m)));
    c.mf = static_cast<float *>(take(4 * size_t(m)));
    c.tmean = static_cast<double *>(take(16 * size_t(m)));
    c.tamp = static_cast<float *>(take(size_t(m) * T2 * T2 * 4));
    c.samp = static_cast<float *>(take(size_t(m) * P * P * 4));
    if (size_t(s - static_cast<unsigned char *>(scratch_ptr)) > kScratchBytes)
        throw Error(AMPCOR_E_UNSUPPORTED, "no-alloc scratch carve-up exceeds the array");
    for (int w0 = 0; w0 < a.n; w0 += m) {
        const int n = std::min(m, a.n - w0);
        // this chunk's starts, bounds-checked on the device (out-of-image windows get NaN outputs)
        hipLaunchKernelGGL(k_sanitize_starts<int32_t>, dim3(idivup(n, 256)), dim3(256), 0, a.stream, a.ref_d + w0,
                           a.ref_a + w0, a.sec_d + w0, a.sec_a + w0, n, a.ref_h, a.ref_w, a.sec_h, a.sec_w,
                           p.window_h, p.window_w, sz.search_h, sz.search_w, st4, bad);
        if (n < m) {            // the sanitized arrays of a short last chunk start at the same offsets
            c.ref_a = st4 + n; c.sec_d = st4 + 2 * n; c.sec_a = st4 + 3 * n;
        }
        c.offset = a.offset + 2 * size_t(w0);
        c.snr = a.snr + w0;
        c.cov = a.cov + 3 * size_t(w0);
        launch_fused(tb, c, n, 0, a.stream);
        hipLaunchKernelGGL(k_mark_bad<float>, dim3(idivup(n, 256)), dim3(256), 0, a.stream, bad, n, c.offset, c.snr,
                           c.cov);
        HIP_OK(hipGetLastError());
    }
}

}  // namespace ampcor
