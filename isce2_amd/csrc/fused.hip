// Fused per-window pipeline: the fast path for square windows of 32 or 64
// pixels with 2x raw oversampling, a 16x16 zoom surface and 16x FFT surface
// oversampling (the reference's defaults). Three kernels per chunk of windows:
// k_coarse -> k_oversample (reference chip, search zoom chip) -> k_fine.
#include <algorithm>
#include <cmath>
#include <vector>
#include "engine.h"
#include "fused_kernels.h"
#include "stage_kernels.h"

namespace ampcor {
namespace fused {

// byte offsets of LDS regions, shared by host (sizing) and device (carving)
struct CoarseLds {
    int A, work, tb, tst, nrm, corr, red, total;
    __host__ __device__ CoarseLds(int W, int L, int SH, int NL2) {
        auto al = [](int x) { return (x + 15) & ~15; };
        A = 0;
        work = al(A + SH * L * 4);
        tb = al(work + 8 * L * 8);
        tst = al(tb + 4 * L * 8);
        nrm = al(tst + 8 * W * 4);
        corr = al(nrm + NL2 * 4);
        red = al(corr + NL2 * 4);
        total = red + 16 * 8;
    }
};

struct CoarseArgs {
    const void *ref; int64_t ref_pitch;
    const void *sec; int64_t sec_pitch;
    int mode;                      // 0 complex, 1 |z| chips, 2 real image
    const int32_t *ref_d, *ref_a, *sec_d, *sec_a;
    int hs_d, hs_a, hz, patch_h, patch_w;
    int32_t *maxloc; float *maxval; int32_t *start, *shift; float *snr, *cov;
};

// cuAmpcorChunk.cpp:18-106 for one window per workgroup
template <int W, int L, int K>
__global__ void __launch_bounds__(NT) k_coarse(CoarseArgs a, FftDesc d, int w0) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int TPT = (W * W + NT - 1) / NT;
    const int SH = W + 2 * a.hs_d, SW = W + 2 * a.hs_a;
    const int NLh = SH - W + 1, NLw = SW - W + 1;
    const CoarseLds lay(W, L, SH, NLh * NLw);
    float *A = reinterpret_cast<float *>(smem + lay.A);
    cf *work = reinterpret_cast<cf *>(smem + lay.work);
    cf *tb = reinterpret_cast<cf *>(smem + lay.tb);
    float *tst = reinterpret_cast<float *>(smem + lay.tst);
    float *nrm = reinterpret_cast<float *>(smem + lay.nrm);
    float *corr = reinterpret_cast<float *>(smem + lay.corr);
    double *red = reinterpret_cast<double *>(smem + lay.red);
    const int win = w0 + blockIdx.x;

    // template amplitudes in registers; FP64 mean and sum of squares (cuCorrNormalization.cu:210-235)
    const int64_t rd = a.ref_d[win], ra = a.ref_a[win];
    float t[TPT];
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
        const int e = threadIdx.x + NT * k;
        t[k] = 0.f;
        if (e < W * W) {
            const int r = e / W, c = e - r * W;
            t[k] = pixel_amp(a.ref, a.mode, (rd + r) * a.ref_pitch + ra + c);
            s += t[k];
        }
    }
    const double mean = block_sum(s, red) / double(W * W);
    double t2 = 0.0;
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
        const int e = threadIdx.x + NT * k;
        if (e < W * W) {
            const double dv = double(t[k]) - mean;
            t2 += dv * dv;
            t[k] = float(dv);
        }
    }
    t2 = block_sum(t2, red);

    // search amplitudes -> A[SH][SW]
    const int64_t sd = a.sec_d[win], sa = a.sec_a[win];
    for (int i = threadIdx.x; i < SH * SW; i += NT) {
        const int r = i / SW, c = i - r * SW;
        A[i] = pixel_amp(a.sec, a.mode, (sd + r) * a.sec_pitch + sa + c);
    }
    __syncthreads();
    window_norms<(L + 63) / 64>(A, SW, SW, W, W, NLh, NLw, t2, nrm);
    __syncthreads();
    search_spectra_inplace<L>(A, SH, SW, work, d);

    correlate_rows<L, K>(
        [&](int i0, int nr) {
#pragma unroll
            for (int k = 0; k < TPT; ++k) {
                const int e = threadIdx.x + NT * k - i0 * W;
                if (e >= 0 && e < nr * W) tst[e] = t[k];
            }
        },
        [&](int i, int x) -> float { return x < W ? tst[(i & 7) * W + x] : 0.f; },
        W, SH, NLh, NLw, reinterpret_cast<cf *>(A), work, tb, d, corr);

    for (int i = threadIdx.x; i < NLh * NLw; i += NT) corr[i] *= nrm[i];
    __syncthreads();
    int idx;
    float val;
    block_argmax(corr, NLh * NLw, idx, val, reinterpret_cast<float *>(red + 8), reinterpret_cast<uint64_t *>(red + 12));
    if (threadIdx.x == 0) {
        const int px = idx / NLw, py = idx - px * NLw;
        window_stats(corr, NLh, NLw, px, py, double(val), W * W, a.patch_h, a.patch_w, a.cov + 3 * win,
                     a.snr + win);
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
// x2 chip oversampling: deramp, then columns and rows through the half-sample
// shift  odd[m] = (1/N) FFT_-(FFT_+(x) * e^{-i pi k_s/N})[m], even[m] = x[m]
// (the quadrant-padded FFT interpolator of cuOverSampler.cpp:72-78 and
// cuArraysPadding.cu:223-266, separable per axis), then |.| -> HBM.
// ---------------------------------------------------------------------------
constexpr int OG = 8;   // columns / rows per FFT batch

struct OsArgs {
    const void *img; int64_t pitch; int mode;
    const int32_t *sd, *sa, *start;   // start may be null
    int deramp;
    const cf *hsh;                    // [N] e^{-i pi k_s/N} / N
    float *out;                       // [n][2N][2N]
    double *mean;                     // [n] or null
};

template <int N>
__global__ void __launch_bounds__(NT) k_oversample(OsArgs a, FftDesc d, int w0) {
    extern __shared__ __align__(16) unsigned char smem[];
    cf *X = reinterpret_cast<cf *>(smem);
    cf *Y = X + N * N;
    cf *work = Y + N * N;
    double *red = reinterpret_cast<double *>(work + 2 * OG * N);
    const int win = w0 + blockIdx.x;
    int64_t r0 = a.sd[win], c0 = a.sa[win];
    if (a.start) { r0 += a.start[2 * win]; c0 += a.start[2 * win + 1]; }
    for (int i = threadIdx.x; i < N * N; i += NT) {
        const int r = i / N, c = i - r * N;
        X[i] = pixel_cx(a.img, a.mode, (r0 + r) * a.pitch + c0 + c);
    }
    __syncthreads();
    if (a.deramp == 1) {
        // cuDeramp.cu:343-390 with float2's operator*= (imaginary part uses the new real part)
        double ayr = 0, ayi = 0, axr = 0, axi = 0;
        for (int i = threadIdx.x; i < N * N; i += NT) {
            const int r = i / N, c = i - r * N;
            const double zr = X[i].x, zi = X[i].y;
            if (c < N - 1) {
                const double br = X[i + 1].x, bi = X[i + 1].y;
                ayr += zr * br + zi * bi; ayi += zi * br - zr * bi;
            }
            if (r < N - 1) {
                const double br = X[i + N].x, bi = X[i + N].y;
                axr += zr * br + zi * bi; axi += zi * br - zr * bi;
            }
        }
        ayr = block_sum(ayr, red); ayi = block_sum(ayi, red);
        axr = block_sum(axr, red); axi = block_sum(axi, red);
        const double phy = atan2(ayi, ayr), phx = atan2(axi, axr);
        for (int i = threadIdx.x; i < N * N; i += NT) {
            const int r = i / N, c = i - r * N;
            double sn, cs;
            sincos(double(c) * phx + double(r) * phy, &sn, &cs);
            const double x = X[i].x, y = X[i].y;
            const double nx = x * cs - y * sn;
            X[i] = {float(nx), float(y * cs + nx * sn)};
        }
        __syncthreads();
    }
    // column pass: odd rows -> Y
    for (int cb = 0; cb < N; cb += OG) {
        for (int i = threadIdx.x; i < N * OG; i += NT) {
            const int e = i / OG, j = i - e * OG;
            work[i] = X[e * N + cb + j];
        }
        __syncthreads();
        cf *Z = fft_lds<float, +1>(work, work + OG * N, d, OG, OG, 1);
        for (int i = threadIdx.x; i < N * OG; i += NT) Z[i] = cmul(Z[i], a.hsh[i / OG]);
        __syncthreads();
        cf *Z2 = fft_lds<float, -1>(Z, Z == work ? work + OG * N : work, d, OG, OG, 1);
        for (int i = threadIdx.x; i < N * OG; i += NT) {
            const int e = i / OG, j = i - e * OG;
            Y[e * N + cb + j] = Z2[i];
        }
        __syncthreads();
    }
    // row pass over the 2N rows (even: X, odd: Y), amplitudes to HBM
    double sum = 0.0;
    float *out = a.out + size_t(win) * 4 * N * N;
    for (int R0 = 0; R0 < 2 * N; R0 += OG) {
        for (int i = threadIdx.x; i < OG * N; i += NT) {
            const int j = i / N, x = i - j * N;
            const int R = R0 + j;
            work[i] = ((R & 1) ? Y : X)[(R >> 1) * N + x];
        }
        __syncthreads();
        cf *Z = fft_lds<float, +1>(work, work + OG * N, d, OG, 1, N);
        for (int i = threadIdx.x; i < OG * N; i += NT) Z[i] = cmul(Z[i], a.hsh[i % N]);
        __syncthreads();
        cf *Z2 = fft_lds<float, -1>(Z, Z == work ? work + OG * N : work, d, OG, 1, N);
        for (int i = threadIdx.x; i < OG * 2 * N; i += NT) {
            const int j = i / (2 * N), x = i - j * 2 * N;
            const int R = R0 + j;
            const cf v = (x & 1) ? Z2[j * N + (x >> 1)] : ((R & 1) ? Y : X)[(R >> 1) * N + (x >> 1)];
            const float amp = sqrtf(v.x * v.x + v.y * v.y);
            out[size_t(R) * 2 * N + x] = amp;
            sum += amp;
        }
        __syncthreads();
    }
    if (a.mean) {
        sum = block_sum(sum, red);
        if (threadIdx.x == 0) a.mean[win] = sum / double(4 * N * N);
    }
}

// ---------------------------------------------------------------------------
// fine stage: oversampled correlation, normalization, 16x16 trim, 16x surface
// interpolation and argmax, sub-pixel offset (cuAmpcorChunk.cpp:148-207)
// ---------------------------------------------------------------------------
struct FineLds {
    int A, work, tb, nrm, corr, tw, red, total;
    __host__ __device__ FineLds(int P, int L, int NL2) {
        auto al = [](int x) { return (x + 15) & ~15; };
        A = 0;
        int a_bytes = P * L * 4;
        if (a_bytes < 16 * 256 * 8) a_bytes = 16 * 256 * 8;   // surface G[16][256] reuses A
        work = al(A + a_bytes);
        tb = al(work + 8 * L * 8);
        nrm = al(tb + 4 * L * 8);
        corr = al(nrm + NL2 * 4);
        tw = al(corr + NL2 * 4);
        red = al(tw + 256 * 8);
        total = red + 16 * 8;
    }
};

struct FineArgs {
    const float *tamp; const double *tmean;   // [n][T][T], [n]
    const float *samp;                         // [n][P][P]
    const int32_t *start;
    int hs_d, hs_a;
    float ratio;                               // 1/(surf_ovs*raw_ovs)
    const cf *tw256;                           // e^{-2 pi i k/256}
    int32_t *zoom_idx; float *offset;
};

template <int T, int P, int L, int K>
__global__ void __launch_bounds__(NT) k_fine(FineArgs a, FftDesc d, FftDesc d16, int w0) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int NL = P - T + 1;      // 17
    constexpr int Z = NL - 1;          // 16 (trimmed surface)
    constexpr int SO = 256;            // oversampled surface side
    const FineLds lay(P, L, NL * NL);
    float *A = reinterpret_cast<float *>(smem + lay.A);
    cf *work = reinterpret_cast<cf *>(smem + lay.work);
    cf *tb = reinterpret_cast<cf *>(smem + lay.tb);
    float *nrm = reinterpret_cast<float *>(smem + lay.nrm);
    float *corr = reinterpret_cast<float *>(smem + lay.corr);
    cf *tw = reinterpret_cast<cf *>(smem + lay.tw);
    double *red = reinterpret_cast<double *>(smem + lay.red);
    const int win = w0 + blockIdx.x;
    const float *tp = a.tamp + size_t(win) * T * T;
    const float *sp_g = a.samp + size_t(win) * P * P;
    const double mean = a.tmean[win];
    const float meanf = float(mean);

    for (int i = threadIdx.x; i < SO; i += NT) tw[i] = a.tw256[i];
    double t2 = 0.0;
    for (int i = threadIdx.x; i < T * T; i += NT) {
        const double dv = double(tp[i]) - mean;
        t2 += dv * dv;
    }
    t2 = block_sum(t2, red);
    for (int i = threadIdx.x; i < P * P; i += NT) A[i] = sp_g[i];
    __syncthreads();
    window_norms<(P + 63) / 64>(A, P, P, T, T, NL, NL, t2, nrm);
    __syncthreads();
    search_spectra_inplace<L>(A, P, P, work, d);
    correlate_rows<L, K>([](int, int) {},
                         [&](int i, int x) -> float { return x < T ? tp[i * T + x] - meanf : 0.f; },
                         T, P, NL, NL, reinterpret_cast<cf *>(A), work, tb, d, corr);
    // normalize, trim to Z x Z, 2-D +i transform / (Z*Z) (cuOverSampler.cpp:141-145)
    for (int i = threadIdx.x; i < Z * Z; i += NT) {
        const int r = i / Z, c = i - r * Z;
        work[i] = {corr[r * NL + c] * nrm[r * NL + c], 0.f};
    }
    __syncthreads();
    cf *S1 = fft_lds<float, +1>(work, work + Z * Z, d16, Z, 1, Z);
    cf *S2 = fft_lds<float, +1>(S1, S1 == work ? work + Z * Z : work, d16, Z, Z, 1);
    // G[k1][16c+dd] = DFT16_c( S2[k1][k2] e^{-2 pi i dd ks2/256} ) / 256, one (k1, dd) per thread
    cf *G = reinterpret_cast<cf *>(A);
    {
        const int k1 = threadIdx.x >> 4, dd = threadIdx.x & 15;
        cf x[16];
#pragma unroll
        for (int k2 = 0; k2 < 16; ++k2) {
            const int ks = k2 < 8 ? k2 : k2 - 16;
            const cf v = S2[k1 * Z + k2] * (1.0f / 256.0f);
            x[k2] = cmul(v, tw[(dd * ks + 4096) & 255]);
        }
        dft16<float, -1>(x);
#pragma unroll
        for (int c = 0; c < 16; ++c) G[k1 * SO + 16 * c + dd] = x[c];
    }
    __syncthreads();
    // out[16a+b][q] = Re DFT16_a( G[k1][q] e^{-2 pi i b ks1/256} ); running argmax
    float bv = -FLT_MAX;
    uint64_t bk = ~uint64_t(0);
    for (int task = threadIdx.x; task < SO * 16; task += NT) {
        const int q = task & (SO - 1), b = task >> 8;
        cf x[16];
#pragma unroll
        for (int k1 = 0; k1 < 16; ++k1) {
            const int ks = k1 < 8 ? k1 : k1 - 16;
            x[k1] = cmul(G[k1 * SO + q], tw[(b * ks + 4096) & 255]);
        }
        dft16<float, -1>(x);
#pragma unroll
        for (int aa = 0; aa < 16; ++aa) {
            const float v = x[aa].x;
            const uint32_t idx = uint32_t((16 * aa + b) * SO + q);
            const uint64_t key = maxloc_key(idx);
            if (v > -FLT_MAX && better(v, key, bv, bk)) { bv = v; bk = key; }
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const float ov = __shfl_xor(bv, o);
        const uint64_t ok = __shfl_xor(bk, o);
        if (better(ov, ok, bv, bk)) { bv = ov; bk = ok; }
    }
    float *shv = reinterpret_cast<float *>(red + 8);
    uint64_t *shk = reinterpret_cast<uint64_t *>(red + 12);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) { shv[threadIdx.x >> 6] = bv; shk[threadIdx.x >> 6] = bk; }
    __syncthreads();
    if (threadIdx.x == 0) {
        bv = shv[0]; bk = shk[0];
        for (int i = 1; i < NT / 64; ++i)
            if (better(shv[i], shk[i], bv, bk)) { bv = shv[i]; bk = shk[i]; }
        const int idx = (bk == ~uint64_t(0)) ? 0 : int(bk & 0xffffffffu);
        const int px = idx / SO, py = idx % SO;
        a.zoom_idx[2 * win] = px;
        a.zoom_idx[2 * win + 1] = py;
        a.offset[2 * win] = a.ratio * float(px) + float(a.start[2 * win]) - float(a.hs_d);
        a.offset[2 * win + 1] = a.ratio * float(py) + float(a.start[2 * win + 1]) - float(a.hs_a);
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
int coarse_L(int W, int SW) { return SW <= 64 ? 64 : (SW <= 128 ? 128 : 0); }
const int kCoarseK[] = {3, 5, 9, 11, 17};

int pick_k(int nl, int blocks) {
    for (int k : kCoarseK)
        if (k * blocks >= nl) return k;
    return 0;
}

template <int W, int L>
void launch_coarse(int K, const CoarseArgs &a, const FftDesc &d, int n, int w0, size_t lds, hipStream_t st) {
    switch (K) {
        case 3: hipLaunchKernelGGL((k_coarse<W, L, 3>), dim3(n), dim3(NT), lds, st, a, d, w0); break;
        case 5: hipLaunchKernelGGL((k_coarse<W, L, 5>), dim3(n), dim3(NT), lds, st, a, d, w0); break;
        case 9: hipLaunchKernelGGL((k_coarse<W, L, 9>), dim3(n), dim3(NT), lds, st, a, d, w0); break;
        case 11: hipLaunchKernelGGL((k_coarse<W, L, 11>), dim3(n), dim3(NT), lds, st, a, d, w0); break;
        default: hipLaunchKernelGGL((k_coarse<W, L, 17>), dim3(n), dim3(NT), lds, st, a, d, w0); break;
    }
}

}  // namespace fused

bool fused_supported(const ampcor_params &p, const ampcor_sizes &s) {
    using namespace fused;
    if (p.window_h != p.window_w || (p.window_h != 32 && p.window_h != 64)) return false;
    if (p.raw_ovs != 2 || s.zoom != 16 || p.surf_ovs != 16 || p.surf_method != 0) return false;
    const int L = coarse_L(p.window_w, s.search_w);
    if (L == 0 || s.search_w > L) return false;
    if (pick_k(s.raw_corr_h, 256 / (L / 2)) == 0) return false;
    if (CoarseLds(p.window_w, L, s.search_h, s.raw_corr_h * s.raw_corr_w).total > 160 * 1024) return false;
    return true;
}

void Engine::run_fused(const RunArgs &a) {
    using namespace fused;
    const hipStream_t st = a.stream;
    const int W = p.window_w, Zc = sz.zoom_in_w;          // 64 -> 72, 32 -> 40
    const int T2 = 2 * W, P = 2 * Zc;
    const int L1 = coarse_L(W, sz.search_w);
    const int K1 = pick_k(sz.raw_corr_h, NT / (L1 / 2));
    const int L2 = P;                                      // 144 or 80
    const int mode = p.image_complex ? (p.deramp == 0 ? 1 : 0) : 2;
    const int chunk = std::min(a.chunk > 0 ? a.chunk : 2048, 65535);
    const int nb = std::min(chunk, a.n);

    int32_t *loc = a.maxloc ? a.maxloc : buf<int32_t>("f_loc", size_t(a.n) * 2);
    float *val = a.maxval ? a.maxval : buf<float>("f_val", a.n);
    int32_t *start = a.start ? a.start : buf<int32_t>("f_start", size_t(a.n) * 2);
    int32_t *shift = a.shift ? a.shift : buf<int32_t>("f_shift", size_t(a.n) * 2);
    int32_t *zidx = a.zoom_idx ? a.zoom_idx : buf<int32_t>("f_zidx", size_t(a.n) * 2);
    float *tamp = buf<float>("f_tamp", size_t(nb) * T2 * T2);
    double *tmean = buf<double>("f_tmean", nb);
    float *samp = buf<float>("f_samp", size_t(nb) * P * P);

    // half-sample shift tables h[k] = e^{-i pi k_s/N}/N for the chip sizes
    auto hshift = [&](int N) -> const cf * {
        const std::string key = "f_hsh" + std::to_string(N);
        auto &b = bufs[key];
        if (!b) {
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
        return b->as<cf>();
    };
    const cf *h_ref = hshift(W), *h_zoom = hshift(Zc);
    const FftPlan &pl1 = plan(L1), &plW = plan(W), &plZ = plan(Zc), &pl2 = plan(L2), &pl16 = plan(16),
                  &pl256 = plan(256);

    CoarseArgs ca{a.ref, a.ref_pitch, a.sec, a.sec_pitch, mode, a.ref_d, a.ref_a, a.sec_d, a.sec_a,
                  p.hs_down, p.hs_across, sz.half_zoom, sz.corr_raw_zoom_h, sz.corr_raw_zoom_w,
                  loc, val, start, shift, a.snr, a.cov};
    const size_t lds_c = CoarseLds(W, L1, sz.search_h, sz.raw_corr_h * sz.raw_corr_w).total;
    const size_t lds_or = (2 * size_t(W) * W + 2 * OG * W) * sizeof(cf) + 16 * 8;
    const size_t lds_oz = (2 * size_t(Zc) * Zc + 2 * OG * Zc) * sizeof(cf) + 16 * 8;
    const size_t lds_f = FineLds(P, L2, (P - T2 + 1) * (P - T2 + 1)).total;

    for (int w0 = 0; w0 < a.n; w0 += nb) {
        const int n = std::min(nb, a.n - w0);
        timed("k_coarse", n, st, [&] {
            if (W == 64) {
                if (L1 == 128) launch_coarse<64, 128>(K1, ca, pl1.desc32, n, w0, lds_c, st);
                else launch_coarse<64, 64>(K1, ca, pl1.desc32, n, w0, lds_c, st);
            } else {
                if (L1 == 128) launch_coarse<32, 128>(K1, ca, pl1.desc32, n, w0, lds_c, st);
                else launch_coarse<32, 64>(K1, ca, pl1.desc32, n, w0, lds_c, st);
            }
        });
        // chips are addressed with absolute window indices; outputs of the
        // oversampler are per chunk, hence the pointer shifts by w0
        OsArgs oref{a.ref, a.ref_pitch, mode, a.ref_d, a.ref_a, nullptr, p.deramp, h_ref,
                    tamp - size_t(w0) * T2 * T2, tmean - w0};
        OsArgs osec{a.sec, a.sec_pitch, mode, a.sec_d, a.sec_a, start, p.deramp, h_zoom,
                    samp - size_t(w0) * P * P, nullptr};
        FineArgs fa{tamp - size_t(w0) * T2 * T2, tmean - w0, samp - size_t(w0) * P * P, start, p.hs_down,
                    p.hs_across, 1.0f / float(p.surf_ovs * p.raw_ovs), static_cast<const cf *>(pl256.tw32.ptr),
                    zidx, a.offset};
        if (W == 64) {
            timed("k_oversample_ref", n, st, [&] {
                hipLaunchKernelGGL((k_oversample<64>), dim3(n), dim3(NT), lds_or, st, oref, plW.desc32, w0);
            });
            timed("k_oversample_search", n, st, [&] {
                hipLaunchKernelGGL((k_oversample<72>), dim3(n), dim3(NT), lds_oz, st, osec, plZ.desc32, w0);
            });
            timed("k_fine", n, st, [&] {
                hipLaunchKernelGGL((k_fine<128, 144, 144, 6>), dim3(n), dim3(NT), lds_f, st, fa, pl2.desc32,
                                   pl16.desc32, w0);
            });
        } else {
            timed("k_oversample_ref", n, st, [&] {
                hipLaunchKernelGGL((k_oversample<32>), dim3(n), dim3(NT), lds_or, st, oref, plW.desc32, w0);
            });
            timed("k_oversample_search", n, st, [&] {
                hipLaunchKernelGGL((k_oversample<40>), dim3(n), dim3(NT), lds_oz, st, osec, plZ.desc32, w0);
            });
            timed("k_fine", n, st, [&] {
                hipLaunchKernelGGL((k_fine<64, 80, 80, 3>), dim3(n), dim3(NT), lds_f, st, fa, pl2.desc32,
                                   pl16.desc32, w0);
            });
        }
        HIP_OK(hipGetLastError());
    }
}

}  // namespace ampcor
